// Single-launch blocked triangular solve for few right-hand sides.
//
// Counterpart of TMV's column-sweep vector solves DoColTriLDivEq
// (src/TMV_TriDiv_V.cpp:147-218, 321-383) as used by LU_LDivEq for a vector
// or a thin matrix (src/TMV_LUDiv.cpp:185-226): the same substitution, but
// blocked by SV_NB = 64 rows and run as one persistent launch.
//
// Workgroups take row blocks in ticket order (an atomic counter), so a
// workgroup only ever waits for blocks held by workgroups that started
// before it -- the schedule cannot deadlock.  Block b (in solve order)
// accumulates acc = sum_c op(A)[b, c] x_c over finished blocks c, then
// solves its diagonal block with one lane per row (row of the diagonal
// block in registers, x_k broadcast by readlane) and publishes x_b.  Up to
// SV_RHS right-hand sides per launch, one wave each.  Every wait is bounded.
//
// x_b is published as tagged words: each 64-bit word carries the launch's
// epoch in its upper half and 32 bits of x in its lower half (two words per
// double), written and read with device-coherent 64-bit accesses.  A
// consumer polls the words themselves, so a block's result is one one-way
// trip away from the next block, with no ready flag and no cache
// write-back or invalidation on the chain.
#include "tmvx_lu.h"

namespace tmvx {

constexpr int SV_NB = 64;
constexpr int SV_RHS = 4;          // RHS per launch (one wave per RHS)
constexpr unsigned SV_SPIN_LIMIT = 1u << 22;

typedef unsigned long long u64;

template <typename T> struct SvWords { static constexpr int n = (int)(sizeof(T) / 4); };

__device__ __forceinline__ void sv_cstore(u64* p, u64 v)
{ __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ __forceinline__ u64 sv_cload(const u64* p)
{ return __hip_atomic_load(const_cast<u64*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

__device__ __forceinline__ void sv_publish(u64* p, float x, unsigned epoch)
{ sv_cstore(p, ((u64)epoch << 32) | (u64)__float_as_uint(x)); }
__device__ __forceinline__ void sv_publish(u64* p, double x, unsigned epoch)
{
    const u64 b = (u64)__double_as_longlong(x);
    sv_cstore(p, ((u64)epoch << 32) | (b & 0xffffffffull));
    sv_cstore(p + 1, ((u64)epoch << 32) | (b >> 32));
}
// true (and x set) once every word of the value carries `epoch`
__device__ __forceinline__ bool sv_fetch(const u64* p, float& x, unsigned epoch)
{
    const u64 w = sv_cload(p);
    if ((unsigned)(w >> 32) != epoch) return false;
    x = __uint_as_float((unsigned)w);
    return true;
}
__device__ __forceinline__ bool sv_fetch(const u64* p, double& x, unsigned epoch)
{
    const u64 lo = sv_cload(p), hi = sv_cload(p + 1);
    if ((unsigned)(lo >> 32) != epoch || (unsigned)(hi >> 32) != epoch) return false;
    x = __longlong_as_double((long long)((hi << 32) | (lo & 0xffffffffull)));
    return true;
}

// -DTMVX_TRSV_TRACE: wall-clock marks per block (solve order q), 8 slots:
// 0 walk reaches block q-1, 1 x_{q-1} seen, 2 partial sums added,
// 3 diagonal block solved, 4 x_q published, 5 workgroup start
// (trsv_inv_kernel: 0 inverse and M formed, 1 partial sums added, 2 X y
// formed, 3 x_p seen, 4 x_b published, 5 workgroup start).
#ifdef TMVX_TRSV_TRACE
__device__ unsigned long long* g_trsv_trace;
#define TRSV_TRACE(qq, slot)                                                   \
    do {                                                                       \
        if (g_trsv_trace && lane == 0) g_trsv_trace[(qq) * 8 + (slot)] = wall_clock64(); \
    } while (0)
#else
#define TRSV_TRACE(qq, slot) do {} while (0)
#endif

// sum_k row[k] xw[k] for a vector xw in LDS that every lane reads whole
// (broadcast reads, 16 bytes each): per chunk all reads are issued before
// the first multiply, so their latencies overlap; four partial sums, k
// taken in order within each.
template <typename T, typename R, int NB>
__device__ __forceinline__ T bcast_dot(const T* __restrict__ xw, const R (&row)[NB])
{
    constexpr int VW = 16 / (int)sizeof(T);           // values per 16-byte read
    constexpr int CH = sizeof(T) == 8 ? NB / 2 : NB;   // values per chunk
    typedef T VT __attribute__((ext_vector_type(VW)));
    T s0 = T(0), s1 = T(0), s2 = T(0), s3 = T(0);
    static_for<0, NB / CH>([&](auto cc) __attribute__((always_inline)) {
        constexpr int c0 = decltype(cc)::value * CH;
        VT xv[CH / VW];
        static_for<0, CH / VW>([&](auto rc) __attribute__((always_inline)) {
            constexpr int r = decltype(rc)::value;
            xv[r] = reinterpret_cast<const VT*>(xw + c0)[r];
        });
        __builtin_amdgcn_sched_barrier(0);
        static_for<0, CH / 4>([&](auto kc) __attribute__((always_inline)) {
            constexpr int k = 4 * decltype(kc)::value;
            s0 += (T)row[c0 + k] * xv[k / VW][k % VW];
            s1 += (T)row[c0 + k + 1] * xv[(k + 1) / VW][(k + 1) % VW];
            s2 += (T)row[c0 + k + 2] * xv[(k + 2) / VW][(k + 2) % VW];
            s3 += (T)row[c0 + k + 3] * xv[(k + 3) / VW][(k + 3) % VW];
        });
    });
    return (s0 + s1) + (s2 + s3);
}

struct SvWork {
    unsigned* ticket;     // next block ticket
    int* err;             // timeout
    int* singular;        // zero diagonal met a non-zero entry
    unsigned epoch;       // tag of this launch's published words
};

// Solve op(A) X = B for rows of B given through `in_perm` (gather), X
// written to Xout through `out_perm` (scatter), nr <= SV_RHS columns.
// `lower`: op(A) lower triangular (forward order), else upper (backward).
// op(A)(i, j) = trans ? A[j + i*lda] : A[i + j*lda].
//
// The workgroup's 4 waves are shared out as wpr = 4 / nr waves per
// right-hand side; lane i owns row b0 + i.  Sub-wave s of a right-hand
// side walks the finished blocks qc = s, s + wpr, ... (each block of op(A)
// loaded before the wait for the x it multiplies; x arrives as the lane's
// tagged word(s) and is broadcast with readlane).  The partial sums are
// added in sub-wave order through LDS, so results are deterministic, and
// the leading sub-wave solves the diagonal block (loaded first of all, off
// the dependency chain) and publishes x_b.
//
// TA is the storage type of A (the arithmetic is T's): float factors with
// double right-hand sides give FP64 solves with FP32-stored factors.
template <typename T, typename TA>
__global__ void __launch_bounds__(256)
trsv_persistent_kernel(int lower, int trans, int unit, i64 n, int nr,
                       const TA* __restrict__ A, i64 lda,
                       const T* __restrict__ Bin, i64 ldb, const i64* __restrict__ in_perm,
                       T* __restrict__ Xout, i64 ldx, const i64* __restrict__ out_perm,
                       u64* __restrict__ xpub, SvWork wk)
{
    constexpr int NB = SV_NB, W = SvWords<T>::n, NW = 4;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const i64 nblk = (n + NB - 1) / NB;
    __shared__ unsigned s_tk;
    __shared__ int s_fail;
    __shared__ __align__(16) T s_part[NW][NB];   // broadcast row, then partial sums
    if (t == 0) { s_tk = atomicAdd(wk.ticket, 1u); s_fail = 0; }
    __syncthreads();
    const i64 q = s_tk;                       // position in solve order
    if (q >= nblk) return;                    // (whole workgroup)
    if (wave == 0) TRSV_TRACE(q, 5);
    const int wpr = nr >= 1 && nr <= NW ? NW / nr : 1;   // waves per right-hand side
    const int rhs = wave / wpr, sub = wave % wpr;
    const bool active = rhs < nr;
    const i64 blk = lower ? q : nblk - 1 - q;
    const i64 b0 = blk * NB;
    const int bs = (int)((n - b0) < NB ? (n - b0) : NB);
    const bool row_ok = lane < bs;
    // op(A)(b0 + lane, j) = arow[j * astr]
    const TA* arow = trans ? A + (b0 + lane) * lda : A + (b0 + lane);
    const i64 astr = trans ? 1 : lda;
    // op(A)(b0 + lane, c0 + k) for k < cs (0 beyond, and on rows past n)
    auto load_block = [&](T (&blk_)[NB], i64 c0, int cs) __attribute__((always_inline)) {
        const TA* ap = arow + c0 * astr;
        if (row_ok && cs == NB) {
            static_for<0, NB>([&](auto kc) __attribute__((always_inline)) {
                constexpr int k = decltype(kc)::value;
                blk_[k] = (T)ap[k * astr];
            });
        } else {
            static_for<0, NB>([&](auto kc) __attribute__((always_inline)) {
                constexpr int k = decltype(kc)::value;
                blk_[k] = (row_ok && k < cs) ? (T)ap[k * astr] : T(0);
            });
        }
    };
    const bool lead = active && sub == 0;

    // diagonal block row and reciprocal of the diagonal, loaded first
    T drow[NB];
    if (lead) load_block(drow, b0, bs);
    T rdiag = T(1);
    if (lead && !unit && row_ok) {
        const T d = (T)arow[(b0 + lane) * astr];
        rdiag = d == T(0) ? T(0) : T(1) / d;
    }
    T acc = T(0);
    if (lead && row_ok) {
        i64 src = b0 + lane;
        if (in_perm) { src = in_perm[b0 + lane]; if (src < 0 || src >= n) src = b0 + lane; }
        acc = Bin[src + rhs * ldb];
    }
    u64* const xcol = xpub + (i64)(active ? rhs : 0) * n * W;
    bool failed = false;
    // wait until lanes < cs see the tagged words of x at c0 (bounded)
    auto wait_x = [&](i64 c0, int cs, T& xc) __attribute__((always_inline)) -> bool {
        unsigned spins = 0;
        while (true) {
            const bool ok = lane >= cs || sv_fetch(xcol + (c0 + lane) * W, xc, wk.epoch);
            if (__ballot(!ok) == 0ull) return true;
            if (++spins > SV_SPIN_LIMIT ||
                ((spins & 255) == 0 &&
                 __hip_atomic_load(wk.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
                if (lane == 0) {
                    __hip_atomic_store(wk.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    atomicOr(wk.singular, 2);      // bit 1: a bounded wait expired
                    s_fail = 1;
                }
                return false;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    };
    // op(A)[b, c] x_c for the finished block at solve position qc, through
    // the wave's LDS row (64 readlanes would each wait for their SGPR)
    auto block_product = [&](i64 qc, bool trace) __attribute__((always_inline)) -> bool {
        const i64 cb = lower ? qc : nblk - 1 - qc;
        const i64 c0 = cb * NB;
        const int cs = (int)((n - c0) < NB ? (n - c0) : NB);
        T ablk[NB];
        load_block(ablk, c0, cs);
        T xc = T(0);
        if (trace) TRSV_TRACE(q, 0);
        if (!wait_x(c0, cs, xc)) return false;
        if (trace) TRSV_TRACE(q, 1);
        s_part[wave][lane] = xc;
        __builtin_amdgcn_wave_barrier();
        acc -= bcast_dot<T, T, NB>(&s_part[wave][0], ablk);
        __builtin_amdgcn_wave_barrier();
        return true;
    };
    // the blocks solved before q - 1, shared out over the sub-waves; the
    // leading wave takes block q - 1 itself after the partial sums, so the
    // barrier and the exchange are off the chain
    for (i64 qc = sub; active && qc + 1 < q && !failed; qc += wpr)
        failed = !block_product(qc, false);
    if (wpr > 1) {                           // partial sums, in sub-wave order
        if (active && sub > 0) s_part[wave][lane] = acc;
        __syncthreads();
        if (lead)
            for (int u = 1; u < wpr; ++u) acc += s_part[wave + u][lane];
    } else {
        __syncthreads();
    }
    if (!lead || s_fail) return;
    if (q >= 1 && !block_product(q - 1, true)) return;
    TRSV_TRACE(q, 2);
    // Diagonal block: lane i keeps row i of op(A)[b, b].  Step k: x_k is
    // scaled by the reciprocal of its diagonal (a zero x_k stays as it is;
    // a zero diagonal under a non-zero x_k is singular) and broadcast, and
    // the rows still to come subtract op(A)(i, k) x_k.  The 64 steps are
    // the dependency chain, so it carries no branch and no select: every
    // lane takes the subtraction (rows already solved hold their result in
    // xs and their x is not read again).
    T x = acc, xs = acc;
    bool sing = false;
    auto dstep = [&](auto kc) __attribute__((always_inline)) {
        constexpr int k = decltype(kc)::value;
        const T xr = readlane(x, k);
        const T rk = readlane(rdiag, k);      // 1 for a unit diagonal
        sing |= (xr != T(0)) & (rk == T(0));
        const T xk = xr != T(0) ? xr * rk : xr;
        x -= drow[k] * xk;
        if (lane == k) xs = xk;
    };
    // When every reciprocal is finite and non-zero (no zero or subnormal
    // diagonal entry in the block) the select and the singular test leave
    // the chain: x_k = x_r (1/d) is then 0 for x_r = 0 as well.
    auto dstep_fast = [&](auto kc) __attribute__((always_inline)) {
        constexpr int k = decltype(kc)::value;
        const T xk = readlane(x, k) * readlane(rdiag, k);
        x -= drow[k] * xk;
        if (lane == k) xs = xk;
    };
    auto sweep = [&](auto&& step) __attribute__((always_inline)) {
        if (lower) {
            if (bs == NB)
                static_for<0, NB>([&](auto kc) __attribute__((always_inline)) { step(kc); });
            else
                static_for<0, NB>([&](auto kc) __attribute__((always_inline)) {
                    if (decltype(kc)::value < bs) step(kc);
                });
        } else {
            // upper: k = NB-1 .. 0 (rows past bs are skipped)
            if (bs == NB)
                static_for<0, NB>([&](auto kc) __attribute__((always_inline)) {
                    step(std::integral_constant<int, NB - 1 - decltype(kc)::value>{});
                });
            else
                static_for<0, NB>([&](auto kc) __attribute__((always_inline)) {
                    constexpr int k = NB - 1 - decltype(kc)::value;
                    if (k < bs) step(std::integral_constant<int, k>{});
                });
        }
    };
    const bool fast =
        __ballot(row_ok && !(rdiag != T(0) && __builtin_isfinite(rdiag))) == 0ull;
    if (fast) sweep(dstep_fast);
    else sweep(dstep);
    if (sing) atomicOr(wk.singular, 1);
    TRSV_TRACE(q, 3);
    x = xs;
    if (row_ok) {
        sv_publish(xcol + (b0 + lane) * W, x, wk.epoch);
        TRSV_TRACE(q, 4);
        i64 dst = b0 + lane;
        if (out_perm) { dst = out_perm[b0 + lane]; if (dst < 0 || dst >= n) dst = b0 + lane; }
        Xout[dst + rhs * ldx] = x;
    }
}

// The same solve with the diagonal substitution taken off the chain, for
// the mixed-precision solver's correction solves (refinement absorbs the
// different rounding; the public solves keep the substitution above).
// With D = op(A)[b, b], A1 = op(A)[b, p] and A2 = op(A)[b, p2] for the
// blocks p and p2 solved one and two before b, X = D^-1, M = X A1 and
// N = X A2 (formed by the workgroup while earlier blocks are still being
// solved: X by substitution on the columns of I, 16 columns per wave; M and
// N by 64 x 64 x 64 products):
//   y   = B_b - sum over the blocks c solved before p2 of op(A)[b, c] x_c
//   x_b = X y - N x_p2 - M x_p,
// X y is formed as soon as y is (two block steps before x_p arrives, so
// the walk's last wait is off the chain) and N x_p2 one step before, so
// the chain carries only x_p's arrival, one 64-term product per lane and
// the publication of x_b.  Rounding differs from substitution by the
// explicit inverse (its error grows with the condition of the 64 x 64
// diagonal blocks, as for BLAS solves that invert diagonal blocks).  A zero
// diagonal sets bit 0 of `singular` and acts as reciprocal 0.
template <typename T, typename TA>
__global__ void __launch_bounds__(256)
trsv_inv_kernel(int lower, int trans, int unit, i64 n, int nr,
                const TA* __restrict__ A, i64 lda,
                const T* __restrict__ Bin, i64 ldb, const i64* __restrict__ in_perm,
                T* __restrict__ Xout, i64 ldx, const i64* __restrict__ out_perm,
                u64* __restrict__ xpub, SvWork wk)
{
    constexpr int NB = SV_NB, W = SvWords<T>::n, NW = 4, LP = NB + 1, CW = NB / NW;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const i64 nblk = (n + NB - 1) / NB;
    __shared__ unsigned s_tk;
    __shared__ int s_fail;
    // s_part[w]: wave w's broadcast row (a vector one entry per lane, read
    // back whole by every lane: an LDS broadcast instead of 64 readlanes,
    // each of which waits for its SGPR), then its partial sums
    __shared__ __align__(16) T s_part[NW][NB];
    __shared__ T s_d[NB * LP];               // D, then X = D^-1 (row i at i * LP)
    __shared__ T s_a[NB * LP];               // A1, then M = X A1
    __shared__ T s_n[NB * LP];               // A2, then N = X A2
    if (t == 0) { s_tk = atomicAdd(wk.ticket, 1u); s_fail = 0; }
    __syncthreads();
    const i64 q = s_tk;                       // position in solve order
    if (q >= nblk) return;                    // (whole workgroup)
    if (wave == 0) TRSV_TRACE(q, 5);
    const int wpr = nr >= 1 && nr <= NW ? NW / nr : 1;
    const int rhs = wave / wpr, sub = wave % wpr;
    const bool active = rhs < nr;
    const i64 blk = lower ? q : nblk - 1 - q;
    const i64 b0 = blk * NB;
    const int bs = (int)((n - b0) < NB ? (n - b0) : NB);
    const bool has_prev = q > 0;
    const i64 p0 = (lower ? blk - 1 : blk + 1) * NB;
    const int ps = has_prev ? (int)((n - p0) < NB ? (n - p0) : NB) : 0;
    const bool has_prev2 = q > 1;                 // block solved two before b
    const i64 p20 = (lower ? blk - 2 : blk + 2) * NB;
    const int ps2 = has_prev2 ? (int)((n - p20) < NB ? (n - p20) : NB) : 0;
    // op(A)(r, c), in A's storage type
    auto opa = [&](i64 r, i64 c) __attribute__((always_inline)) -> TA {
        return trans ? A[c + r * lda] : A[r + c * lda];
    };

    // 1. D (the other triangle zeroed, unit diagonal as 1, identity past
    // the block's rows) and A1 (zero past bs rows and ps columns) into LDS;
    // the index order follows the storage so that the loads coalesce.  All
    // loads are issued before any value is converted and stored, so they
    // overlap (a conversion next to each load would wait for it).
    constexpr int PER = NB * NB / 256;
    {
        TA dl[PER], al[PER], a2[PER];
        static_for<0, PER>([&](auto rc) __attribute__((always_inline)) {
            constexpr int r = decltype(rc)::value;
            const int e = t + 256 * r;
            const int i = trans ? (e >> 6) : (e & 63);
            const int k = trans ? (e & 63) : (e >> 6);
            dl[r] = (i < bs && k < bs && (lower ? k <= i : k >= i)) ? opa(b0 + i, b0 + k) : TA(0);
            al[r] = (has_prev && i < bs && k < ps) ? opa(b0 + i, p0 + k) : TA(0);
            a2[r] = (has_prev2 && i < bs && k < ps2) ? opa(b0 + i, p20 + k) : TA(0);
        });
        static_for<0, PER>([&](auto rc) __attribute__((always_inline)) {
            constexpr int r = decltype(rc)::value;
            const int e = t + 256 * r;
            const int i = trans ? (e >> 6) : (e & 63);
            const int k = trans ? (e & 63) : (e >> 6);
            T dv = (T)dl[r];
            if (i == k && (unit || i >= bs)) dv = T(1);
            s_d[i * LP + k] = dv;
            if (has_prev) s_a[i * LP + k] = (T)al[r];
            if (has_prev2) s_n[i * LP + k] = (T)a2[r];
        });
    }
    __syncthreads();
    // 2. X = D^-1, wave w solving D X = I for columns CW w .. CW w + CW-1
    // (lane i holds row i; step k scales row k by its reciprocal, broadcasts
    // it and the later rows subtract D(i, k) times it).
    {
        const T dii = s_d[lane * LP + lane];
        const T rdi = dii == T(0) ? T(0) : T(1) / dii;
        if (wave == 0 && lane < bs && dii == T(0)) atomicOr(wk.singular, 1);
        T wv[CW];
        static_for<0, CW>([&](auto jc) __attribute__((always_inline)) {
            constexpr int jj = decltype(jc)::value;
            wv[jj] = lane == wave * CW + jj ? T(1) : T(0);
        });
        for (int kk = 0; kk < NB; ++kk) {
            const int k = lower ? kk : NB - 1 - kk;
            const T dik = s_d[lane * LP + k];
            const T rk = readlane(rdi, k);
            static_for<0, CW>([&](auto jc) __attribute__((always_inline)) {
                constexpr int jj = decltype(jc)::value;
                const T xk = readlane(wv[jj], k) * rk;
                wv[jj] = lane == k ? xk : wv[jj] - dik * xk;
            });
        }
        __syncthreads();                      // every wave done with D
        static_for<0, CW>([&](auto jc) __attribute__((always_inline)) {
            constexpr int jj = decltype(jc)::value;
            s_d[lane * LP + wave * CW + jj] = wv[jj];
        });
        __syncthreads();
    }
    // 3. M = X A1 and N = X A2, wave w forming columns CW w .. CW w + CW-1
    // (one product at a time: each overwrites its own operand in LDS)
    auto times_x = [&](T* sm) __attribute__((always_inline)) {
        T m[CW];
        static_for<0, CW>([&](auto jc) __attribute__((always_inline)) {
            m[decltype(jc)::value] = T(0);
        });
        for (int k = 0; k < NB; ++k) {
            const T xik = s_d[lane * LP + k];
            const T* ak = sm + k * LP + wave * CW;
            static_for<0, CW>([&](auto jc) __attribute__((always_inline)) {
                constexpr int jj = decltype(jc)::value;
                m[jj] += xik * ak[jj];
            });
        }
        __syncthreads();                      // every wave done with the operand
        static_for<0, CW>([&](auto jc) __attribute__((always_inline)) {
            constexpr int jj = decltype(jc)::value;
            sm[lane * LP + wave * CW + jj] = m[jj];
        });
    };
    if (has_prev) times_x(s_a);
    if (has_prev2) times_x(s_n);
    __syncthreads();
    if (wave == 0) TRSV_TRACE(q, 0);

    // 4. The walk over the blocks solved before p, shared out as in the
    // substitution kernel; partial sums added in sub-wave order.
    const bool row_ok = lane < bs;
    const TA* arow = trans ? A + (b0 + lane) * lda : A + (b0 + lane);
    const i64 astr = trans ? 1 : lda;
    const bool lead = active && sub == 0;
    T* const xw = &s_part[wave][0];
    T acc = T(0);
    if (lead && row_ok) {
        i64 src = b0 + lane;
        if (in_perm) { src = in_perm[b0 + lane]; if (src < 0 || src >= n) src = b0 + lane; }
        acc = Bin[src + rhs * ldb];
    }
    u64* const xcol = xpub + (i64)(active ? rhs : 0) * n * W;
    bool failed = false;
    // wait until lanes < cs see the tagged words of x at c0 (bounded)
    auto wait_x = [&](i64 c0, int cs, T& xc) __attribute__((always_inline)) -> bool {
        unsigned spins = 0;
        while (true) {
            const bool ok = lane >= cs || sv_fetch(xcol + (c0 + lane) * W, xc, wk.epoch);
            if (__ballot(!ok) == 0ull) return true;
            if (++spins > SV_SPIN_LIMIT ||
                ((spins & 255) == 0 &&
                 __hip_atomic_load(wk.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
                if (lane == 0) {
                    __hip_atomic_store(wk.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    atomicOr(wk.singular, 2);      // bit 1: a bounded wait expired
                    s_fail = 1;
                }
                return false;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    };
    for (i64 qc = sub; active && qc + 2 < q && !failed; qc += wpr) {
        const i64 cb = lower ? qc : nblk - 1 - qc;
        const i64 c0 = cb * NB;
        const int cs = (int)((n - c0) < NB ? (n - c0) : NB);
        TA ablk[NB];                          // storage type: converted at use
        const TA* ap = arow + c0 * astr;
        static_for<0, NB>([&](auto kc) __attribute__((always_inline)) {
            constexpr int k = decltype(kc)::value;
            ablk[k] = (row_ok && k < cs) ? ap[k * astr] : TA(0);
        });
        T xc = T(0);
        if (!wait_x(c0, cs, xc)) { failed = true; break; }
        xw[lane] = xc;
        __builtin_amdgcn_wave_barrier();
        acc -= bcast_dot<T, TA, NB>(xw, ablk);
        __builtin_amdgcn_wave_barrier();
    }
    if (wpr > 1) {
        if (active && sub > 0) s_part[wave][lane] = acc;
        __syncthreads();
        if (lead)
            for (int u = 1; u < wpr; ++u) acc += s_part[wave + u][lane];
    } else {
        __syncthreads();
    }
    if (!lead || s_fail) return;
    TRSV_TRACE(q, 1);
    // 5. X y (before x_p arrives; X's row in registers), then
    // x_b = X y - M x_p
    T x;
    {
        T xrow[NB];
        static_for<0, NB>([&](auto kc) __attribute__((always_inline)) {
            constexpr int k = decltype(kc)::value;
            xrow[k] = s_d[lane * LP + k];
        });
        xw[lane] = acc;
        __builtin_amdgcn_wave_barrier();
        x = bcast_dot<T, T, NB>(xw, xrow);
        __builtin_amdgcn_wave_barrier();
    }
    TRSV_TRACE(q, 2);
    if (has_prev2) {
        T nrow[NB];
        static_for<0, NB>([&](auto kc) __attribute__((always_inline)) {
            constexpr int k = decltype(kc)::value;
            nrow[k] = s_n[lane * LP + k];
        });
        T xp2 = T(0);
        if (!wait_x(p20, ps2, xp2)) return;
        xw[lane] = xp2;
        __builtin_amdgcn_wave_barrier();
        x -= bcast_dot<T, T, NB>(xw, nrow);
        __builtin_amdgcn_wave_barrier();
    }
    if (has_prev) {
        T mrow[NB];
        static_for<0, NB>([&](auto kc) __attribute__((always_inline)) {
            constexpr int k = decltype(kc)::value;
            mrow[k] = s_a[lane * LP + k];
        });
        T xp = T(0);
        if (!wait_x(p0, ps, xp)) return;
        TRSV_TRACE(q, 3);
        xw[lane] = xp;
        __builtin_amdgcn_wave_barrier();
        x -= bcast_dot<T, T, NB>(xw, mrow);
    }
    if (row_ok) {
        sv_publish(xcol + (b0 + lane) * W, x, wk.epoch);
        TRSV_TRACE(q, 4);
        i64 dst = b0 + lane;
        if (out_perm) { dst = out_perm[b0 + lane]; if (dst < 0 || dst >= n) dst = b0 + lane; }
        Xout[dst + rhs * ldx] = x;
    }
}

// Per launch: ticket and error flag reset; before the first launch of a
// solve also the published words (a fresh workspace may hold any tags).
__global__ void sv_init_kernel(unsigned* ticket, u64* words, i64 nwords, int* err)
{
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nwords) words[i] = 0ull;
    if (i == 0) { *ticket = 0u; *err = 0; }
}

// op(A) X = B with X and B possibly aliased only when out_perm == in_perm
// == nullptr and ldx == ldb (each block reads its B rows before writing its
// X rows, and no other block reads them).
template <typename T, typename TA>
int trsv_multi(int upper, int trans, int unit, i64 n, i64 nrhs, const TA* A, i64 lda,
               const T* B, i64 ldb, const i64* in_perm, T* X, i64 ldx, const i64* out_perm,
               int* d_singular, hipStream_t s, int inv)
{
    if (n <= 0 || nrhs <= 0) return TMVX_OK;
    const int lower = (upper == 0) != (trans != 0);
    const i64 nblk = (n + SV_NB - 1) / SV_NB;
    if (nblk > (i64)0xffffffffu) return TMVX_ERR_UNSUPPORTED;
    char* ws = nullptr;
    const i64 nwords = n * SV_RHS * SvWords<T>::n;
    const size_t szx = sizeof(u64) * (size_t)nwords;
    if (hipMallocAsync((void**)&ws, szx + 256, s) != hipSuccess) return TMVX_ERR_ALLOC;
    u64* xpub = reinterpret_cast<u64*>(ws);
    SvWork wk;
    wk.ticket = reinterpret_cast<unsigned*>(ws + szx);
    wk.err = reinterpret_cast<int*>(ws + szx + 64);
    wk.singular = d_singular;
    int rc = TMVX_OK;
    unsigned epoch = 0;
    for (i64 j0 = 0; j0 < nrhs && rc == TMVX_OK; j0 += SV_RHS) {
        const int nr = (int)((nrhs - j0) < SV_RHS ? (nrhs - j0) : SV_RHS);
        ++epoch;
        const i64 nz = epoch == 1 ? nwords : 0;
        hipLaunchKernelGGL(sv_init_kernel, dim3((unsigned)((nz + 255) / 256 > 0 ? (nz + 255) / 256 : 1)),
                           dim3(256), 0, s, wk.ticket, xpub, nz, wk.err);
        if ((rc = hip_status(hipGetLastError())) != TMVX_OK) break;
        wk.epoch = epoch;
        if (inv)
            hipLaunchKernelGGL((trsv_inv_kernel<T, TA>), dim3((unsigned)nblk), dim3(256), 0, s,
                               lower, trans, unit, n, nr, A, lda, B + j0 * ldb, ldb, in_perm,
                               X + j0 * ldx, ldx, out_perm, xpub, wk);
        else
            hipLaunchKernelGGL((trsv_persistent_kernel<T, TA>), dim3((unsigned)nblk), dim3(256),
                               0, s, lower, trans, unit, n, nr, A, lda, B + j0 * ldb, ldb,
                               in_perm, X + j0 * ldx, ldx, out_perm, xpub, wk);
        rc = hip_status(hipGetLastError());
    }
    // The timeout flag is reported by the callers that synchronise.
    hipFreeAsync(ws, s);
    return rc;
}

template int trsv_multi<double, double>(int, int, int, i64, i64, const double*, i64,
                                        const double*, i64, const i64*, double*, i64, const i64*,
                                        int*, hipStream_t, int);
template int trsv_multi<float, float>(int, int, int, i64, i64, const float*, i64, const float*,
                                      i64, const i64*, float*, i64, const i64*, int*, hipStream_t, int);
template int trsv_multi<double, float>(int, int, int, i64, i64, const float*, i64,
                                       const double*, i64, const i64*, double*, i64, const i64*,
                                       int*, hipStream_t, int);

} // namespace tmvx

#ifdef TMVX_TRSV_TRACE
extern "C" int tmvx_trsv_trace_set(unsigned long long* buf)
{
    return hipMemcpyToSymbol(HIP_SYMBOL(tmvx::g_trsv_trace), &buf, sizeof(buf)) == hipSuccess
               ? 0 : -3;
}
#endif
