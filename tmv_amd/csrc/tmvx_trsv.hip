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
// accumulates acc = sum_c op(A)[b, c] x_c over finished blocks c as their
// ready flags appear, then solves its diagonal block with one lane per row
// (row of the diagonal block in registers, x_k broadcast by readlane),
// publishes x_b with device-coherent stores and raises its flag.  Up to
// SV_RHS right-hand sides per launch, one wave each.  Every wait is bounded.
#include "tmvx_lu.h"

namespace tmvx {

constexpr int SV_NB = 64;
constexpr int SV_RHS = 4;          // RHS per launch (one wave per RHS)
constexpr unsigned SV_SPIN_LIMIT = 1u << 22;

template <typename T>
__device__ __forceinline__ void sv_cstore(T* p, T v)
{ __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
template <typename T>
__device__ __forceinline__ T sv_cload(const T* p)
{ return __hip_atomic_load(const_cast<T*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

struct SvWork {
    unsigned* ticket;     // next block ticket
    unsigned* flags;      // [nblk] = epoch when the block's x is published
    int* err;             // timeout
    int* singular;        // zero diagonal met a non-zero entry
    unsigned epoch;       // value a finished block's flag takes
};

// Solve op(A) X = B for rows of B given through `in_perm` (gather), X
// written to Xout through `out_perm` (scatter), nr <= SV_RHS columns.
// `lower`: op(A) lower triangular (forward order), else upper (backward).
// op(A)(i, j) = trans ? A[j + i*lda] : A[i + j*lda].
//
// Wave w of the block's workgroup owns right-hand side w; lane i owns row
// b0 + i.  A finished block's x arrives as one coherent load per lane and
// is broadcast with readlane, so contributions need no barrier.  Ready
// flags are checked in bulk: one poll reads the flags of the next 64
// blocks in solve order and the longest finished prefix is consumed
// without further polls (contributions stay in solve order, so results
// are deterministic).  The diagonal block and the reciprocals of its
// diagonal are loaded first, off the dependency chain.
template <typename T>
__global__ void __launch_bounds__(256)
trsv_persistent_kernel(int lower, int trans, int unit, i64 n, int nr,
                       const T* __restrict__ A, i64 lda,
                       const T* __restrict__ Bin, i64 ldb, const i64* __restrict__ in_perm,
                       T* __restrict__ Xout, i64 ldx, const i64* __restrict__ out_perm,
                       T* __restrict__ xpub, SvWork wk)
{
    constexpr int NB = SV_NB;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;   // wave = RHS column
    const i64 nblk = (n + NB - 1) / NB;
    __shared__ unsigned s_tk;
    if (t == 0) s_tk = atomicAdd(wk.ticket, 1u);
    __syncthreads();
    const i64 q = s_tk;                       // position in solve order
    if (q >= nblk) return;
    const i64 blk = lower ? q : nblk - 1 - q;
    const i64 b0 = blk * NB;
    const int bs = (int)((n - b0) < NB ? (n - b0) : NB);
    const bool row_ok = lane < bs;
    auto opA = [&](i64 i, i64 j) -> T { return trans ? A[j + i * lda] : A[i + j * lda]; };

    if (wave < nr) {
        // diagonal block row and reciprocal of the diagonal, loaded first
        T drow[NB];
        static_for<0, NB>([&](auto kc) __attribute__((always_inline)) {
            constexpr int k = decltype(kc)::value;
            drow[k] = (row_ok && k < bs) ? opA(b0 + lane, b0 + k) : T(0);
        });
        T rdiag = T(1);
        if (!unit && row_ok) {
            const T d = opA(b0 + lane, b0 + lane);
            rdiag = d == T(0) ? T(0) : T(1) / d;
        }
        T acc = T(0);
        if (row_ok) {
            i64 src = b0 + lane;
            if (in_perm) { src = in_perm[b0 + lane]; if (src < 0 || src >= n) src = b0 + lane; }
            acc = Bin[src + wave * ldb];
        }
        bool failed = false;
        i64 ready = 0;                        // blocks [0, ready) in solve order are done
        for (i64 qc = 0; qc < q && !failed; ++qc) {
            const i64 cb = lower ? qc : nblk - 1 - qc;
            const i64 c0 = cb * NB;
            const int cs = (int)((n - c0) < NB ? (n - c0) : NB);
            T ablk[NB];
            static_for<0, NB>([&](auto kc) __attribute__((always_inline)) {
                constexpr int k = decltype(kc)::value;
                ablk[k] = (row_ok && k < cs) ? opA(b0 + lane, c0 + k) : T(0);
            });
            if (qc >= ready) {
                unsigned spins = 0;
                while (true) {
                    const i64 qq = qc + lane;
                    bool done = true;
                    if (qq < q) {
                        const i64 fb = lower ? qq : nblk - 1 - qq;
                        done = __hip_atomic_load(wk.flags + fb, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT) == wk.epoch;
                    }
                    const unsigned long long bal = __ballot(done);
                    const int run = bal == ~0ull ? 64 : __builtin_ctzll(~bal);
                    if (run > 0) { ready = qc + run; break; }
                    if (++spins > SV_SPIN_LIMIT ||
                        __hip_atomic_load(wk.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                        if (lane == 0) {
                            __hip_atomic_store(wk.err, 1, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
                            atomicOr(wk.singular, 2);      // bit 1: a bounded wait expired
                        }
                        failed = true;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
                if (failed) break;
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            }
            const T xc = lane < cs ? sv_cload(xpub + c0 + lane + (i64)wave * n) : T(0);
            T s0 = T(0), s1 = T(0);
            static_for<0, NB / 2>([&](auto kc) __attribute__((always_inline)) {
                constexpr int k = 2 * decltype(kc)::value;
                s0 += ablk[k] * readlane(xc, k);
                s1 += ablk[k + 1] * readlane(xc, k + 1);
            });
            acc -= s0 + s1;
        }
        // Diagonal block: lane i keeps row i of op(A)[b, b].
        T x = acc;
        bool sing = false;
        if (lower) {
            static_for<0, NB>([&](auto kc) __attribute__((always_inline)) {
                constexpr int k = decltype(kc)::value;
                if (k < bs) {
                    if (!unit && lane == k && x != T(0)) {
                        if (rdiag == T(0)) sing = true; else x = x * rdiag;
                    }
                    const T xk = readlane(x, k);
                    if (lane > k) x -= drow[k] * xk;
                }
            });
        } else {
            static_for<0, NB>([&](auto kc) __attribute__((always_inline)) {
                constexpr int k = NB - 1 - decltype(kc)::value;
                if (k < bs) {
                    if (!unit && lane == k && x != T(0)) {
                        if (rdiag == T(0)) sing = true; else x = x * rdiag;
                    }
                    const T xk = readlane(x, k);
                    if (lane < k) x -= drow[k] * xk;
                }
            });
        }
        if (sing) atomicOr(wk.singular, 1);
        if (row_ok && !failed) {
            sv_cstore(xpub + b0 + lane + (i64)wave * n, x);
            i64 dst = b0 + lane;
            if (out_perm) { dst = out_perm[b0 + lane]; if (dst < 0 || dst >= n) dst = b0 + lane; }
            Xout[dst + wave * ldx] = x;
        }
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (t == 0)
        __hip_atomic_store(wk.flags + blk, wk.epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void sv_init_kernel(unsigned* ticket, unsigned* flags, i64 nblk, int* err)
{
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nblk) flags[i] = 0u;
    if (i == 0) { *ticket = 0u; *err = 0; }
}

// op(A) X = B with X and B possibly aliased only when out_perm == in_perm
// == nullptr and ldx == ldb (each block reads its B rows before writing its
// X rows, and no other block reads them).
template <typename T>
int trsv_multi(int upper, int trans, int unit, i64 n, i64 nrhs, const T* A, i64 lda,
               const T* B, i64 ldb, const i64* in_perm, T* X, i64 ldx, const i64* out_perm,
               int* d_singular, hipStream_t s)
{
    if (n <= 0 || nrhs <= 0) return TMVX_OK;
    const int lower = (upper == 0) != (trans != 0);
    const i64 nblk = (n + SV_NB - 1) / SV_NB;
    char* ws = nullptr;
    const size_t szx = sizeof(T) * (size_t)n * SV_RHS;
    const size_t szf = sizeof(unsigned) * (size_t)nblk;
    if (hipMallocAsync((void**)&ws, szx + szf + 256, s) != hipSuccess) return TMVX_ERR_ALLOC;
    T* xpub = reinterpret_cast<T*>(ws);
    SvWork wk;
    wk.flags = reinterpret_cast<unsigned*>(ws + szx);
    wk.ticket = reinterpret_cast<unsigned*>(ws + szx + szf);
    wk.err = reinterpret_cast<int*>(ws + szx + szf + 64);
    wk.singular = d_singular;
    int rc = TMVX_OK;
    unsigned epoch = 0;
    for (i64 j0 = 0; j0 < nrhs && rc == TMVX_OK; j0 += SV_RHS) {
        const int nr = (int)((nrhs - j0) < SV_RHS ? (nrhs - j0) : SV_RHS);
        ++epoch;
        if (epoch == 1) {
            hipLaunchKernelGGL(sv_init_kernel, dim3((unsigned)((nblk + 255) / 256)), dim3(256),
                               0, s, wk.ticket, wk.flags, nblk, wk.err);
        } else {
            hipLaunchKernelGGL(sv_init_kernel, dim3(1), dim3(1), 0, s, wk.ticket, wk.flags,
                               (i64)0, wk.err);
        }
        if ((rc = hip_status(hipGetLastError())) != TMVX_OK) break;
        wk.epoch = epoch;
        hipLaunchKernelGGL(trsv_persistent_kernel<T>, dim3((unsigned)nblk), dim3(256), 0, s,
                           lower, trans, unit, n, nr, A, lda, B + j0 * ldb, ldb, in_perm,
                           X + j0 * ldx, ldx, out_perm, xpub, wk);
        rc = hip_status(hipGetLastError());
    }
    // The timeout flag is reported by the callers that synchronise.
    hipFreeAsync(ws, s);
    return rc;
}

template int trsv_multi<double>(int, int, int, i64, i64, const double*, i64, const double*, i64,
                                const i64*, double*, i64, const i64*, int*, hipStream_t);
template int trsv_multi<float>(int, int, int, i64, i64, const float*, i64, const float*, i64,
                               const i64*, float*, i64, const i64*, int*, hipStream_t);

} // namespace tmvx
