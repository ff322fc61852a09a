// FP32 GEMM computed on FP16 or BF16 MFMA with FP32 accumulation, for the
// trailing updates of the low-precision LU factorisations.
//
//   C := beta C + alpha A B,   A m x k, B k x n, C m x n, all FP32 in HBM,
//   column-major.  Operands are rounded to FP16 / BF16 while they are
//   staged in LDS; products accumulate in FP32 (v_mfma_f32_16x16x32_f16 /
//   _bf16).  With per-call scaling, B is multiplied by a power of two s
//   chosen so that max |s B| sits near 2^14 (FP16 keeps its range and
//   leading bits; for BF16 it only recentres the exponents), and the result
//   is multiplied by alpha / s: both exact in binary arithmetic.  In the LU
//   the A operand is L21, whose entries are bounded by 1 under partial
//   pivoting, and B is the U12 block row, so one scale per block row is the
//   per-panel scaling of the factorisation.
//
// Split operands (TMVX_LOWPREC_F16X2 / _BF16X2): each FP32 operand x is
// carried as two 16-bit terms, hi = fl16(x) and lo = fl16((x - hi) 2^S)
// (x - hi is exact in FP32; S = 11 for FP16, 8 for BF16, so lo keeps the
// next 11 / 8 bits at hi's magnitude), and the product is
// hi_a hi_b + 2^-S (hi_a lo_b + lo_a hi_b): three MFMA products per tile
// into two FP32 accumulators, the lo x lo term dropped.  The operands then
// carry about 22 (FP16) / 16 (BF16) significant bits instead of 11 / 8
// (the technique of H. Ootomo and R. Yokota, Int. J. High Perform. Comput.
// Appl. 36 (2022) 475-491, on 16-bit matrix units).
//
// Tile: 128 x 128 per workgroup (4 waves, 2 x 2, 64 x 64 each = 4 x 4 MFMA
// tiles), K step 32.  A is read with one coalesced scalar load per row per
// k; B columns are contiguous in k.  Both tiles are stored k-contiguous in
// LDS so each lane reads its 8 k-values with one 16-byte LDS load.  The k
// index inside an MFMA is k = 8 g + v for lane group g and slot v, the
// same for both operands (any common bijection gives the same sum).
#include "tmvx_lu.h"

#include <algorithm>

namespace tmvx {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <int MT> struct LowpT;
template <> struct LowpT<TMVX_LOWPREC_F16> {
    typedef _Float16 type;
    typedef h8 vec;
    static __device__ __forceinline__ f4 mac(vec a, vec b, f4 c)
    { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }
};
template <> struct LowpT<TMVX_LOWPREC_BF16> {
    typedef __bf16 type;
    typedef b8 vec;
    static __device__ __forceinline__ f4 mac(vec a, vec b, f4 c)
    { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
};

// Arithmetic of a lowprec value: the 16-bit MFMA type, the number of terms
// per operand and the exponent shift of the second term.
template <int MT> struct LowpMode {
    static constexpr int BASE = MT, TERMS = 1, SHIFT = 0;
};
template <> struct LowpMode<TMVX_LOWPREC_F16X2> {
    static constexpr int BASE = TMVX_LOWPREC_F16, TERMS = 2, SHIFT = 11;
};
template <> struct LowpMode<TMVX_LOWPREC_BF16X2> {
    static constexpr int BASE = TMVX_LOWPREC_BF16, TERMS = 2, SHIFT = 8;
};

constexpr int LP_BM = 128, LP_BN = 128, LP_BK = 32, LP_PAD = 8;
constexpr float LP_TARGET = 16384.0f;       // 2^14

// Scale s = 2^e with max|B| * s in (TARGET/2, TARGET]; 1 when max is 0/inf/nan.
__device__ __forceinline__ float lowp_scale(const unsigned* bmax)
{
    if (!bmax) return 1.0f;
    const float mx = __uint_as_float(*bmax);
    if (!(mx > 0.0f) || !(mx < 3.0e38f)) return 1.0f;
    int e;
    (void)frexpf(LP_TARGET / mx, &e);         // TARGET/mx = f 2^e, f in [0.5, 1)
    e -= 1;
    e = e < -100 ? -100 : (e > 100 ? 100 : e);
    return ldexpf(1.0f, e);
}

template <int MT>
__global__ void __launch_bounds__(256)
gemm_lowp_kernel(int m, int n, int k, float alpha, const float* __restrict__ A, i64 lda,
                 const float* __restrict__ B, i64 ldb, float beta, float* __restrict__ C, i64 ldc,
                 const unsigned* __restrict__ bmax)
{
    typedef LowpMode<MT> Mode;
    typedef LowpT<Mode::BASE> LT;
    typedef typename LT::type ht;
    typedef typename LT::vec hv;
    constexpr int NT = Mode::TERMS;
    __shared__ __attribute__((aligned(16))) ht As[NT][LP_BM][LP_BK + LP_PAD];
    __shared__ __attribute__((aligned(16))) ht Bs[NT][LP_BN][LP_BK + LP_PAD];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int li = lane & 15, g = lane >> 4;
    const int wm = wave & 1, wn = wave >> 1;
    const int nbm = (m + LP_BM - 1) / LP_BM;
    const int r0 = (blockIdx.x % nbm) * LP_BM, c0 = (blockIdx.x / nbm) * LP_BN;
    const float s = lowp_scale(bmax);
    const float aeff = alpha / s;
    // loader roles: row/column t % 128, k half t / 128 (16 k-values)
    const int lr = t & 127, lk = (t >> 7) * 16;
    const bool b_vec = (ldb % 4) == 0 && ((uintptr_t)B % 16) == 0;

    f4 acc[NT][4][4];          // [0]: hi x hi; [1]: the cross terms, scaled by 2^S
    static_for<0, NT>([&](auto tc) {
        static_for<0, 4>([&](auto ic) {
            static_for<0, 4>([&](auto jc) {
                acc[decltype(tc)::value][decltype(ic)::value][decltype(jc)::value] =
                    f4{0, 0, 0, 0};
            });
        });
    });
    float ra[16], rb[16];
    auto gload = [&](int k0) {
        const int ar = r0 + lr;
        static_for<0, 16>([&](auto vc) {
            constexpr int v = decltype(vc)::value;
            const int kk = k0 + lk + v;
            ra[v] = (ar < m && kk < k) ? A[ar + (i64)kk * lda] : 0.0f;
        });
        const int bc = c0 + lr;
        const float* bp = B + (i64)bc * ldb + k0 + lk;
        if (b_vec && bc < n && k0 + lk + 16 <= k) {
            static_for<0, 4>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                const float4 x = *reinterpret_cast<const float4*>(bp + 4 * q);
                rb[4 * q] = x.x; rb[4 * q + 1] = x.y; rb[4 * q + 2] = x.z; rb[4 * q + 3] = x.w;
            });
        } else {
            static_for<0, 16>([&](auto vc) {
                constexpr int v = decltype(vc)::value;
                rb[v] = (bc < n && k0 + lk + v < k) ? bp[v] : 0.0f;
            });
        }
    };
    auto lstore = [&]() {
        constexpr float up = (float)(1 << Mode::SHIFT);
        hv a0[NT], a1[NT], b0[NT], b1[NT];
        static_for<0, 8>([&](auto vc) {
            constexpr int v = decltype(vc)::value;
            const float x[4] = {ra[v], ra[8 + v], rb[v] * s, rb[8 + v] * s};
            ht h[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) h[e] = (ht)x[e];
            a0[0][v] = h[0];
            a1[0][v] = h[1];
            b0[0][v] = h[2];
            b1[0][v] = h[3];
            if constexpr (NT == 2) {
                a0[1][v] = (ht)((x[0] - (float)h[0]) * up);
                a1[1][v] = (ht)((x[1] - (float)h[1]) * up);
                b0[1][v] = (ht)((x[2] - (float)h[2]) * up);
                b1[1][v] = (ht)((x[3] - (float)h[3]) * up);
            }
        });
        static_for<0, NT>([&](auto tc) {
            constexpr int q = decltype(tc)::value;
            *reinterpret_cast<hv*>(&As[q][lr][lk]) = a0[q];
            *reinterpret_cast<hv*>(&As[q][lr][lk + 8]) = a1[q];
            *reinterpret_cast<hv*>(&Bs[q][lr][lk]) = b0[q];
            *reinterpret_cast<hv*>(&Bs[q][lr][lk + 8]) = b1[q];
        });
    };

    gload(0);
    for (int k0 = 0; k0 < k; k0 += LP_BK) {
        __syncthreads();                 // previous tile consumed
        lstore();
        __syncthreads();
        if (k0 + LP_BK < k) gload(k0 + LP_BK);    // next tile in flight
        hv af[NT][4], bf[NT][4];
        static_for<0, NT>([&](auto tc) {
            constexpr int q = decltype(tc)::value;
            static_for<0, 4>([&](auto ic) {
                constexpr int i = decltype(ic)::value;
                af[q][i] = *reinterpret_cast<const hv*>(&As[q][wm * 64 + 16 * i + li][8 * g]);
                bf[q][i] = *reinterpret_cast<const hv*>(&Bs[q][wn * 64 + 16 * i + li][8 * g]);
            });
        });
        static_for<0, 4>([&](auto ic) {
            constexpr int i = decltype(ic)::value;
            static_for<0, 4>([&](auto jc) {
                constexpr int j = decltype(jc)::value;
                acc[0][i][j] = LT::mac(af[0][i], bf[0][j], acc[0][i][j]);
                if constexpr (NT == 2) {
                    acc[NT - 1][i][j] = LT::mac(af[0][i], bf[NT - 1][j], acc[NT - 1][i][j]);
                    acc[NT - 1][i][j] = LT::mac(af[NT - 1][i], bf[0][j], acc[NT - 1][i][j]);
                }
            });
        });
    }
    constexpr float down = 1.0f / (float)(1 << Mode::SHIFT);
    static_for<0, 4>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        static_for<0, 4>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            const int col = c0 + wn * 64 + 16 * j + li;
            static_for<0, 4>([&](auto rc) {
                constexpr int r = decltype(rc)::value;
                const int row = r0 + wm * 64 + 16 * i + 4 * g + r;
                if (row < m && col < n) {
                    float* p = C + row + (i64)col * ldc;
                    float sum = acc[0][i][j][r];
                    if constexpr (NT == 2) sum += down * acc[NT - 1][i][j][r];
                    const float v = aeff * sum;
                    *p = beta == 0.0f ? v : beta * *p + v;
                }
            });
        });
    });
}

// max |B| over a k x n column-major block into *out (float bits; B >= 0
// compares as unsigned).  *out must start at 0.
__global__ void __launch_bounds__(256)
absmax_kernel(int k, int n, const float* __restrict__ B, i64 ldb, unsigned* __restrict__ out)
{
    float mx = 0.0f;
    const i64 total = (i64)k * n;
    for (i64 e = (i64)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (i64)gridDim.x * blockDim.x) {
        const i64 c = e / k, r = e % k;
        const float v = fabsf(B[r + c * ldb]);
        mx = v > mx ? v : mx;          // NaN is dropped
    }
    static_for<0, 6>([&](auto sc) {
        constexpr int sh = 1 << decltype(sc)::value;
        const float o = __shfl_xor(mx, sh);
        mx = o > mx ? o : mx;
    });
    if ((threadIdx.x & 63) == 0 && mx > 0.0f) atomicMax(out, __float_as_uint(mx));
}

__global__ void word_zero_kernel(unsigned* p) { *p = 0u; }

int gemm_lowp_launch(int lowprec, i64 m, i64 n, i64 k, float alpha, const float* A, i64 lda,
                     const float* B, i64 ldb, float beta, float* C, i64 ldc, unsigned* bmax_word,
                     hipStream_t s)
{
    if (m < 0 || n < 0 || k < 0 || m > INT32_MAX || n > INT32_MAX || k > INT32_MAX)
        return TMVX_ERR_ARG;
    if (m == 0 || n == 0) return TMVX_OK;
    if (lowprec != TMVX_LOWPREC_F16 && lowprec != TMVX_LOWPREC_BF16 &&
        lowprec != TMVX_LOWPREC_F16X2 && lowprec != TMVX_LOWPREC_BF16X2)
        return TMVX_ERR_ARG;
    if (bmax_word) {
        hipLaunchKernelGGL(word_zero_kernel, dim3(1), dim3(1), 0, s, bmax_word);
        const i64 total = k * n;
        const unsigned grid = (unsigned)std::min<i64>((total + 255) / 256, 2048);
        if (total > 0)
            hipLaunchKernelGGL(absmax_kernel, dim3(grid), dim3(256), 0, s, (int)k, (int)n, B, ldb,
                               bmax_word);
        TMVX_CHECK_LAUNCH();
    }
    const i64 tiles = ((m + LP_BM - 1) / LP_BM) * ((n + LP_BN - 1) / LP_BN);
    if (tiles > INT32_MAX) return TMVX_ERR_UNSUPPORTED;
    auto launch = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3((unsigned)tiles), dim3(256), 0, s, (int)m, (int)n, (int)k,
                           alpha, A, lda, B, ldb, beta, C, ldc, (const unsigned*)bmax_word);
    };
    switch (lowprec) {
    case TMVX_LOWPREC_F16: launch(gemm_lowp_kernel<TMVX_LOWPREC_F16>); break;
    case TMVX_LOWPREC_BF16: launch(gemm_lowp_kernel<TMVX_LOWPREC_BF16>); break;
    case TMVX_LOWPREC_F16X2: launch(gemm_lowp_kernel<TMVX_LOWPREC_F16X2>); break;
    default: launch(gemm_lowp_kernel<TMVX_LOWPREC_BF16X2>); break;
    }
    return hip_status(hipGetLastError());
}

} // namespace tmvx

extern "C" int tmvx_gemm_f32_lowp(int lowprec, int64_t m, int64_t n, int64_t k, float alpha,
                                  const float* A, int64_t lda, const float* B, int64_t ldb,
                                  float beta, float* C, int64_t ldc, int scale_b, void* stream)
{
    using namespace tmvx;
    if (m < 0 || n < 0 || k < 0 || lda < (m > 0 ? m : 1) || ldb < (k > 0 ? k : 1) ||
        ldc < (m > 0 ? m : 1))
        return TMVX_ERR_ARG;
    if (lowprec != TMVX_LOWPREC_F16 && lowprec != TMVX_LOWPREC_BF16 &&
        lowprec != TMVX_LOWPREC_F16X2 && lowprec != TMVX_LOWPREC_BF16X2)
        return TMVX_ERR_ARG;
    if (m == 0 || n == 0) return TMVX_OK;
    if (!A || !B || !C) return TMVX_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    unsigned* w = nullptr;
    if (scale_b && hipMallocAsync((void**)&w, 256, s) != hipSuccess) return TMVX_ERR_ALLOC;
    const int rc = gemm_lowp_launch(lowprec, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, w, s);
    if (w) hipFreeAsync(w, s);
    return rc;
}
