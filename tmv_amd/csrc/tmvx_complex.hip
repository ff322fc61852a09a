// Complex GEMM: C := alpha op(A) op(B) + beta C for interleaved complex
// storage (std::complex<double> / std::complex<float>), op = identity or
// conjugate.
//
// TMV computes a complex product as four real products of the real and
// imaginary parts (DoBlockMultMM's complex overloads, src/TMV_MultMM_Block.cpp:
// 334-392), with the conjugation flags folded into the signs.  Here the same
// four real products run on the real MFMA GEMM: the operands are first split
// into planar real and imaginary parts (one pass each; the real GEMM needs a
// unit stride), then
//     Pr = Ar Br + sR Ai Bi,   Pi = sB Ar Bi + sA Ai Br,
// sA = -1 for conj(A), sB = -1 for conj(B), sR = -sA sB; finally one pass
// forms C = beta C + alpha (Pr + i Pi) in the interleaved layout.  The
// arithmetic is 8mnk real flops, as TMV's (no 3-multiplication variant, whose
// rounding differs).
#include "tmvx_common.h"

namespace tmvx {

template <typename R> struct Cplx;
template <> struct Cplx<double> { typedef double2 type; };
template <> struct Cplx<float> { typedef float2 type; };

// X: `outer` lines of `inner` contiguous complex elements, line stride ldx
// (in complex elements) -> planar Xr, Xi with line stride `inner`.
template <typename R>
__global__ void split_kernel(i64 outer, i64 inner, const typename Cplx<R>::type* __restrict__ X,
                             i64 ldx, R* __restrict__ Xr, R* __restrict__ Xi)
{
    for (i64 o = blockIdx.y; o < outer; o += gridDim.y)
        for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < inner;
             i += (i64)gridDim.x * blockDim.x) {
            const typename Cplx<R>::type z = X[i + o * ldx];
            Xr[i + o * inner] = z.x;
            Xi[i + o * inner] = z.y;
        }
}

// C(i, j) = beta C(i, j) + alpha (Pr + i Pi)(i, j); C column-major, P with
// leading dimension m (P = 0 when !have_p: k == 0).  beta == 0: C is not read.
template <typename R>
__global__ void combine_kernel(i64 m, i64 n, const R* __restrict__ Pr, const R* __restrict__ Pi,
                               bool have_p, R ar, R ai, R br, R bi, bool beta_zero,
                               typename Cplx<R>::type* __restrict__ C, i64 ldc)
{
    for (i64 j = blockIdx.y; j < n; j += gridDim.y)
        for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < m;
             i += (i64)gridDim.x * blockDim.x) {
            const R pr = have_p ? Pr[i + j * m] : R(0), pi = have_p ? Pi[i + j * m] : R(0);
            R cr = ar * pr - ai * pi, ci = ar * pi + ai * pr;
            if (!beta_zero) {
                const typename Cplx<R>::type c = C[i + j * ldc];
                cr += br * c.x - bi * c.y;
                ci += br * c.y + bi * c.x;
            }
            typename Cplx<R>::type z;
            z.x = cr;
            z.y = ci;
            C[i + j * ldc] = z;
        }
}

static dim3 grid2(i64 inner, i64 outer)
{
    const i64 gx = (inner + 255) / 256;
    return dim3((unsigned)(gx < 1024 ? (gx > 0 ? gx : 1) : 1024),
                (unsigned)(outer < 65535 ? (outer > 0 ? outer : 1) : 65535));
}

template <typename R>
int gemm_complex(int la, int lb, int conj_a, int conj_b, i64 m, i64 n, i64 k, R ar, R ai,
                 const R* A, i64 lda, const R* B, i64 ldb, R br, R bi, R* C, i64 ldc,
                 hipStream_t s)
{
    typedef typename Cplx<R>::type Z;
    if ((la != TMVX_COLMAJOR && la != TMVX_ROWMAJOR) ||
        (lb != TMVX_COLMAJOR && lb != TMVX_ROWMAJOR))
        return TMVX_ERR_ARG;
    if (m < 0 || n < 0 || k < 0) return TMVX_ERR_ARG;
    if (m == 0 || n == 0) return TMVX_OK;
    if (ldc < m || lda < (la ? k : m) || ldb < (lb ? n : k) || lda < 1 || ldb < 1)
        return TMVX_ERR_ARG;
    if (!C || (k > 0 && (!A || !B))) return TMVX_ERR_ARG;
    // whole complex elements (the kernels move them as one 2-wide vector)
    auto misaligned = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) % sizeof(Z)) != 0; };
    if (misaligned(C) || (k > 0 && (misaligned(A) || misaligned(B)))) return TMVX_ERR_ARG;
    const i64 na = m * k, nb = k * n, np = m * n;   // k == 0: only P's space (unused)
    R* ws = nullptr;
    const size_t bytes = sizeof(R) * (size_t)(2 * na + 2 * nb + 2 * np);
    if (hipMallocAsync((void**)&ws, bytes, s) != hipSuccess) return TMVX_ERR_ALLOC;
    R* Ar = ws;
    R* Ai = Ar + na;
    R* Br = Ai + na;
    R* Bi = Br + nb;
    R* Pr = Bi + nb;
    R* Pi = Pr + np;
    int rc = TMVX_OK;
    if (k > 0) {
        // A: m x k, lines are columns (col-major) or rows (row-major)
        const i64 a_out = la ? m : k, a_in = la ? k : m;
        const i64 b_out = lb ? k : n, b_in = lb ? n : k;
        hipLaunchKernelGGL(split_kernel<R>, grid2(a_in, a_out), dim3(256), 0, s, a_out, a_in,
                           reinterpret_cast<const Z*>(A), lda, Ar, Ai);
        rc = hip_status(hipGetLastError());
        if (rc == TMVX_OK) {
            hipLaunchKernelGGL(split_kernel<R>, grid2(b_in, b_out), dim3(256), 0, s, b_out, b_in,
                               reinterpret_cast<const Z*>(B), ldb, Br, Bi);
            rc = hip_status(hipGetLastError());
        }
    }
    const i64 lda2 = la ? k : m, ldb2 = lb ? n : k;
    const R sA = conj_a ? R(-1) : R(1), sB = conj_b ? R(-1) : R(1), sR = -sA * sB;
    if (rc == TMVX_OK && k > 0) {
        rc = gemm_launch<R>(la, lb, m, n, k, R(1), Ar, lda2, Br, ldb2, R(0), Pr, m, s);
        if (rc == TMVX_OK)
            rc = gemm_launch<R>(la, lb, m, n, k, sR, Ai, lda2, Bi, ldb2, R(1), Pr, m, s);
        if (rc == TMVX_OK)
            rc = gemm_launch<R>(la, lb, m, n, k, sB, Ar, lda2, Bi, ldb2, R(0), Pi, m, s);
        if (rc == TMVX_OK)
            rc = gemm_launch<R>(la, lb, m, n, k, sA, Ai, lda2, Br, ldb2, R(1), Pi, m, s);
    }
    if (rc == TMVX_OK) {
        hipLaunchKernelGGL(combine_kernel<R>, grid2(m, n), dim3(256), 0, s, m, n, Pr, Pi, k > 0,
                           ar, ai, br, bi, br == R(0) && bi == R(0), reinterpret_cast<Z*>(C),
                           ldc);
        rc = hip_status(hipGetLastError());
    }
    hipFreeAsync(ws, s);
    return rc;
}

template int gemm_complex<double>(int, int, int, int, i64, i64, i64, double, double,
                                  const double*, i64, const double*, i64, double, double,
                                  double*, i64, hipStream_t);
template int gemm_complex<float>(int, int, int, int, i64, i64, i64, float, float,
                                 const float*, i64, const float*, i64, float, float, float*,
                                 i64, hipStream_t);

} // namespace tmvx

using namespace tmvx;

extern "C" {

int tmvx_gemm_z(int layout_a, int layout_b, int conj_a, int conj_b, int64_t m, int64_t n,
                int64_t k, double alpha_re, double alpha_im, const double* A, int64_t lda,
                const double* B, int64_t ldb, double beta_re, double beta_im, double* C,
                int64_t ldc, void* stream)
{
    return gemm_complex<double>(layout_a, layout_b, conj_a, conj_b, m, n, k, alpha_re, alpha_im,
                                A, lda, B, ldb, beta_re, beta_im, C, ldc, (hipStream_t)stream);
}

int tmvx_gemm_c(int layout_a, int layout_b, int conj_a, int conj_b, int64_t m, int64_t n,
                int64_t k, float alpha_re, float alpha_im, const float* A, int64_t lda,
                const float* B, int64_t ldb, float beta_re, float beta_im, float* C,
                int64_t ldc, void* stream)
{
    return gemm_complex<float>(layout_a, layout_b, conj_a, conj_b, m, n, k, alpha_re, alpha_im,
                               A, lda, B, ldb, beta_re, beta_im, C, ldc, (hipStream_t)stream);
}

} // extern "C"
