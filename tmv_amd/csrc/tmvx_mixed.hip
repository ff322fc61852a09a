// Mixed-precision LU solve with iterative refinement (FP32 factorisation on
// MFMA, FP64 residual and update on the GPU).
//
// Extends TMV's x = b/A (LUDiv::doLDivEq -> LU_LDivEq, src/TMV_LUD.cpp:
// 92-100, src/TMV_LUDiv.cpp:185-226) with iterative refinement: the O(n^3)
// work runs in low precision and the two steps whose stability needs the
// working precision -- r = b - A x and x += d -- stay in FP64.
//
// Two refinement methods share the factorisation:
//   classical:  d = (P^T L U)^-1 r in FP32, x += d;
//   GMRES-IR (Carson & Higham, SIAM J. Sci. Comput. 40 (2018) A817-A847):
//               A d = r is solved by GMRES in FP64 on the left-preconditioned
//               system M^-1 A d = M^-1 r, M = P^T L U the low-precision
//               factors; x += d.  It converges where classical refinement
//               does not (16-bit trailing updates, kappa u_low >~ 1).
#include "tmvx_lu.h"
#include <cmath>
#include <cstring>
#include <vector>

namespace tmvx {

// GMRES restart length (a build option for experiments: -DTMVX_GMRES_RESTART=k).
#ifndef TMVX_GMRES_RESTART
#define TMVX_GMRES_RESTART 50
#endif

// A32 = (float) A, and per-(row, column chunk) partial sums of |A| for the
// infinity norm.  grid = (ceil(n / 256), nchunk).
__global__ void lowprec_copy_kernel(i64 n, const double* __restrict__ A, i64 lda,
                                    float* __restrict__ A32, double* __restrict__ rowpart,
                                    int nchunk)
{
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const i64 cw = (n + nchunk - 1) / nchunk;
    const i64 j0 = (i64)blockIdx.y * cw, j1 = j0 + cw < n ? j0 + cw : n;
    double s = 0.0;
    for (i64 j = j0; j < j1; ++j) {
        const double a = A[i + j * lda];
        A32[i + j * n] = (float)a;
        s += fabs(a);
    }
    rowpart[i + (i64)blockIdx.y * n] = s;
}

// ||A||_inf from the partial row sums (one value; atomicMax on the bit
// pattern, valid for non-negative doubles).
__global__ void rownorm_kernel(i64 n, const double* __restrict__ rowpart, int nchunk,
                               unsigned long long* __restrict__ out)
{
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    double s = 0.0;
    if (i < n)
        for (int c = 0; c < nchunk; ++c) s += rowpart[i + (i64)c * n];
    // wave max, then one atomic per wave
    for (int off = 32; off >= 1; off >>= 1) s = fmax(s, __shfl_xor(s, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(s));
}

// Column max-abs of an n x nr matrix into out[j] (bit-pattern atomicMax).
__global__ void colmax_kernel(i64 n, i64 nr, const double* __restrict__ M, i64 ldm,
                              unsigned long long* __restrict__ out)
{
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    const i64 j = blockIdx.y;
    double v = (i < n && j < nr) ? fabs(M[i + j * ldm]) : 0.0;
    for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    if ((threadIdx.x & 63) == 0 && j < nr)
        atomicMax(out + j, (unsigned long long)__double_as_longlong(v));
}

// Residual R = B - A X in FP64.  Split over column chunks: each workgroup
// takes 256 rows and one chunk of columns and writes partial products;
// residual_reduce_kernel adds them in a fixed order (deterministic).
template <int NR>
__global__ void __launch_bounds__(256)
residual_partial_kernel(i64 n, i64 nr, const double* __restrict__ A, i64 lda,
                        const double* __restrict__ X, i64 ldx, double* __restrict__ P,
                        int nchunk)
{
    constexpr int KT = 64;                     // columns staged per LDS tile
    __shared__ double xs[KT][NR];
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    const i64 cw = (n + nchunk - 1) / nchunk;
    const i64 j0 = (i64)blockIdx.y * cw, j1 = j0 + cw < n ? j0 + cw : n;
    double acc[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) acc[r] = 0.0;
    for (i64 jt = j0; jt < j1; jt += KT) {
        const int kt = (int)(j1 - jt < KT ? j1 - jt : KT);
        __syncthreads();
        for (int e = threadIdx.x; e < KT * NR; e += blockDim.x) {
            const int k = e / NR, r = e % NR;
            xs[k][r] = (k < kt && r < nr) ? X[(jt + k) + r * ldx] : 0.0;
        }
        __syncthreads();
        if (i < n) {
            for (int k = 0; k < kt; ++k) {
                const double a = A[i + (jt + k) * lda];
#pragma unroll
                for (int r = 0; r < NR; ++r) acc[r] = fma(a, xs[k][r], acc[r]);
            }
        }
    }
    if (i < n)
        for (int r = 0; r < NR && r < nr; ++r) P[i + ((i64)blockIdx.y * nr + r) * n] = acc[r];
}

__global__ void residual_reduce_kernel(i64 n, i64 nr, const double* __restrict__ B, i64 ldb,
                                       const double* __restrict__ P, int nchunk,
                                       double* __restrict__ R)
{
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    const i64 r = blockIdx.y;
    if (i >= n || r >= nr) return;
    double s = 0.0;
    for (int c = 0; c < nchunk; ++c) s += P[i + ((i64)c * nr + r) * n];
    R[i + r * n] = B[i + r * ldb] - s;
}

// R32 = (float)(R / scale[j]) (scale = max|R(:,j)|, 1 when zero).
__global__ void scale_down_kernel(i64 n, i64 nr, const double* __restrict__ R,
                                  const unsigned long long* __restrict__ rmax,
                                  float* __restrict__ R32)
{
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    const i64 j = blockIdx.y;
    if (i >= n || j >= nr) return;
    double s = __longlong_as_double((long long)rmax[j]);
    if (!(s > 0.0)) s = 1.0;
    R32[i + j * n] = (float)(R[i + j * n] / s);
}

// X += scale[j] * D32 (FP64 update).
__global__ void update_kernel(i64 n, i64 nr, double* __restrict__ X, i64 ldx,
                              const float* __restrict__ D32,
                              const unsigned long long* __restrict__ rmax)
{
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    const i64 j = blockIdx.y;
    if (i >= n || j >= nr) return;
    double s = __longlong_as_double((long long)rmax[j]);
    if (!(s > 0.0)) s = 1.0;
    X[i + j * ldx] += s * (double)D32[i + j * n];
}

__global__ void zero_x_kernel(i64 n, i64 nr, double* __restrict__ X, i64 ldx)
{
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    const i64 j = blockIdx.y;
    if (i < n && j < nr) X[i + j * ldx] = 0.0;
}

// w = A v from the column-chunk partials of residual_partial_kernel<1>
// (fixed summation order, as residual_reduce_kernel).
__global__ void matvec_reduce_kernel(i64 n, const double* __restrict__ P, int nchunk,
                                     double* __restrict__ w)
{
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int c = 0; c < nchunk; ++c) s += P[i + (i64)c * n];
    w[i] = s;
}

// h[c] = V(:, c) . w for c < k (one workgroup per column, fixed-order
// reduction: deterministic).  out[c] is written, not accumulated.
__global__ void __launch_bounds__(256)
vdot_kernel(i64 n, const double* __restrict__ V, i64 ldv, const double* __restrict__ w,
            double* __restrict__ out)
{
    __shared__ double part[256];
    const i64 c = blockIdx.x;
    const double* v = V + c * ldv;
    double s = 0.0;
    for (i64 i = threadIdx.x; i < n; i += 256) s = fma(v[i], w[i], s);
    part[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = part[0];
}

// w -= V(:, 0:k) h, h on the device.
__global__ void vsub_kernel(i64 n, int k, const double* __restrict__ V, i64 ldv,
                            const double* __restrict__ h, double* __restrict__ w)
{
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = w[i];
    for (int c = 0; c < k; ++c) s = fma(-V[i + (i64)c * ldv], h[c], s);
    w[i] = s;
}

// out = alpha * w.
__global__ void vscale_kernel(i64 n, const double* __restrict__ w, const double* __restrict__ alpha,
                              double* __restrict__ out)
{
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = alpha[0] * w[i];
}

static int launch_residual(i64 n, i64 nr, const double* A, i64 lda, const double* X, i64 ldx,
                           const double* B, i64 ldb, double* P, int nchunk, double* R,
                           hipStream_t s)
{
    const dim3 grid((unsigned)((n + 255) / 256), (unsigned)nchunk);
    for (i64 r0 = 0; r0 < nr; r0 += 16) {
        const i64 rr = nr - r0 < 16 ? nr - r0 : 16;
        if (rr == 1)
            hipLaunchKernelGGL(residual_partial_kernel<1>, grid, dim3(256), 0, s, n, rr, A, lda,
                               X + r0 * ldx, ldx, P, nchunk);
        else if (rr <= 4)
            hipLaunchKernelGGL(residual_partial_kernel<4>, grid, dim3(256), 0, s, n, rr, A, lda,
                               X + r0 * ldx, ldx, P, nchunk);
        else
            hipLaunchKernelGGL(residual_partial_kernel<16>, grid, dim3(256), 0, s, n, rr, A, lda,
                               X + r0 * ldx, ldx, P, nchunk);
        TMVX_CHECK_LAUNCH();
        hipLaunchKernelGGL(residual_reduce_kernel, dim3((unsigned)((n + 255) / 256), (unsigned)rr),
                           dim3(256), 0, s, n, rr, B + r0 * ldb, ldb, P, nchunk, R + r0 * n);
        TMVX_CHECK_LAUNCH();
    }
    return TMVX_OK;
}

// GMRES-IR for one right-hand side: A d = r (FP64) by left-preconditioned
// GMRES, M^-1 = the low-precision solve.  State lives in the caller's
// workspace; the host keeps the (k+1) x k Hessenberg matrix and the Givens
// rotations.  Arnoldi uses classical Gram-Schmidt twice (CGS2) so that each
// inner step costs one host synchronisation.  Adds d to x.
struct GmresWork {
    double* V;                 // n x (M + 1) Krylov basis
    double* w;                 // n
    double* t;                 // n (A v)
    double* h;                 // 2 (M + 1) + 8: CGS coefficients, norms, scalars
    double* P;                 // n x nchunk matvec partials
    int nchunk;
    int M;                     // restart length (columns of V - 1)
};

static int gmres_correction(i64 n, const double* A, i64 lda, const float* LU, const i64* perm,
                            const double* r, double* x, double tol, int max_inner, GmresWork& g,
                            int* d_flag, int* inner_done, hipStream_t s)
{
    const dim3 g1((unsigned)((n + 255) / 256)), b1(256);
    const int M = g.M;
    // z = M^-1 y with FP64 arithmetic on the FP32-stored factors: applied
    // in FP32 the preconditioner's own error (about kappa n u_32) would
    // leave GMRES a perturbed operator and stall it near 1e-4 for BF16
    // updates at kappa = 1e4.
    auto precond = [&](const double* y, double* z) -> int {
        TMVX_TRY((trsv_multi<double, float>(0, 0, 1, n, 1, LU, n, y, n, perm, z, n, nullptr,
                                            d_flag, s)));
        return trsv_multi<double, float>(1, 0, 0, n, 1, LU, n, z, n, nullptr, z, n, nullptr,
                                         d_flag, s);
    };
    auto matvec = [&](const double* v, double* out) -> int {
        hipLaunchKernelGGL(residual_partial_kernel<1>, dim3(g1.x, g.nchunk), b1, 0, s, n, (i64)1,
                           A, lda, v, n, g.P, g.nchunk);
        TMVX_CHECK_LAUNCH();
        hipLaunchKernelGGL(matvec_reduce_kernel, g1, b1, 0, s, n, g.P, g.nchunk, out);
        TMVX_CHECK_LAUNCH();
        return TMVX_OK;
    };
    auto norm_to_host = [&](const double* v, double* out) -> int {
        hipLaunchKernelGGL(vdot_kernel, dim3(1), b1, 0, s, n, v, n, v, g.h + 2 * (M + 1));
        TMVX_CHECK_LAUNCH();
        TMVX_HIP(hipMemcpyAsync(out, g.h + 2 * (M + 1), sizeof(double), hipMemcpyDeviceToHost, s));
        TMVX_HIP(hipStreamSynchronize(s));
        *out = std::sqrt(*out);
        return TMVX_OK;
    };
    // v0 = M^-1 r / beta
    TMVX_TRY(precond(r, g.V));
    double beta = 0.0;
    TMVX_TRY(norm_to_host(g.V, &beta));
    if (!(beta > 0.0)) return TMVX_OK;                       // r = 0 (or M^-1 r = 0)
    if (!std::isfinite(beta)) return TMVX_ERR_NOT_CONVERGED;
    std::vector<double> H((size_t)(M + 1) * M, 0.0), cs(M), sn(M), gv(M + 1, 0.0), hh(2 * (M + 1) + 1);
    gv[0] = beta;
    {
        const double inv = 1.0 / beta;
        TMVX_HIP(hipMemcpyAsync(g.h + 2 * (M + 1) + 1, &inv, sizeof(double), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(vscale_kernel, g1, b1, 0, s, n, g.V, g.h + 2 * (M + 1) + 1, g.V);
        TMVX_CHECK_LAUNCH();
    }
    int k = 0;
    for (; k < M && k < max_inner; ++k) {
        double* vk = g.V + (i64)k * n;
        double* vn = g.V + (i64)(k + 1) * n;
        // w = M^-1 A v_k
        TMVX_TRY(matvec(vk, g.t));
        TMVX_TRY(precond(g.t, g.w));
        // CGS2 against v_0..v_k: h1 = V^T w, w -= V h1, h2 = V^T w, w -= V h2
        hipLaunchKernelGGL(vdot_kernel, dim3((unsigned)(k + 1)), b1, 0, s, n, g.V, n, g.w, g.h);
        TMVX_CHECK_LAUNCH();
        hipLaunchKernelGGL(vsub_kernel, g1, b1, 0, s, n, k + 1, g.V, n, g.h, g.w);
        TMVX_CHECK_LAUNCH();
        hipLaunchKernelGGL(vdot_kernel, dim3((unsigned)(k + 1)), b1, 0, s, n, g.V, n, g.w,
                           g.h + (M + 1));
        TMVX_CHECK_LAUNCH();
        hipLaunchKernelGGL(vsub_kernel, g1, b1, 0, s, n, k + 1, g.V, n, g.h + (M + 1), g.w);
        TMVX_CHECK_LAUNCH();
        hipLaunchKernelGGL(vdot_kernel, dim3(1), b1, 0, s, n, g.w, n, g.w, g.h + 2 * (M + 1));
        TMVX_CHECK_LAUNCH();
        TMVX_HIP(hipMemcpyAsync(hh.data(), g.h, sizeof(double) * (2 * (M + 1) + 1),
                                hipMemcpyDeviceToHost, s));
        TMVX_HIP(hipStreamSynchronize(s));
        double* Hk = &H[(size_t)k * (M + 1)];                 // column k of H
        for (int i = 0; i <= k; ++i) Hk[i] = hh[i] + hh[(M + 1) + i];
        const double hn = std::sqrt(hh[2 * (M + 1)]);
        Hk[k + 1] = hn;
        // previous rotations, then a new one zeroing H(k+1, k)
        for (int i = 0; i < k; ++i) {
            const double a = Hk[i], b = Hk[i + 1];
            Hk[i] = cs[i] * a + sn[i] * b;
            Hk[i + 1] = -sn[i] * a + cs[i] * b;
        }
        const double a = Hk[k], b = Hk[k + 1];
        const double rho = std::hypot(a, b);
        cs[k] = rho > 0.0 ? a / rho : 1.0;
        sn[k] = rho > 0.0 ? b / rho : 0.0;
        Hk[k] = rho;
        Hk[k + 1] = 0.0;
        gv[k + 1] = -sn[k] * gv[k];
        gv[k] = cs[k] * gv[k];
        if (!std::isfinite(rho)) return TMVX_ERR_NOT_CONVERGED;
        const bool done = std::fabs(gv[k + 1]) <= tol * beta || !(hn > 0.0);
        if (!done && k + 1 < M && k + 1 < max_inner) {
            const double inv = 1.0 / hn;
            TMVX_HIP(hipMemcpyAsync(g.h + 2 * (M + 1) + 1, &inv, sizeof(double),
                                    hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(vscale_kernel, g1, b1, 0, s, n, g.w, g.h + 2 * (M + 1) + 1, vn);
            TMVX_CHECK_LAUNCH();
        }
        if (done) { ++k; break; }
    }
    if (k > M) k = M;
    // y = H(0:k, 0:k)^-1 g(0:k) (upper triangular), x += V y
    std::vector<double> y(k, 0.0);
    for (int i = k - 1; i >= 0; --i) {
        double v = gv[i];
        for (int j = i + 1; j < k; ++j) v -= H[(size_t)j * (M + 1) + i] * y[j];
        const double d = H[(size_t)i * (M + 1) + i];
        y[i] = d != 0.0 ? v / d : 0.0;
    }
    for (int i = 0; i < k; ++i) y[i] = -y[i];                 // vsub adds V (-(-y))
    TMVX_HIP(hipMemcpyAsync(g.h, y.data(), sizeof(double) * k, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(vsub_kernel, g1, b1, 0, s, n, k, g.V, n, g.h, x);
    TMVX_CHECK_LAUNCH();
    *inner_done += k;
    return TMVX_OK;
}

int gesv_mixed_f32(int lowprec, i64 n, i64 nrhs, const double* A, i64 lda, const double* B, i64 ldb,
                   double* X, i64 ldx, double tol, int max_iter, tmvx_ir_info* info,
                   hipStream_t s)
{
    if (!info) return TMVX_ERR_ARG;
    *info = tmvx_ir_info{0, 0, 0, 0, 0.0, 0.0, 0.0};   // iterations, converged, ...
    if (n < 0 || nrhs < 0 || lda < (n > 0 ? n : 1) || ldb < (n > 0 ? n : 1) ||
        ldx < (n > 0 ? n : 1) || max_iter < 0)
        return TMVX_ERR_ARG;
    if (n == 0 || nrhs == 0) { info->converged = 1; return TMVX_OK; }
    if (!A || !B || !X) return TMVX_ERR_ARG;
    if (nrhs > 65535) return TMVX_ERR_UNSUPPORTED;

    const int prec = lowprec & 0xf;
    const bool gmres = (lowprec & TMVX_REFINE_GMRES) != 0 ||
                       (prec != TMVX_LOWPREC_F32 && (lowprec & TMVX_REFINE_CLASSIC) == 0);
    const int nchunk = 16;
    constexpr int GM = TMVX_GMRES_RESTART;   // GMRES restart length
    constexpr double GTOL = 1e-8;            // inner relative tolerance
    // Workspace: A32 | piv | perm | R | P | R32 | rowpart | norms | GMRES.
    const size_t nn = (size_t)n * (size_t)n;
    const size_t sz_a32 = sizeof(float) * nn;
    const size_t sz_piv = sizeof(i64) * (size_t)n;
    const size_t sz_r = sizeof(double) * (size_t)n * (size_t)nrhs;
    const size_t sz_p = sizeof(double) * (size_t)n * (size_t)(nrhs < 16 ? nrhs : 16) * nchunk;
    const size_t sz_r32 = sizeof(float) * (size_t)n * (size_t)nrhs;
    const size_t sz_rp = sizeof(double) * (size_t)n * nchunk;
    const size_t sz_nrm = sizeof(unsigned long long) * (size_t)(3 * nrhs + 8) + 64;
    const size_t sz_gv = gmres ? sizeof(double) * (size_t)n * (GM + 3) : 0;   // V, w, t
    const size_t sz_gh = gmres ? sizeof(double) * (2 * (GM + 1) + 8) : 0;
    auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
    const size_t total = al(sz_a32) + 2 * al(sz_piv) + al(sz_r) + al(sz_p) + al(sz_r32) +
                         al(sz_rp) + al(sz_nrm) + al(sz_gv) + al(sz_gh);
    char* ws = nullptr;
    if (hipMallocAsync((void**)&ws, total, s) != hipSuccess) return TMVX_ERR_ALLOC;
    char* q = ws;
    float* A32 = reinterpret_cast<float*>(q); q += al(sz_a32);
    i64* piv = reinterpret_cast<i64*>(q); q += al(sz_piv);
    i64* perm = reinterpret_cast<i64*>(q); q += al(sz_piv);
    double* R = reinterpret_cast<double*>(q); q += al(sz_r);
    double* P = reinterpret_cast<double*>(q); q += al(sz_p);
    float* R32 = reinterpret_cast<float*>(q); q += al(sz_r32);
    double* rowpart = reinterpret_cast<double*>(q); q += al(sz_rp);
    unsigned long long* nrm = reinterpret_cast<unsigned long long*>(q); q += al(sz_nrm);
    GmresWork gw{};
    if (gmres) {
        gw.V = reinterpret_cast<double*>(q);
        gw.w = gw.V + (size_t)n * (GM + 1);
        gw.t = gw.w + n;
        q += al(sz_gv);
        gw.h = reinterpret_cast<double*>(q); q += al(sz_gh);
        gw.P = P;
        gw.nchunk = nchunk;
        gw.M = GM;
    }
    unsigned long long* d_anorm = nrm;               // [1]
    unsigned long long* d_bmax = nrm + 8;            // [nrhs]
    unsigned long long* d_rmax = d_bmax + nrhs;      // [nrhs]
    unsigned long long* d_xmax = d_rmax + nrhs;      // [nrhs]
    int* d_flag = reinterpret_cast<int*>(nrm + 8 + 3 * nrhs);

    hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventCreate(&e2);
    int rc = TMVX_OK;
    const dim3 gcol((unsigned)((n + 255) / 256), (unsigned)nrhs);
    auto chk = [&]() { if (rc == TMVX_OK) rc = hip_status(hipGetLastError()); };

    rc = zero_bytes(nrm, sz_nrm, s);          // norms and the solve flag
    (void)hipEventRecord(e0, s);
    // 1. A32 = fl32(A), ||A||_inf, ||b||_inf per column.
    if (rc == TMVX_OK) {
        hipLaunchKernelGGL(lowprec_copy_kernel, dim3((unsigned)((n + 255) / 256), nchunk), dim3(256),
                           0, s, n, A, lda, A32, rowpart, nchunk);
        chk();
        hipLaunchKernelGGL(rownorm_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n,
                           rowpart, nchunk, d_anorm);
        chk();
        hipLaunchKernelGGL(colmax_kernel, gcol, dim3(256), 0, s, n, nrhs, B, ldb, d_bmax);
        chk();
    }
    // 2. FP32 LU with partial pivoting (synchronises).
    i64 finfo = 0;
    // partial pivoting (the tournament-pivoted panel of tmvx_tslu.hip is
    // within 3% of it on this GPU: its local selection costs about as much
    // per column as the partial-pivoting panel's exchange and local work)
    if (rc == TMVX_OK) rc = getrf_impl<float>(n, A32, n, piv, perm, &finfo, s, prec, 0);
    (void)hipEventRecord(e1, s);
    info->factor_info = (int)finfo;
    if (rc == TMVX_OK && finfo != 0) rc = TMVX_ERR_SINGULAR;

    // 3. Refinement from x0 = 0 (r0 = b).
    std::vector<unsigned long long> h(3 * nrhs + 1);
    double anorm = 0.0, worst = INFINITY;
    int it = 0;
    if (rc == TMVX_OK) {
        hipLaunchKernelGGL(zero_x_kernel, gcol, dim3(256), 0, s, n, nrhs, X, ldx);
        chk();
        hipLaunchKernelGGL(residual_reduce_kernel, gcol, dim3(256), 0, s, n, nrhs, B, ldb, P, 0, R);
        chk();
    }
    while (rc == TMVX_OK) {
        // max|r|, max|x| per column
        rc = zero_bytes(d_rmax, sizeof(unsigned long long) * 2 * nrhs, s);
        if (rc != TMVX_OK) break;
        hipLaunchKernelGGL(colmax_kernel, gcol, dim3(256), 0, s, n, nrhs, R, n, d_rmax);
        chk();
        hipLaunchKernelGGL(colmax_kernel, gcol, dim3(256), 0, s, n, nrhs, X, ldx, d_xmax);
        chk();
        if (rc != TMVX_OK) break;
        if (hipMemcpyAsync(h.data(), d_anorm, sizeof(unsigned long long), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipMemcpyAsync(h.data() + 1, d_bmax, sizeof(unsigned long long) * 3 * nrhs,
                           hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) { rc = TMVX_ERR_HIP; break; }
        auto dv = [](unsigned long long b) { double d; memcpy(&d, &b, 8); return d; };
        anorm = dv(h[0]);
        worst = 0.0;
        for (i64 j = 0; j < nrhs; ++j) {
            const double bm = dv(h[1 + j]), rm = dv(h[1 + nrhs + j]), xm = dv(h[1 + 2 * nrhs + j]);
            const double den = anorm * xm + bm;
            const double be = den > 0.0 ? rm / den : (rm > 0.0 ? INFINITY : 0.0);
            if (!(be <= worst)) worst = be;     // NaN propagates as not converged
        }
        if (worst <= tol) { info->converged = 1; break; }
        if (it >= max_iter) break;
        if (gmres) {
            // A d = r by preconditioned GMRES, column by column; x += d
            for (i64 j = 0; j < nrhs && rc == TMVX_OK; ++j)
                rc = gmres_correction(n, A, lda, A32, perm, R + j * n, X + j * ldx, GTOL, GM, gw,
                                      d_flag, &info->inner_iterations, s);
            if (rc != TMVX_OK) break;
        } else {
            // d = (LU)^-1 P r / s in FP32;  x += s d
            hipLaunchKernelGGL(scale_down_kernel, gcol, dim3(256), 0, s, n, nrhs, R, d_rmax, R32);
            chk();
            if (rc == TMVX_OK) rc = getrs_perm<float>(0, n, nrhs, A32, n, perm, R32, n, d_flag, s);
            if (rc != TMVX_OK) break;
            hipLaunchKernelGGL(update_kernel, gcol, dim3(256), 0, s, n, nrhs, X, ldx, R32, d_rmax);
            chk();
        }
        // r = b - A x
        if (rc == TMVX_OK) rc = launch_residual(n, nrhs, A, lda, X, ldx, B, ldb, P, nchunk, R, s);
        ++it;
    }
    (void)hipEventRecord(e2, s);
    int hflag = 0;
    if (hipMemcpyAsync(&hflag, d_flag, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess &&
        rc == TMVX_OK)
        rc = TMVX_ERR_HIP;
    hipFreeAsync(ws, s);
    if (hipStreamSynchronize(s) != hipSuccess && rc == TMVX_OK) rc = TMVX_ERR_HIP;
    if (rc == TMVX_OK && (hflag & 2)) rc = TMVX_ERR_TIMEOUT;
    float f_ms = 0.f, r_ms = 0.f;
    if (hipEventElapsedTime(&f_ms, e0, e1) == hipSuccess) info->factor_ms = f_ms;
    if (hipEventElapsedTime(&r_ms, e1, e2) == hipSuccess) info->refine_ms = r_ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(e2);
    info->iterations = it;
    info->backward_error = worst;
    if (rc == TMVX_OK && !info->converged) rc = TMVX_ERR_NOT_CONVERGED;
    return rc;
}

} // namespace tmvx

extern "C" int tmvx_gesv_mixed_f64(int lowprec, int64_t n, int64_t nrhs, const double* A,
                                   int64_t lda, const double* B, int64_t ldb, double* X,
                                   int64_t ldx, double tol, int max_iter, tmvx_ir_info* info,
                                   void* stream)
{
    const int prec = lowprec & 0xf, flags = lowprec & ~0xf;
    if ((prec != TMVX_LOWPREC_F32 && prec != TMVX_LOWPREC_F16 && prec != TMVX_LOWPREC_BF16) ||
        (flags & ~(TMVX_REFINE_GMRES | TMVX_REFINE_CLASSIC)) != 0 ||
        flags == (TMVX_REFINE_GMRES | TMVX_REFINE_CLASSIC))
        return TMVX_ERR_UNSUPPORTED;
    return tmvx::gesv_mixed_f32(lowprec, n, nrhs, A, lda, B, ldb, X, ldx, tol, max_iter, info,
                                (hipStream_t)stream);
}
