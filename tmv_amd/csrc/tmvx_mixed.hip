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
//
// Two factorisations:
//   partial pivoting (default):  P A = L U on an FP32 copy of A (the blocked
//               LU of tmvx_lu.hip with TMV's pivot rule);
//   random butterfly transformation (RBT) + no interchanges, on request (D. S. Parker,
//               UCLA CSD-950023, 1995; M. Baboulin, J. Dongarra,
//               J. Herrmann, S. Tomov, ACM TOMS 39(2), 2013, art. 8):
//               A' = U^T A V with U, V random recursive butterflies of depth
//               two, then A' = L U without interchanges; A x = b becomes
//               x = V A'^-1 U^T b.  No panel needs a grid-wide pivot
//               search, so the factorisation runs at the speed of its
//               trailing updates.  Refinement against the original A in FP64
//               removes the effect of the missing pivoting on accuracy; a
//               breakdown (zero or non-finite pivot), a stagnating or a
//               non-converging refinement falls back to partial pivoting.
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

// ---- random butterfly transformation --------------------------------
//
// A butterfly of order m (even) is B = 2^-1/2 [R0 R1; R0 -R1] with R0, R1
// random diagonal matrices of order m/2; a recursive butterfly of depth two
// is W = diag(B1, B2) B, B of order N and B1, B2 of order N/2 (N = n rounded
// up to a multiple of 4; rows and columns past n are an identity block
// scaled by `pad`).  W only mixes the four positions k + a N/4 (a = 0..3,
// k < N/4): restricted to them it is the 4 x 4 matrix
//   1/2 [ s0 r0a   s1 r0b   s0 r1a   s1 r1b ]
//       [ s0 r0a  -s1 r0b   s0 r1a  -s1 r1b ]
//       [ t0 r0a   t1 r0b  -t0 r1a  -t1 r1b ]
//       [ t0 r0a  -t1 r0b  -t0 r1a   t1 r1b ]
// with r0a = R0[k], r0b = R0[k + N/4] (r1 likewise from R1), s0, s1 the
// diagonals of B1 and t0, t1 those of B2 at k.  Every diagonal entry is
// exp(u / 10), u uniform in [-1/2, 1/2) (Baboulin et al.'s choice), from a
// SplitMix64 hash of (seed, diagonal, index).  U uses diagonals 0-5, V 6-11.
__device__ __forceinline__ double rbt_entry(unsigned long long seed, int diag, i64 idx)
{
    unsigned long long z =
        seed + 0x9E3779B97F4A7C15ull * ((((unsigned long long)diag) << 40) + (unsigned long long)idx + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    const double u = (double)(z >> 11) * 0x1.0p-53 - 0.5;
    return exp(0.1 * u);
}

// W[(set * N4 + k) * 16 + 4 a + b] = the 4 x 4 group matrix of butterfly
// `set` (0: U, 1: V) at k.
__global__ void rbt_table_kernel(unsigned long long seed, i64 N4, double* __restrict__ W)
{
    const i64 g = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= 2 * N4) return;
    const int set = g >= N4 ? 1 : 0;
    const i64 k = g - set * N4;
    const int d0 = 6 * set;
    const double r0a = rbt_entry(seed, d0 + 0, k), r0b = rbt_entry(seed, d0 + 0, k + N4);
    const double r1a = rbt_entry(seed, d0 + 1, k), r1b = rbt_entry(seed, d0 + 1, k + N4);
    const double s0 = rbt_entry(seed, d0 + 2, k), s1 = rbt_entry(seed, d0 + 3, k);
    const double t0 = rbt_entry(seed, d0 + 4, k), t1 = rbt_entry(seed, d0 + 5, k);
    const double w[16] = {s0 * r0a, s1 * r0b,  s0 * r1a,  s1 * r1b,
                          s0 * r0a, -s1 * r0b, s0 * r1a,  -s1 * r1b,
                          t0 * r0a, t1 * r0b,  -t0 * r1a, -t1 * r1b,
                          t0 * r0a, -t1 * r0b, -t0 * r1a, t1 * r1b};
    double* out = W + g * 16;
#pragma unroll
    for (int e = 0; e < 16; ++e) out[e] = 0.5 * w[e];
}

// Scale of the identity block that pads n up to N: a power of two near the
// largest |A(i, i)| (else the largest |A(i, 0)|, else 1), so that the block
// is of the matrix's own magnitude.  One workgroup of 256.
__global__ void __launch_bounds__(256)
rbt_pad_kernel(i64 n, const double* __restrict__ A, i64 lda, double* __restrict__ pad)
{
    __shared__ double sd[256], sc[256];
    double d = 0.0, c = 0.0;
    for (i64 i = threadIdx.x; i < n; i += 256) {
        d = fmax(d, fabs(A[i + i * lda]));
        c = fmax(c, fabs(A[i]));
    }
    sd[threadIdx.x] = d;
    sc[threadIdx.x] = c;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            sd[threadIdx.x] = fmax(sd[threadIdx.x], sd[threadIdx.x + h]);
            sc[threadIdx.x] = fmax(sc[threadIdx.x], sc[threadIdx.x + h]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double v = sd[0] > 0.0 ? sd[0] : sc[0];
        if (!(v > 0.0) || !isfinite(v)) v = 1.0;
        int e;
        (void)frexp(v, &e);
        *pad = ldexp(1.0, e - 1);
    }
}

// A32 (N x N) = fl32(U^T A_pad V), and per-(row, chunk) partial sums of |A|
// for the infinity norm (rowpart[i + chunk n], rows i < n).  Thread k takes
// the four rows k + a N4 and, for each column group l of its chunk, the 16
// entries at (k + a N4, l + b N4): one 4 x 4 product on each side.  Loads
// and stores are coalesced over k; V's group matrix is the same for the
// whole wave.  grid = (ceil(N4 / 256), nchunk).
__global__ void __launch_bounds__(256)
rbt_copy_kernel(i64 n, i64 N4, const double* __restrict__ A, i64 lda,
                const double* __restrict__ W, const double* __restrict__ pad,
                float* __restrict__ A32, double* __restrict__ rowpart, int nchunk)
{
    const i64 k = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N4) return;
    const i64 N = 4 * N4;
    const i64 cw = (N4 + nchunk - 1) / nchunk;
    const i64 l0 = (i64)blockIdx.y * cw, l1 = l0 + cw < N4 ? l0 + cw : N4;
    const double pv = N != n ? *pad : 0.0;
    double wu[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) wu[e] = W[k * 16 + e];
    double rs[4] = {0.0, 0.0, 0.0, 0.0};
    for (i64 l = l0; l < l1; ++l) {
        const double* wv = W + (N4 + l) * 16;
        double a[4][4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const i64 j = l + c * N4;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const i64 i = k + r * N4;
                const bool in = i < n && j < n;
                a[r][c] = in ? A[i + j * lda] : (i == j ? pv : 0.0);
                if (in) rs[r] += fabs(a[r][c]);
            }
        }
        double t[4][4];                            // t = Wu^T a
#pragma unroll
        for (int p = 0; p < 4; ++p)
#pragma unroll
            for (int c = 0; c < 4; ++c)
                t[p][c] = wu[0 * 4 + p] * a[0][c] + wu[1 * 4 + p] * a[1][c] +
                          wu[2 * 4 + p] * a[2][c] + wu[3 * 4 + p] * a[3][c];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const i64 j = l + q * N4;
#pragma unroll
            for (int p = 0; p < 4; ++p) {
                const double o = t[p][0] * wv[0 * 4 + q] + t[p][1] * wv[1 * 4 + q] +
                                 t[p][2] * wv[2 * 4 + q] + t[p][3] * wv[3 * 4 + q];
                A32[(k + p * N4) + j * N] = (float)o;
            }
        }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r)
        if (k + r * N4 < n) rowpart[(k + r * N4) + (i64)blockIdx.y * n] = rs[r];
}

// out(:, j) = U^T [in(:, j) / s_j; 0] (N rows; s_j = max|in(:, j)| from
// rmax, 1 when rmax is null or the column is zero).  grid = (ceil(N4/256), nr).
template <typename TO>
__global__ void rbt_left_kernel(i64 n, i64 N4, i64 nr, const double* __restrict__ in, i64 ldi,
                                const unsigned long long* __restrict__ rmax,
                                const double* __restrict__ W, TO* __restrict__ out, i64 ldo)
{
    const i64 k = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    const i64 j = blockIdx.y;
    if (k >= N4 || j >= nr) return;
    double s = rmax ? __longlong_as_double((long long)rmax[j]) : 1.0;
    if (!(s > 0.0)) s = 1.0;
    double v[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) v[r] = k + r * N4 < n ? in[(k + r * N4) + j * ldi] / s : 0.0;
    const double* wu = W + k * 16;
#pragma unroll
    for (int a = 0; a < 4; ++a)
        out[(k + a * N4) + j * ldo] =
            (TO)(wu[0 * 4 + a] * v[0] + wu[1 * 4 + a] * v[1] + wu[2 * 4 + a] * v[2] +
                 wu[3 * 4 + a] * v[3]);
}

// d = (V y(:, j))[0:n]; x(:, j) += s_j d (accumulate) or x(:, j) = d.
template <typename TI>
__global__ void rbt_right_kernel(i64 n, i64 N4, i64 nr, const TI* __restrict__ y, i64 ldy,
                                 const unsigned long long* __restrict__ rmax,
                                 const double* __restrict__ W, double* __restrict__ x, i64 ldx,
                                 int accumulate)
{
    const i64 l = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    const i64 j = blockIdx.y;
    if (l >= N4 || j >= nr) return;
    double s = rmax ? __longlong_as_double((long long)rmax[j]) : 1.0;
    if (!(s > 0.0)) s = 1.0;
    double v[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) v[c] = (double)y[(l + c * N4) + j * ldy];
    const double* wv = W + (N4 + l) * 16;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const i64 i = l + b * N4;
        if (i >= n) continue;
        const double d = wv[b * 4 + 0] * v[0] + wv[b * 4 + 1] * v[1] + wv[b * 4 + 2] * v[2] +
                         wv[b * 4 + 3] * v[3];
        double* xp = x + i + j * ldx;
        *xp = accumulate ? *xp + s * d : d;
    }
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

// The butterflies of an RBT factorisation: group matrices W (see
// rbt_table_kernel), N = 4 N4 >= n, and an N-vector of FP64 scratch.
struct RbtWork {
    const double* W;
    i64 N4;
    double* tmp;
};

static int gmres_correction(i64 n, const double* A, i64 lda, const float* LU, const i64* perm,
                            const RbtWork* rbt, const double* r, double* x, double tol,
                            int max_inner, GmresWork& g, int* d_flag, int* inner_done,
                            hipStream_t s)
{
    const dim3 g1((unsigned)((n + 255) / 256)), b1(256);
    const int M = g.M;
    // z = M^-1 y with FP64 arithmetic on the FP32-stored factors: applied
    // in FP32 the preconditioner's own error (about kappa n u_32) would
    // leave GMRES a perturbed operator and stall it near 1e-4 for BF16
    // updates at kappa = 1e4.  After an RBT, M^-1 = V (L U)^-1 U^T.  The
    // triangular solves apply the diagonal blocks through their inverses
    // (formed in FP64, off the solves' dependency chain): M^-1 stays one
    // fixed linear operator, within about u_64 times the blocks'
    // condition of the substitution.
    auto precond = [&](const double* y, double* z) -> int {
        if (rbt) {
            const i64 N = 4 * rbt->N4;
            const dim3 g4((unsigned)((rbt->N4 + 255) / 256), 1);
            hipLaunchKernelGGL(rbt_left_kernel<double>, g4, b1, 0, s, n, rbt->N4, (i64)1, y, n,
                               nullptr, rbt->W, rbt->tmp, N);
            TMVX_CHECK_LAUNCH();
            TMVX_TRY((trsv_multi<double, float>(0, 0, 1, N, 1, LU, N, rbt->tmp, N, nullptr,
                                                rbt->tmp, N, nullptr, d_flag, s, 1)));
            TMVX_TRY((trsv_multi<double, float>(1, 0, 0, N, 1, LU, N, rbt->tmp, N, nullptr,
                                                rbt->tmp, N, nullptr, d_flag, s, 1)));
            hipLaunchKernelGGL(rbt_right_kernel<double>, g4, b1, 0, s, n, rbt->N4, (i64)1,
                               rbt->tmp, N, nullptr, rbt->W, z, n, 0);
            TMVX_CHECK_LAUNCH();
            return TMVX_OK;
        }
        TMVX_TRY((trsv_multi<double, float>(0, 0, 1, n, 1, LU, n, y, n, perm, z, n, nullptr,
                                            d_flag, s, 1)));
        return trsv_multi<double, float>(1, 0, 0, n, 1, LU, n, z, n, nullptr, z, n, nullptr,
                                         d_flag, s, 1);
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
    *info = tmvx_ir_info{0, 0, 0, 0, 0.0, 0.0, 0.0, TMVX_PIVOT_PARTIAL, 0};
    if (n < 0 || nrhs < 0 || lda < (n > 0 ? n : 1) || ldb < (n > 0 ? n : 1) ||
        ldx < (n > 0 ? n : 1) || max_iter < 0)
        return TMVX_ERR_ARG;
    if (n == 0 || nrhs == 0) { info->converged = 1; return TMVX_OK; }
    if (!A || !B || !X) return TMVX_ERR_ARG;
    if (nrhs > 65535) return TMVX_ERR_UNSUPPORTED;

    const int prec = lowprec & 0xf;
    // GMRES-IR by default where the factors carry 16 bits or fewer
    const bool gmres_first = (lowprec & TMVX_REFINE_GMRES) != 0 ||
                             (prec != TMVX_LOWPREC_F32 && prec != TMVX_LOWPREC_F16X2 &&
                              (lowprec & TMVX_REFINE_CLASSIC) == 0);
    const bool try_rbt = (lowprec & TMVX_PIVOT_RBT) != 0;
    // after an RBT factorisation, classical refinement that contracts the
    // error by less than 2x per step hands over to GMRES-IR on the same
    // factors (unless classical refinement was asked for)
    const bool may_switch = try_rbt && (lowprec & (TMVX_REFINE_GMRES | TMVX_REFINE_CLASSIC)) == 0;
    const bool gmres = gmres_first || may_switch;        // GMRES workspace needed
    const i64 N4 = (n + 3) / 4, N = try_rbt ? 4 * N4 : n;   // RBT order (n padded to 4k)
    constexpr int NCHUNK = 16;               // column chunks of the residual partials
    constexpr int RBT_CHUNKS = 32;           // column chunks of the RBT copy
    constexpr int GM = TMVX_GMRES_RESTART;   // GMRES restart length
    constexpr double GTOL = 1e-8;            // inner relative tolerance
    // Workspace: A32 | piv | perm | R | P | R32 | rowpart | norms | GMRES | RBT.
    const size_t nn = (size_t)N * (size_t)N;
    const size_t sz_a32 = sizeof(float) * nn;
    const size_t sz_piv = sizeof(i64) * (size_t)N;
    const size_t sz_r = sizeof(double) * (size_t)n * (size_t)nrhs;
    const size_t sz_p = sizeof(double) * (size_t)n * (size_t)(nrhs < 16 ? nrhs : 16) * NCHUNK;
    const size_t sz_r32 = sizeof(float) * (size_t)N * (size_t)nrhs;
    const size_t sz_rp = sizeof(double) * (size_t)n * (try_rbt ? RBT_CHUNKS : NCHUNK);
    const size_t sz_nrm = sizeof(unsigned long long) * (size_t)(3 * nrhs + 8) + 64;
    const size_t sz_gv = gmres ? sizeof(double) * (size_t)n * (GM + 3) : 0;   // V, w, t
    const size_t sz_gh = gmres ? sizeof(double) * (2 * (GM + 1) + 8) : 0;
    const size_t sz_w = try_rbt ? sizeof(double) * (size_t)(32 * N4 + 8) : 0;  // W, pad
    const size_t sz_tmp = try_rbt && gmres ? sizeof(double) * (size_t)N : 0;
    auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
    const size_t total = al(sz_a32) + 2 * al(sz_piv) + al(sz_r) + al(sz_p) + al(sz_r32) +
                         al(sz_rp) + al(sz_nrm) + al(sz_gv) + al(sz_gh) + al(sz_w) + al(sz_tmp);
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
        gw.nchunk = NCHUNK;
        gw.M = GM;
    }
    double* Wrbt = reinterpret_cast<double*>(q); q += al(sz_w);
    double* d_pad = try_rbt ? Wrbt + 32 * N4 : nullptr;
    RbtWork rw{Wrbt, N4, reinterpret_cast<double*>(q)};
    q += al(sz_tmp);
    unsigned long long* d_anorm = nrm;               // [1]
    unsigned long long* d_bmax = nrm + 8;            // [nrhs]
    unsigned long long* d_rmax = d_bmax + nrhs;      // [nrhs]
    unsigned long long* d_xmax = d_rmax + nrhs;      // [nrhs]
    int* d_flag = reinterpret_cast<int*>(nrm + 8 + 3 * nrhs);

    hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventCreate(&e2);
    const dim3 gcol((unsigned)((n + 255) / 256), (unsigned)nrhs);
    const dim3 g4col((unsigned)((N4 + 255) / 256), (unsigned)nrhs);
    std::vector<unsigned long long> h(3 * nrhs + 1);
    double worst = INFINITY;
    (void)hipEventRecord(e0, s);

    // One attempt: factorisation (RBT + no interchanges, or partial
    // pivoting) and refinement from x0 = 0.  *fallback is set when an RBT
    // attempt breaks down or its refinement stagnates or does not converge.
    auto attempt = [&](bool rbt, bool* fallback) -> int {
        int rc = zero_bytes(nrm, sz_nrm, s);          // norms and the solve flag
        auto chk = [&]() { if (rc == TMVX_OK) rc = hip_status(hipGetLastError()); };
        // 1. A32 = fl32(A) or fl32(U^T A V), ||A||_inf, ||b||_inf per column.
        if (rc == TMVX_OK && rbt) {
            hipLaunchKernelGGL(rbt_table_kernel, dim3((unsigned)((2 * N4 + 255) / 256)), dim3(256),
                               0, s, 0x5eed5eed0badcafeull, N4, Wrbt);
            chk();
            if (N != n) {
                hipLaunchKernelGGL(rbt_pad_kernel, dim3(1), dim3(256), 0, s, n, A, lda, d_pad);
                chk();
            }
            hipLaunchKernelGGL(rbt_copy_kernel, dim3((unsigned)((N4 + 255) / 256), RBT_CHUNKS),
                               dim3(256), 0, s, n, N4, A, lda, Wrbt, d_pad, A32, rowpart,
                               RBT_CHUNKS);
            chk();
            hipLaunchKernelGGL(rownorm_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                               n, rowpart, RBT_CHUNKS, d_anorm);
            chk();
        } else if (rc == TMVX_OK) {
            hipLaunchKernelGGL(lowprec_copy_kernel, dim3((unsigned)((n + 255) / 256), NCHUNK),
                               dim3(256), 0, s, n, A, lda, A32, rowpart, NCHUNK);
            chk();
            hipLaunchKernelGGL(rownorm_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                               n, rowpart, NCHUNK, d_anorm);
            chk();
        }
        if (rc == TMVX_OK) {
            hipLaunchKernelGGL(colmax_kernel, gcol, dim3(256), 0, s, n, nrhs, B, ldb, d_bmax);
            chk();
        }
        // 2. FP32 LU (synchronises).  Partial pivoting keeps the composed
        // permutation for the solves; the tournament-pivoted panel of
        // tmvx_tslu.hip is within 3% of partial pivoting on this GPU.
        i64 finfo = 0;
        const i64 nf = rbt ? N : n;
        if (rc == TMVX_OK)
            rc = getrf_impl<float>(nf, A32, nf, piv, rbt ? nullptr : perm, &finfo, s, prec,
                                   rbt ? PIVOT_NONE : PIVOT_PARTIAL);
        (void)hipEventRecord(e1, s);
        info->factor_info = (int)finfo;
        if (rc == TMVX_OK && finfo != 0) {
            if (rbt) { *fallback = true; return TMVX_OK; }
            return TMVX_ERR_SINGULAR;
        }
        // 3. Refinement from x0 = 0 (r0 = b).
        if (rc == TMVX_OK) {
            hipLaunchKernelGGL(zero_x_kernel, gcol, dim3(256), 0, s, n, nrhs, X, ldx);
            chk();
            hipLaunchKernelGGL(residual_reduce_kernel, gcol, dim3(256), 0, s, n, nrhs, B, ldb, P,
                               0, R);
            chk();
        }
        double prev = INFINITY;
        int it = 0;
        bool use_gmres = gmres_first;
        while (rc == TMVX_OK) {
            // max|r|, max|x| per column
            rc = zero_bytes(d_rmax, sizeof(unsigned long long) * 2 * nrhs, s);
            if (rc != TMVX_OK) break;
            hipLaunchKernelGGL(colmax_kernel, gcol, dim3(256), 0, s, n, nrhs, R, n, d_rmax);
            chk();
            hipLaunchKernelGGL(colmax_kernel, gcol, dim3(256), 0, s, n, nrhs, X, ldx, d_xmax);
            chk();
            if (rc != TMVX_OK) break;
            if (hipMemcpyAsync(h.data(), d_anorm, sizeof(unsigned long long),
                               hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipMemcpyAsync(h.data() + 1, d_bmax, sizeof(unsigned long long) * 3 * nrhs,
                               hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess) {
                rc = TMVX_ERR_HIP;
                break;
            }
            auto dv = [](unsigned long long b) { double d; memcpy(&d, &b, 8); return d; };
            const double anorm = dv(h[0]);
            worst = 0.0;
            for (i64 j = 0; j < nrhs; ++j) {
                const double bm = dv(h[1 + j]), rm = dv(h[1 + nrhs + j]),
                             xm = dv(h[1 + 2 * nrhs + j]);
                const double den = anorm * xm + bm;
                const double be = den > 0.0 ? rm / den : (rm > 0.0 ? INFINITY : 0.0);
                if (!(be <= worst)) worst = be;     // NaN propagates as not converged
            }
            if (worst <= tol) { info->converged = 1; break; }
            // RBT: a classical step gaining less than 2x switches to GMRES-IR
            // (when allowed); a step gaining less than 10%, or a non-finite
            // error, hands over to partial pivoting
            if (rbt && !std::isfinite(worst)) {
                *fallback = true;
                break;
            }
            if (rbt && it >= 2 && !use_gmres && may_switch && worst > 0.5 * prev) {
                use_gmres = true;
            } else if (rbt && it >= 2 && worst > 0.9 * prev) {
                *fallback = true;
                break;
            }
            prev = worst;
            if (it >= max_iter) {
                if (rbt) *fallback = true;
                break;
            }
            if (use_gmres) {
                // A d = r by preconditioned GMRES, column by column; x += d
                for (i64 j = 0; j < nrhs && rc == TMVX_OK; ++j)
                    rc = gmres_correction(n, A, lda, A32, perm, rbt ? &rw : nullptr, R + j * n,
                                          X + j * ldx, GTOL, GM, gw, d_flag,
                                          &info->inner_iterations, s);
                if (rc != TMVX_OK) break;
            } else if (rbt) {
                // d = V (LU)^-1 U^T r / s in FP32 (inverted diagonal blocks);
                // x += s d
                hipLaunchKernelGGL(rbt_left_kernel<float>, g4col, dim3(256), 0, s, n, N4, nrhs,
                                   R, n, d_rmax, Wrbt, R32, N);
                chk();
                if (rc == TMVX_OK)
                    rc = getrs_perm<float>(0, N, nrhs, A32, N, nullptr, R32, N, d_flag, s, 1);
                if (rc != TMVX_OK) break;
                hipLaunchKernelGGL(rbt_right_kernel<float>, g4col, dim3(256), 0, s, n, N4, nrhs,
                                   R32, N, d_rmax, Wrbt, X, ldx, 1);
                chk();
            } else {
                // d = (LU)^-1 P r / s in FP32 (diagonal blocks through their
                // inverses, off the solves' chain);  x += s d
                hipLaunchKernelGGL(scale_down_kernel, gcol, dim3(256), 0, s, n, nrhs, R, d_rmax,
                                   R32);
                chk();
                if (rc == TMVX_OK)
                    rc = getrs_perm<float>(0, n, nrhs, A32, n, perm, R32, n, d_flag, s, 1);
                if (rc != TMVX_OK) break;
                hipLaunchKernelGGL(update_kernel, gcol, dim3(256), 0, s, n, nrhs, X, ldx, R32,
                                   d_rmax);
                chk();
            }
            // r = b - A x
            if (rc == TMVX_OK) rc = launch_residual(n, nrhs, A, lda, X, ldx, B, ldb, P, NCHUNK, R, s);
            ++it;
            ++info->iterations;
        }
        return rc;
    };

    int rc = TMVX_OK;
    bool fallback = false;
    if (try_rbt) {
        info->pivoting = TMVX_PIVOT_RBT;
        rc = attempt(true, &fallback);
        if (rc == TMVX_OK && fallback) info->pivoting = TMVX_PIVOT_RBT | TMVX_PIVOT_PARTIAL;
    }
    if (!try_rbt || (rc == TMVX_OK && fallback)) {
        info->converged = 0;
        bool unused = false;
        rc = attempt(false, &unused);
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
    const int refine = flags & (TMVX_REFINE_GMRES | TMVX_REFINE_CLASSIC);
    const int pivot = flags & (TMVX_PIVOT_RBT | TMVX_PIVOT_PARTIAL);
    if ((prec != TMVX_LOWPREC_F32 && prec != TMVX_LOWPREC_F16 && prec != TMVX_LOWPREC_BF16 &&
         prec != TMVX_LOWPREC_F16X2 && prec != TMVX_LOWPREC_BF16X2) ||
        (flags & ~(TMVX_REFINE_GMRES | TMVX_REFINE_CLASSIC | TMVX_PIVOT_RBT |
                   TMVX_PIVOT_PARTIAL)) != 0 ||
        refine == (TMVX_REFINE_GMRES | TMVX_REFINE_CLASSIC) ||
        pivot == (TMVX_PIVOT_RBT | TMVX_PIVOT_PARTIAL))
        return TMVX_ERR_UNSUPPORTED;
    return tmvx::gesv_mixed_f32(lowprec, n, nrhs, A, lda, B, ldb, X, ldx, tol, max_iter, info,
                                (hipStream_t)stream);
}
