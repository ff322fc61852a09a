// QR decomposition by Householder reflections and the least-squares /
// square solve with its factors (QR division).
//
// TMV's QR_Decompose(A, beta, det) (src/TMV_QRDecompose.cpp:603-650) stores
// the reflectors below the diagonal (first element 1, implicit), R on and
// above it, and beta (the reflectors' scale: H = I - beta v v^T).  Each
// reflector follows HouseholderReflect (src/TMV_Householder.cpp:46-180):
// scale by max |x|; y = -|x| when x0 > 0, else +|x|; v = x / (x0 - y) so
// that v(0) = 1; beta = (x0 - y)^2 / (|x|^2 - y x0); R(j, j) = y scale;
// an underflowing column becomes zero (beta = 0), and a column already zero
// below the diagonal needs no reflection (beta = 0).  QR division
// (QR_LDiv, src/TMV_QRDiv.cpp:33-94): y = Q^T b, x = R^-1 y(0:n).
//
// Blocked (32-column panels), LAPACK's compact WY form for the updates:
//   panel: ONE workgroup forms the panel's reflectors column by column
//     (block reductions for the scale, the norm and the dot products with
//     the panel's later columns);
//   trailing update C := (I - V T V^T)^T C = C - V (T^T (V^T C)): T (upper
//     triangular, jb x jb) from G = V^T V by one small kernel, the three
//     products on the MFMA GEMM.
// The same block update applies Q^T to the right-hand sides of the solve.
// Real types only; functional, not tuned (QR is not a measured workload).
#include "tmvx_common.h"

#include <cfloat>

namespace tmvx {

constexpr int QR_NB = 32;        // panel width
constexpr int QR_NTH = 512;      // panel workgroup

template <typename T> struct QrLim;
template <> struct QrLim<double> { static constexpr double rmin = DBL_MIN; };
template <> struct QrLim<float> { static constexpr float rmin = FLT_MIN; };

// Sum (op = 0) or maximum (op = 1) of the workgroup's values v[0..K).
template <typename T, int K>
__device__ __forceinline__ void block_reduce(T (&v)[K], int op, T (*s_red)[K], int nk)
{
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    constexpr int NW = QR_NTH / 64;
    // static indices (a runtime index would put v in scratch memory)
#pragma unroll
    for (int q = 0; q < K; ++q)
        if (q < nk)
            for (int off = 32; off > 0; off >>= 1) {
                const T o = __shfl_xor(v[q], off);
                v[q] = op ? fmax(v[q], o) : v[q] + o;
            }
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < K; ++q)
            if (q < nk) s_red[wave][q] = v[q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < K; ++q)
        if (q < nk) {
            T a = s_red[0][q];
            for (int w = 1; w < NW; ++w) a = op ? fmax(a, s_red[w][q]) : a + s_red[w][q];
            v[q] = a;
        }
    __syncthreads();                      // s_red is reused by the next call
}

// Reflectors for panel columns [j0, j1) of the m x n matrix; the panel's own
// later columns are updated as each reflector is formed.
template <typename T>
__global__ void __launch_bounds__(QR_NTH)
qr_panel_kernel(i64 m, T* __restrict__ A, i64 lda, i64 j0, i64 j1, T* __restrict__ beta)
{
    constexpr int NW = QR_NTH / 64;
    const int t = threadIdx.x;
    __shared__ T s_red[NW][QR_NB];
    __shared__ T s_w[QR_NB];
    for (i64 j = j0; j < j1; ++j) {
        T* x = A + j * lda;                 // column j; x[j] is x0
        // scale = max |x_i| over i >= j
        T mv[1] = {T(0)};
        for (i64 i = j + t; i < m; i += QR_NTH) mv[0] = fmax(mv[0], fabs(x[i]));
        block_reduce<T, 1>(mv, 1, reinterpret_cast<T(*)[1]>(s_red), 1);
        const T scale = mv[0];
        if (!(scale >= QrLim<T>::rmin)) {   // hopeless: the column becomes zero
            for (i64 i = j + t; i < m; i += QR_NTH) x[i] = T(0);
            if (t == 0) beta[j] = T(0);
            __syncthreads();
            continue;
        }
        const T invscale = T(1) / scale;
        T ns[1] = {T(0)};
        for (i64 i = j + 1 + t; i < m; i += QR_NTH) {
            const T a = x[i] * invscale;
            ns[0] += a * a;
        }
        block_reduce<T, 1>(ns, 0, reinterpret_cast<T(*)[1]>(s_red), 1);
        const T normsqx = ns[0];
        const T x0 = x[j];
        if (normsqx == T(0)) {              // H = I
            for (i64 i = j + 1 + t; i < m; i += QR_NTH) x[i] = T(0);
            if (t == 0) beta[j] = T(0);
            __syncthreads();
            continue;
        }
        const T x0s = x0 * invscale;
        const T normsq = normsqx + x0s * x0s;
        const T normx = sqrt(normsq);
        const T y = x0s > T(0) ? -normx : normx;
        const T v0 = x0s - y;
        const T b = (v0 * v0) / (normsq - y * x0s);
        const T sinv = (T(1) / v0) * invscale;
        for (i64 i = j + 1 + t; i < m; i += QR_NTH) x[i] *= sinv;
        __syncthreads();
        if (t == 0) { x[j] = y * scale; beta[j] = b; }
        // apply H = I - b v v^T to the panel's later columns
        const int nc = (int)(j1 - j - 1);
        if (nc > 0) {
            T dots[QR_NB];
            for (int c = 0; c < QR_NB; ++c) dots[c] = T(0);
            for (i64 i = j + 1 + t; i < m; i += QR_NTH) {
                const T vi = x[i];
#pragma unroll
                for (int c = 0; c < QR_NB; ++c)
                    if (c < nc) dots[c] += vi * A[i + (j + 1 + c) * lda];
            }
            block_reduce<T, QR_NB>(dots, 0, s_red, nc);
#pragma unroll
            for (int c = 0; c < QR_NB; ++c)       // dots[] is uniform after the reduction
                if (c < nc && t == c) s_w[c] = b * (A[j + (j + 1 + c) * lda] + dots[c]);   // v(0) = 1
            __syncthreads();
            if (t < nc) A[j + (j + 1 + t) * lda] -= s_w[t];
            for (i64 i = j + 1 + t; i < m; i += QR_NTH) {
                const T vi = x[i];
                for (int c = 0; c < nc; ++c) A[i + (j + 1 + c) * lda] -= s_w[c] * vi;
            }
        }
        __syncthreads();
    }
}

// V (m2 x jb, column-major, ld m2): the panel's reflectors with their unit
// first elements and zeros above.
template <typename T>
__global__ void qr_copy_v_kernel(i64 m2, int jb, const T* __restrict__ A, i64 lda, T* __restrict__ V)
{
    const int c = blockIdx.y;
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < m2; i += (i64)gridDim.x * blockDim.x)
        V[i + c * m2] = i < c ? T(0) : (i == c ? T(1) : A[i + (i64)c * lda]);
}

// T (jb x jb upper, column-major) of the compact WY form I - V T V^T from
// G = V^T V and beta (LAPACK larft, forward columnwise):
//   T(i, i) = beta_i,  T(0:i, i) = -beta_i T(0:i, 0:i) G(0:i, i).
template <typename T>
__global__ void qr_form_t_kernel(int jb, const T* __restrict__ G, const T* __restrict__ beta,
                                 T* __restrict__ Tm)
{
    __shared__ T s_t[QR_NB][QR_NB + 1];
    __shared__ T s_g[QR_NB][QR_NB + 1];
    const int t = threadIdx.x;
    for (int e = t; e < QR_NB * QR_NB; e += blockDim.x) {
        const int r = e % QR_NB, c = e / QR_NB;
        s_g[r][c] = (r < jb && c < jb) ? G[r + c * jb] : T(0);
        s_t[r][c] = T(0);
    }
    __syncthreads();
    for (int i = 0; i < jb; ++i) {
        const T bi = beta[i];
        if (t < i) {
            T acc = T(0);
            for (int q = t; q < i; ++q) acc += s_t[t][q] * s_g[q][i];   // T upper: q >= t
            s_t[t][i] = -bi * acc;
        }
        if (t == 0) s_t[i][i] = bi;
        __syncthreads();
    }
    for (int e = t; e < jb * jb; e += blockDim.x) {
        const int r = e % jb, c = e / jb;
        Tm[r + c * jb] = s_t[r][c];
    }
}

// C (m2 x nc, ld ldc) := (I - V T V^T)^T C for the panel's V (from A at
// column j0, m2 = m - j0 rows) and beta.  ws: m2*jb + 2*jb*jb + jb*nc
// elements.
template <typename T>
int qr_apply_block(i64 m2, int jb, const T* Apanel, i64 lda, const T* beta, T* C, i64 ldc, i64 nc,
                   T* ws, hipStream_t s)
{
    if (nc <= 0 || m2 <= 0 || jb <= 0) return TMVX_OK;
    T* V = ws;
    T* G = V + m2 * jb;
    T* Tm = G + (i64)jb * jb;
    T* W = Tm + (i64)jb * jb;
    const unsigned gx = (unsigned)((m2 + 255) / 256 < 1024 ? (m2 + 255) / 256 : 1024);
    hipLaunchKernelGGL(qr_copy_v_kernel<T>, dim3(gx, (unsigned)jb), dim3(256), 0, s, m2, jb, Apanel,
                       lda, V);
    TMVX_CHECK_LAUNCH();
    // G = V^T V (V^T is V read row-major)
    TMVX_TRY(gemm_launch<T>(1, 0, jb, jb, m2, T(1), V, m2, V, m2, T(0), G, jb, s));
    hipLaunchKernelGGL(qr_form_t_kernel<T>, dim3(1), dim3(64), 0, s, jb, G, beta, Tm);
    TMVX_CHECK_LAUNCH();
    // W = V^T C  (jb x nc)
    TMVX_TRY(gemm_launch<T>(1, 0, jb, nc, m2, T(1), V, m2, C, ldc, T(0), W, jb, s));
    // W := T^T W  (T^T is T read row-major); G's space holds the product
    T* W2 = G;
    if ((i64)jb * nc > (i64)jb * jb) {
        // G is only jb x jb: put the product after W instead
        W2 = W + (i64)jb * nc;
    }
    TMVX_TRY(gemm_launch<T>(1, 0, jb, nc, jb, T(1), Tm, jb, W, jb, T(0), W2, jb, s));
    // C -= V W2
    TMVX_TRY(gemm_launch<T>(0, 0, m2, nc, jb, T(-1), V, m2, W2, jb, T(1), C, ldc, s));
    return TMVX_OK;
}

static i64 qr_ws_elems(i64 m, i64 ncols)
{
    return m * QR_NB + 2 * (i64)QR_NB * QR_NB + 2 * (i64)QR_NB * ncols;
}

template <typename T>
int geqrf_impl(i64 m, i64 n, T* A, i64 lda, T* beta, hipStream_t s)
{
    if (m < 0 || n < 0 || m < n || lda < (m > 0 ? m : 1)) return TMVX_ERR_ARG;
    if (n == 0) return TMVX_OK;
    if (!A || !beta) return TMVX_ERR_ARG;
    T* ws = nullptr;
    if (hipMallocAsync((void**)&ws, sizeof(T) * (size_t)qr_ws_elems(m, n), s) != hipSuccess)
        return TMVX_ERR_ALLOC;
    int rc = TMVX_OK;
    for (i64 j0 = 0; j0 < n && rc == TMVX_OK; j0 += QR_NB) {
        const int jb = (int)(n - j0 < QR_NB ? n - j0 : QR_NB);
        hipLaunchKernelGGL(qr_panel_kernel<T>, dim3(1), dim3(QR_NTH), 0, s, m, A, lda, j0, j0 + jb,
                           beta);
        if ((rc = hip_status(hipGetLastError())) != TMVX_OK) break;
        rc = qr_apply_block<T>(m - j0, jb, A + j0 + j0 * lda, lda, beta + j0,
                               A + j0 + (j0 + jb) * lda, lda, n - j0 - jb, ws, s);
    }
    hipFreeAsync(ws, s);
    return rc;
}

__global__ void qr_diag_check_kernel(i64 n, const double* __restrict__ A, i64 lda, int* __restrict__ bad)
{
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x)
        if (A[i + i * lda] == 0.0) *bad = 1;
}
__global__ void qr_diag_check_kernel_f(i64 n, const float* __restrict__ A, i64 lda, int* __restrict__ bad)
{
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x)
        if (A[i + i * lda] == 0.0f) *bad = 1;
}
__global__ void qr_set_int_kernel(int* p, int v) { *p = v; }

template <typename T>
int geqrs_impl(i64 m, i64 n, i64 nrhs, const T* QR, i64 lda, const T* beta, T* B, i64 ldb,
               hipStream_t s)
{
    if (m < 0 || n < 0 || m < n || nrhs < 0 || lda < (m > 0 ? m : 1) || ldb < (m > 0 ? m : 1))
        return TMVX_ERR_ARG;
    if (n == 0 || nrhs == 0) return TMVX_OK;
    if (!QR || !beta || !B) return TMVX_ERR_ARG;
    T* ws = nullptr;
    const size_t bytes = sizeof(T) * (size_t)qr_ws_elems(m, nrhs) + 64;
    if (hipMallocAsync((void**)&ws, bytes, s) != hipSuccess) return TMVX_ERR_ALLOC;
    int* d_bad = reinterpret_cast<int*>(reinterpret_cast<char*>(ws) + bytes - 64);
    int rc = TMVX_OK;
    // B := Q^T B = H_{n-1} ... H_0 B, a panel at a time
    for (i64 j0 = 0; j0 < n && rc == TMVX_OK; j0 += QR_NB) {
        const int jb = (int)(n - j0 < QR_NB ? n - j0 : QR_NB);
        rc = qr_apply_block<T>(m - j0, jb, QR + j0 + j0 * lda, lda, beta + j0, B + j0, ldb, nrhs,
                               ws, s);
    }
    // a zero diagonal of R: singular (TMV's triangular solve throws)
    if (rc == TMVX_OK) {
        hipLaunchKernelGGL(qr_set_int_kernel, dim3(1), dim3(1), 0, s, d_bad, 0);
        if constexpr (sizeof(T) == 8)
            hipLaunchKernelGGL(qr_diag_check_kernel, dim3((unsigned)((n + 255) / 256 < 256 ? (n + 255) / 256 : 256)),
                               dim3(256), 0, s, n, QR, lda, d_bad);
        else
            hipLaunchKernelGGL(qr_diag_check_kernel_f, dim3((unsigned)((n + 255) / 256 < 256 ? (n + 255) / 256 : 256)),
                               dim3(256), 0, s, n, QR, lda, d_bad);
        rc = hip_status(hipGetLastError());
    }
    int bad = 0;
    if (rc == TMVX_OK &&
        hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess)
        rc = TMVX_ERR_HIP;
    if (rc == TMVX_OK && hipStreamSynchronize(s) != hipSuccess) rc = TMVX_ERR_HIP;
    if (rc == TMVX_OK && bad) rc = TMVX_ERR_SINGULAR;
    // X = R^-1 (Q^T B)(0:n)
    if (rc == TMVX_OK) {
        if constexpr (sizeof(T) == 8)
            rc = tmvx_trsm_f64(1, 0, 0, n, nrhs, QR, lda, B, ldb, s);
        else
            rc = tmvx_trsm_f32(1, 0, 0, n, nrhs, QR, lda, B, ldb, s);
    }
    hipFreeAsync(ws, s);
    return rc;
}

} // namespace tmvx

using namespace tmvx;

extern "C" {

int tmvx_geqrf_f64(int64_t m, int64_t n, double* A, int64_t lda, double* beta, void* stream)
{ return geqrf_impl<double>(m, n, A, lda, beta, (hipStream_t)stream); }
int tmvx_geqrf_f32(int64_t m, int64_t n, float* A, int64_t lda, float* beta, void* stream)
{ return geqrf_impl<float>(m, n, A, lda, beta, (hipStream_t)stream); }

int tmvx_geqrs_f64(int64_t m, int64_t n, int64_t nrhs, const double* QR, int64_t lda,
                   const double* beta, double* B, int64_t ldb, void* stream)
{ return geqrs_impl<double>(m, n, nrhs, QR, lda, beta, B, ldb, (hipStream_t)stream); }
int tmvx_geqrs_f32(int64_t m, int64_t n, int64_t nrhs, const float* QR, int64_t lda,
                   const float* beta, float* B, int64_t ldb, void* stream)
{ return geqrs_impl<float>(m, n, nrhs, QR, lda, beta, B, ldb, (hipStream_t)stream); }

} // extern "C"
