/*
 * tmvx.h -- C ABI of the MI355X (gfx950) dense GEMM + LU library.
 *
 * This is the drop-in seam for TMV 0.76's dense GEMM and LU path.  TMV's own
 * external seam is BLAS/LAPACK shaped (dgemm / dgetrf / dgetrs / dtrsm,
 * src/TMV_MultMM.cpp:102-127, src/TMV_LUDecompose.cpp:418-442,
 * src/TMV_LUDiv.cpp:58-92, src/TMV_TriDiv_M.cpp:268-310); these entry points
 * take the same arguments, with TMV's semantics where the two differ
 * (0-based pivots, interchange lists as TMV's Permutation stores them).
 *
 * Conventions
 *   - Matrices are column-major with a leading dimension unless a layout
 *     argument says row-major.  All matrix/vector pointers are DEVICE
 *     pointers; `stream` is a hipStream_t (NULL = default stream).
 *   - Pivots are int64 (ptrdiff_t), 0-based: piv[j] is the row interchanged
 *     with row j at step j (TMV LU_Decompose, include/tmv/TMV_LUD.h:78-79).
 *   - Calls are asynchronous on `stream` unless documented otherwise; a
 *     host-side out-parameter (e.g. `info`) means the call synchronises.
 *   - Status codes are returned, never exceptions.  The library holds no
 *     mutable global state other than per-device caches set up once.
 */
#ifndef TMVX_H
#define TMVX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum tmvx_status {
    TMVX_OK = 0,
    TMVX_ERR_ARG = -1,          /* invalid argument (LAPACK info < 0)       */
    TMVX_ERR_ALLOC = -2,        /* device allocation failed                 */
    TMVX_ERR_HIP = -3,          /* HIP runtime / launch error               */
    TMVX_ERR_SINGULAR = -4,     /* zero diagonal met a non-zero RHS entry   */
    TMVX_ERR_NOT_CONVERGED = -5,/* iterative refinement missed its target   */
    TMVX_ERR_TIMEOUT = -6,      /* a bounded device-side wait expired       */
    TMVX_ERR_UNSUPPORTED = -7   /* valid request this build does not handle */
};

enum tmvx_layout { TMVX_COLMAJOR = 0, TMVX_ROWMAJOR = 1 };

/* Human-readable text for a status code. */
const char* tmvx_status_str(int status);

/* Library version string. */
const char* tmvx_version(void);

/*
 * C := alpha * op(A) * op(B) + beta * C, C column-major (m x n).
 * layout_a / layout_b give the storage of A (m x k) and B (k x n):
 * TMVX_COLMAJOR (lda >= m resp. ldb >= k) or TMVX_ROWMAJOR (lda >= k resp.
 * ldb >= n).  TMV's MultMM<add>(alpha,A,B,C) is beta = add ? 1 : 0; when
 * beta == 0, C is not read.  Layout pairs map to TMV's kernels:
 * CC = CCCMultMM, CR = CRCMultMM, RC = RCCMultMM, RR = RCC after copying B
 * (src/TMV_MultMM.cpp:79-91).
 * Replaces: tmv::MultMM<add> (include/tmv/TMV_MatrixArithFunc.h:53-57,
 * src/TMV_MultMM.cpp:404-477) and its dgemm seam (src/TMV_MultMM.cpp:102-127).
 */
int tmvx_gemm_f64(int layout_a, int layout_b, int64_t m, int64_t n, int64_t k,
                  double alpha, const double* A, int64_t lda, const double* B,
                  int64_t ldb, double beta, double* C, int64_t ldc,
                  void* stream);
int tmvx_gemm_f32(int layout_a, int layout_b, int64_t m, int64_t n, int64_t k,
                  float alpha, const float* A, int64_t lda, const float* B,
                  int64_t ldb, float beta, float* C, int64_t ldc,
                  void* stream);

/*
 * In-place PA = LU with partial pivoting of the n x n column-major A.
 * On return A holds unit-lower L below the diagonal and U on and above it;
 * piv (device, n entries) holds 0-based interchanges.  Like TMV, a column
 * whose largest candidate underflows (|x| < DBL_MIN / FLT_MIN) is zeroed
 * below the diagonal with piv[j] = j and the factorisation continues
 * (src/TMV_LUDecompose.cpp:162-169, 286-292, 359-373); *info (host) gets
 * the 1-based index of the first zero pivot, or 0.  Synchronises.
 * Replaces: tmv::LU_Decompose(MatrixView<T>, ptrdiff_t*)
 * (include/tmv/TMV_LUD.h:78-79, src/TMV_LUDecompose.cpp:514-559) and its
 * dgetrf seam (src/TMV_LUDecompose.cpp:418-442).
 */
int tmvx_getrf_f64(int64_t n, double* A, int64_t lda, int64_t* piv,
                   int64_t* info, void* stream);
int tmvx_getrf_f32(int64_t n, float* A, int64_t lda, int64_t* piv,
                   int64_t* info, void* stream);

/*
 * The same for an m x n column-major A (lda >= m): min(m, n) pivot steps,
 * piv has min(m, n) entries; a wide A's columns past m take the row
 * interchanges and U = L^-1 A there, a tall A's rows past n end as L.
 * *info as above (first zero pivot among the min(m, n)).  Synchronises.
 * Replaces: tmv::LU_Decompose on a non-square MatrixView (src/
 * TMV_LUDecompose.cpp:217-245,514-538: RecursiveLUDecompose with
 * R = min(M, N)) and its dgetrf seam (src/TMV_LUDecompose.cpp:418-442).
 */
int tmvx_getrf_mn_f64(int64_t m, int64_t n, double* A, int64_t lda, int64_t* piv,
                      int64_t* info, void* stream);
int tmvx_getrf_mn_f32(int64_t m, int64_t n, float* A, int64_t lda, int64_t* piv,
                      int64_t* info, void* stream);

/*
 * tmvx_getrf_f64 that allocates nothing: no counterpart in TMV; for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The same factorisation
 * (same kernels in the same order: identical pivots and factors) on
 * `stream` alone, asynchronous: *d_info (DEVICE int64) receives the 1-based
 * first zero pivot, 0, or -1 when a bounded device-side wait expired; the
 * call neither synchronises nor copies to the host.  Its scratch (panel
 * exchange slots and flags) is a static 64 KiB array per GPU, cleared by
 * each call: n <= 65536 (larger n returns TMVX_ERR_UNSUPPORTED).  Calls
 * that overlap on one GPU (from several streams or threads) race on that
 * array and must be serialised by the caller.
 */
int tmvx_getrf_f64_noalloc(int64_t n, double* A, int64_t lda, int64_t* piv,
                           int64_t* d_info, void* stream);

/*
 * Solve with the factors of tmvx_getrf_*: trans == 0 solves A X = B
 * (permute B's rows by piv, then unit-lower L, then U: TMV LU_LDivEq,
 * src/TMV_LUDiv.cpp:185-226); trans != 0 solves A^T X = B (TMV LU_RDivEq on
 * B^T, src/TMV_LUDiv.cpp:369-410).  B is n x nrhs column-major, overwritten
 * with X.  A zero U diagonal met by a non-zero entry reports
 * TMVX_ERR_SINGULAR (TMV throws SingularUpperTriMatrix,
 * src/TMV_TriDiv_V.cpp:171-217), and an expired bounded device-side wait
 * TMVX_ERR_TIMEOUT; both need a synchronisation.  check_singular = 0 skips
 * the synchronisation and with it BOTH reports: the call is then fully
 * asynchronous and a singular factor leaves inf/NaN in B.
 * Replaces: tmv::LU_LDivEq / LU_RDivEq (src/TMV_LUDiv.h:29-35) and the
 * dgetrs seam (src/TMV_LUDiv.cpp:58-92).
 */
int tmvx_getrs_f64(int trans, int64_t n, int64_t nrhs, const double* LU,
                   int64_t lda, const int64_t* piv, double* B, int64_t ldb,
                   int check_singular, void* stream);
int tmvx_getrs_f32(int trans, int64_t n, int64_t nrhs, const float* LU,
                   int64_t lda, const int64_t* piv, float* B, int64_t ldb,
                   int check_singular, void* stream);

/*
 * Complex GEMM: C := alpha * op(A) * op(B) + beta * C with interleaved
 * complex storage (z: std::complex<double>, c: std::complex<float>; the
 * pointers address the first real part, leading dimensions count complex
 * elements, every pointer aligned to a whole complex element).  op(X) is X,
 * or its conjugate when conj_x != 0 (TMV's conjugate views); layouts as in
 * tmvx_gemm_f64; C column-major.  beta == 0: C is not read.  Computed as
 * TMV does, from four real products of the real and imaginary parts with
 * the conjugations folded into their signs, here on the real MFMA GEMM.
 * Replaces: tmv::MultMM<add> for complex T (include/tmv/TMV_MatrixArithFunc.h:
 * 53-57) and its complex block split (src/TMV_MultMM_Block.cpp:334-392).
 */
int tmvx_gemm_z(int layout_a, int layout_b, int conj_a, int conj_b, int64_t m, int64_t n,
                int64_t k, double alpha_re, double alpha_im, const double* A, int64_t lda,
                const double* B, int64_t ldb, double beta_re, double beta_im, double* C,
                int64_t ldc, void* stream);
int tmvx_gemm_c(int layout_a, int layout_b, int conj_a, int conj_b, int64_t m, int64_t n,
                int64_t k, float alpha_re, float alpha_im, const float* A, int64_t lda,
                const float* B, int64_t ldb, float beta_re, float beta_im, float* C,
                int64_t ldc, void* stream);

/*
 * Complex LU with partial pivoting, in place, interleaved complex storage
 * (z: std::complex<double>, c: std::complex<float>; lda in complex
 * elements).  The pivot is the candidate of largest modulus, first row on
 * ties; a column whose largest candidate underflows is zeroed from the
 * diagonal down with piv[j] = j, and *info (host) gets the 1-based first
 * zero pivot, or 0.  Synchronises.  The complex path is functional, not
 * tuned: one workgroup per panel, trailing updates on the complex GEMM.
 * Replaces: tmv::LU_Decompose for complex T (src/TMV_LUDecompose.cpp:217-407,
 * 443-462).
 */
int tmvx_getrf_z(int64_t n, double* A, int64_t lda, int64_t* piv, int64_t* info,
                 void* stream);
int tmvx_getrf_c(int64_t n, float* A, int64_t lda, int64_t* piv, int64_t* info,
                 void* stream);

/*
 * Solve with the factors of tmvx_getrf_z / _c: trans == 0 solves A X = B,
 * trans != 0 solves A^T X = B (transpose, not conjugate).  B is n x nrhs
 * column-major (complex elements), overwritten with X.  A zero U diagonal
 * met by a non-zero entry reports TMVX_ERR_SINGULAR.  Synchronises.
 * Replaces: tmv::LU_LDivEq / LU_RDivEq for complex T (src/TMV_LUDiv.cpp:
 * 185-226, 369-410).
 */
int tmvx_getrs_z(int trans, int64_t n, int64_t nrhs, const double* LU, int64_t lda,
                 const int64_t* piv, double* B, int64_t ldb, void* stream);
int tmvx_getrs_c(int trans, int64_t n, int64_t nrhs, const float* LU, int64_t lda,
                 const int64_t* piv, float* B, int64_t ldb, void* stream);

/*
 * From the factors of tmvx_getrf_z: log|det A| (FP64) and the complex sign
 * det(P) * prod U_ii / |U_ii| (host outputs; synchronises), a zero diagonal
 * giving logdet = -inf and sign 0 -- LUDiv::logDet for complex T
 * (src/TMV_LUD.cpp:145-156; DiagMatrix::logDet, src/TMV_DiagMatrix.cpp:
 * 59-90); and X := A^-1 (n x n, not aliasing LU; n <= 65535) as the solve
 * of A X = I, the matrix TMV's LU_Inverse returns (src/TMV_LUInverse.cpp:
 * 235-275).  A zero U diagonal makes tmvx_getri_z report TMVX_ERR_SINGULAR.
 */
int tmvx_logdet_z(int64_t n, const double* LU, int64_t lda, const int64_t* piv,
                  double* logdet, double* sign_re, double* sign_im, void* stream);
int tmvx_getri_z(int64_t n, const double* LU, int64_t lda, const int64_t* piv, double* X,
                 int64_t ldx, void* stream);

/*
 * QR decomposition by Householder reflections, in place: A (m x n, m >= n,
 * column-major) gets R on and above the diagonal and the reflectors below
 * it (first elements 1, implicit); beta (device, n) the reflectors' scales,
 * H_j = I - beta_j v_j v_j^T, A = H_0 H_1 ... H_{n-1} R.  Reflectors as
 * TMV's HouseholderReflect builds them (R(j,j) = -sign(x0) |x|, zero or
 * underflowing columns give beta = 0).  Functional, not tuned.
 * Replaces: tmv::QR_Decompose(MatrixView A, VectorView beta, T& det)
 * (src/TMV_QRDecompose.cpp:603-650; reflector src/TMV_Householder.cpp:46-180).
 */
int tmvx_geqrf_f64(int64_t m, int64_t n, double* A, int64_t lda, double* beta, void* stream);
int tmvx_geqrf_f32(int64_t m, int64_t n, float* A, int64_t lda, float* beta, void* stream);

/*
 * QR division with the factors of tmvx_geqrf_*: B (m x nrhs) := Q^T B, then
 * its first n rows := R^-1 of themselves -- the least-squares solution of
 * A X = B for m > n, the solution for m == n (rows n..m-1 of B keep the
 * residual's Q^T components).  A zero diagonal of R reports
 * TMVX_ERR_SINGULAR.  Synchronises.
 * Replaces: tmv::QR_LDiv (src/TMV_QRDiv.cpp:33-94) behind x = b/m with
 * m.divideUsing(tmv::QR).
 */
int tmvx_geqrs_f64(int64_t m, int64_t n, int64_t nrhs, const double* QR, int64_t lda,
                   const double* beta, double* B, int64_t ldb, void* stream);
int tmvx_geqrs_f32(int64_t m, int64_t n, int64_t nrhs, const float* QR, int64_t lda,
                   const float* beta, float* B, int64_t ldb, void* stream);

/*
 * X := A^-1 from the factors of tmvx_getrf_* (X n x n column-major, not
 * aliasing LU).  The same matrix TMV's LU_Inverse returns (U^-1 L^-1 with
 * the columns reverse-permuted by piv); formed here as the solve of
 * A X = I with all n columns at once.  A zero U diagonal reports
 * TMVX_ERR_SINGULAR.  n <= 65535.
 * Replaces: tmv::LU_Inverse (src/TMV_LUDiv.h:38, src/TMV_LUInverse.cpp:
 * 235-275) behind LUDiv::doMakeInverse (src/TMV_LUD.cpp:158-169), i.e.
 * m.inverse() with LU division.
 */
int tmvx_getri_f64(int64_t n, const double* LU, int64_t lda, const int64_t* piv,
                   double* X, int64_t ldx, void* stream);
int tmvx_getri_f32(int64_t n, const float* LU, int64_t lda, const int64_t* piv,
                   float* X, int64_t ldx, void* stream);

/*
 * log|det A| and sign(det A) from the factors of tmvx_getrf_*, written to
 * host memory (the call synchronises the stream): logdet = sum log|U_ii|
 * (in FP64), sign = det(P) * prod sign(U_ii), det(P) = (-1)^(number of i
 * with piv[i] != i).  A zero diagonal gives logdet = -inf and sign = 0;
 * det A = sign * exp(logdet), or 0 when sign is 0.
 * Replaces: LUDiv::det / LUDiv::logDet (src/TMV_LUD.cpp:133-156) with
 * DiagMatrix::logDet (src/TMV_DiagMatrix.cpp:59-90) and Permutation::det
 * (include/tmv/TMV_Permutation.h:273-278).
 */
int tmvx_logdet_f64(int64_t n, const double* LU, int64_t lda, const int64_t* piv,
                    double* logdet, double* sign, void* stream);
int tmvx_logdet_f32(int64_t n, const float* LU, int64_t lda, const int64_t* piv,
                    double* logdet, double* sign, void* stream);

/*
 * min_i |U_ii| and max_i |U_ii| of the packed factors (for complex
 * entries |re| + |im|, TMV's Abs2), written to host memory (the call
 * synchronises the stream).  n = 0 gives min = +inf, max = 0.
 * Replaces: the diagonal scan of LUDiv::isSingular (src/TMV_LUD.cpp:
 * 195-200: minAbs2Element <= eps * maxAbs2Element).  _z / _c take
 * interleaved complex storage as tmvx_getrf_z / _c write it.
 */
int tmvx_lu_diag_range_f64(int64_t n, const double* LU, int64_t lda, double* dmin,
                           double* dmax, void* stream);
int tmvx_lu_diag_range_f32(int64_t n, const float* LU, int64_t lda, double* dmin,
                           double* dmax, void* stream);
int tmvx_lu_diag_range_z(int64_t n, const double* LU, int64_t lda, double* dmin,
                         double* dmax, void* stream);
int tmvx_lu_diag_range_c(int64_t n, const float* LU, int64_t lda, double* dmin,
                         double* dmax, void* stream);

/*
 * B := op(A)^-1 B with A n x n triangular (upper != 0: upper), unit != 0
 * for an implicit unit diagonal, trans != 0 for A^T.  B is n x nrhs.
 * Entries are divided by the diagonal as TMV does.  A zero diagonal met by
 * a non-zero entry reports TMVX_ERR_SINGULAR (TMV throws
 * SingularUpperTriMatrix / SingularLowerTriMatrix, src/TMV_TriDiv_V.cpp:
 * 171-217).  Synchronises.
 * Replaces: tmv::TriLDivEq(Tri, MatrixView) (include/tmv/TMV_TriDiv.h:72-75,
 * src/TMV_TriDiv_M.cpp:190-266) and the dtrsm seam (src/TMV_TriDiv_M.cpp:
 * 268-310).
 */
int tmvx_trsm_f64(int upper, int trans, int unit, int64_t n, int64_t nrhs,
                  const double* A, int64_t lda, double* B, int64_t ldb,
                  void* stream);
int tmvx_trsm_f32(int upper, int trans, int unit, int64_t n, int64_t nrhs,
                  const float* A, int64_t lda, float* B, int64_t ldb,
                  void* stream);

/*
 * Apply interchanges piv[k1..k2) in order (forward != 0) or in reverse
 * order to the rows of the ncols columns of A.  TMV permuteRows /
 * reversePermuteRows (src/TMV_Matrix.cpp:865-930).
 */
int tmvx_laswp_f64(int forward, int64_t ncols, double* A, int64_t lda,
                   int64_t k1, int64_t k2, const int64_t* piv, void* stream);
int tmvx_laswp_f32(int forward, int64_t ncols, float* A, int64_t lda,
                   int64_t k1, int64_t k2, const int64_t* piv, void* stream);

/*
 * Mixed-precision solve with iterative refinement (no counterpart in TMV,
 * whose solves are all in the matrix's own precision; the seam it extends
 * is x = b/A with divideUsing(LU), src/TMV_LUDiv.cpp:185-226).
 * A (n x n, FP64, not modified) is factorised in low precision on MFMA
 * (after a random butterfly transformation without interchanges, or with
 * partial pivoting: tmvx_pivoting below; on an FP32 copy; lowprec selects the
 * arithmetic of the trailing updates: FP32, or FP16 / BF16 with FP32
 * accumulation and the U block row scaled by a power of two per outer
 * panel), then for each of the nrhs columns
 *   r = b - A x  (FP64, on the GPU),  then the correction d:
 *     classical: solve LU d = P r / s in FP32 with s = max|r|, d *= s;
 *     GMRES-IR:  A d = r by GMRES in FP64 preconditioned by the factors
 *                (see tmvx_refine below);
 *   x += d  (FP64)
 * until the normwise backward error ||r||_inf / (||A||_inf ||x||_inf +
 * ||b||_inf) of every column is <= tol, or max_iter refinement steps.
 * X (n x nrhs, FP64) receives the solution.  Synchronises.  Returns
 * TMVX_ERR_NOT_CONVERGED (with X holding the last iterate and
 * info->backward_error its backward error) when the target is not met.
 * lowprec may carry a tmvx_refine and a tmvx_pivoting flag.  iterations and
 * inner_iterations count both attempts when RBT falls back; factor_ms then
 * runs to the end of the partial-pivoting factorisation.
 */
enum tmvx_lowprec {
    TMVX_LOWPREC_F32 = 0, TMVX_LOWPREC_F16 = 1, TMVX_LOWPREC_BF16 = 2,
    /* split operands on the same 16-bit MFMA: x = hi + 2^-S lo, three
       products per tile (hi hi + hi lo + lo hi), FP32 accumulation; about
       22 (FP16, S = 11) / 16 (BF16, S = 8) significant operand bits */
    TMVX_LOWPREC_F16X2 = 3, TMVX_LOWPREC_BF16X2 = 4
};

/*
 * Refinement method, or-ed into lowprec.  Default: classical refinement
 * after an FP32 or FP16X2 factorisation; GMRES-IR after FP16, BF16 or
 * BF16X2 trailing updates.
 * GMRES-IR (Carson & Higham's three-precision refinement) solves the
 * correction equation A d = r by GMRES in FP64, left-preconditioned by the
 * low-precision factors (restart 50, relative inner tolerance 1e-8, one
 * right-hand side at a time); it converges where kappa * u_low >~ 1 makes
 * classical refinement stall or diverge.
 */
enum tmvx_refine { TMVX_REFINE_GMRES = 16, TMVX_REFINE_CLASSIC = 32 };

/*
 * Factorisation, or-ed into lowprec.  Default (neither flag) is
 * TMVX_PIVOT_PARTIAL.
 *   TMVX_PIVOT_RBT: a random butterfly transformation A' = U^T A V (U, V
 *     recursive butterflies of depth two; Parker 1995, Baboulin et al.,
 *     ACM TOMS 39(2), 2013), then A' = LU without interchanges, so no panel
 *     waits on a grid-wide pivot search; the corrections are
 *     d = V (LU)^-1 U^T r.  A zero or non-finite pivot, or a refinement that
 *     stagnates (a step gaining less than 10%) or reaches max_iter, falls
 *     back to partial pivoting from x0 = 0 (info->pivoting reports it).
 *     Without interchanges the element growth of dense random matrices
 *     rises with n (10^3 - 10^4 at n = 500 - 2000), which FP32 cannot
 *     absorb at large n and kappa: at n = 16384, kappa = 1e4 the attempt
 *     falls back.  Classical refinement after an RBT hands over to GMRES-IR
 *     when a step gains less than 2x (unless TMVX_REFINE_CLASSIC is set).
 *   TMVX_PIVOT_PARTIAL: partial pivoting (TMV's rule, on FP32).
 */
enum tmvx_pivoting { TMVX_PIVOT_PARTIAL = 64, TMVX_PIVOT_RBT = 128 };

typedef struct tmvx_ir_info {
    int iterations;            /* refinement steps taken                  */
    int converged;             /* 1 when every column met tol             */
    int factor_info;           /* first zero pivot of the low-precision LU */
    int inner_iterations;      /* GMRES steps over all refinement steps   */
    double backward_error;     /* worst column, final iterate             */
    double factor_ms;          /* device time of the factorisation        */
    double refine_ms;          /* device time of all refinement steps     */
    int pivoting;              /* TMVX_PIVOT_RBT, TMVX_PIVOT_PARTIAL, or  */
                               /* both: RBT fell back to partial pivoting */
    int reserved;
} tmvx_ir_info;

int tmvx_gesv_mixed_f64(int lowprec, int64_t n, int64_t nrhs, const double* A,
                        int64_t lda, const double* B, int64_t ldb, double* X,
                        int64_t ldx, double tol, int max_iter,
                        tmvx_ir_info* info, void* stream);

/*
 * C := alpha A B + beta C for FP32 column-major operands computed on FP16
 * or BF16 MFMA (lowprec = TMVX_LOWPREC_F16 / _BF16, or the split-operand
 * _F16X2 / _BF16X2) with FP32 accumulation: the trailing-update kernel of
 * the low-precision factorisations.  A and B are rounded to the 16-bit
 * format (one or two terms each); scale_b != 0 first multiplies B by a
 * power of two fitted to max |B| (and divides the product by it), so FP16
 * keeps B's leading bits whatever its magnitude.  No counterpart in TMV.
 */
int tmvx_gemm_f32_lowp(int lowprec, int64_t m, int64_t n, int64_t k, float alpha,
                       const float* A, int64_t lda, const float* B, int64_t ldb, float beta,
                       float* C, int64_t ldc, int scale_b, void* stream);

/*
 * Adjoint of the solve X = A^-1 B with the factors of tmvx_getrf_*: given
 * the cotangent G = dL/dX (n x nrhs),
 *   gradB = A^-T G                 (n x nrhs; may alias G)
 *   gradA = -gradB X^T (+ gradA when accumulate != 0)      (n x n)
 * No counterpart in TMV, which has no derivatives; the transposed solve
 * is TMV's LU_RDivEq (x % A, src/TMV_LUDiv.cpp:369-410) and the rank-nrhs
 * update its MultMM (src/TMV_MultMM.cpp:404-477).  Synchronises: the
 * transposed solve reports a singular factor (TMVX_ERR_SINGULAR) or an
 * expired bounded wait (TMVX_ERR_TIMEOUT) as tmvx_getrs_* with
 * check_singular != 0 does, before gradA is formed.
 */
int tmvx_gesv_adjoint_f64(int64_t n, int64_t nrhs, const double* LU, int64_t lda,
                          const int64_t* piv, const double* X, int64_t ldx, const double* G,
                          int64_t ldg, double* gradB, int64_t ldgb, double* gradA, int64_t ldga,
                          int accumulate, void* stream);
int tmvx_gesv_adjoint_f32(int64_t n, int64_t nrhs, const float* LU, int64_t lda,
                          const int64_t* piv, const float* X, int64_t ldx, const float* G,
                          int64_t ldg, float* gradB, int64_t ldgb, float* gradA, int64_t ldga,
                          int accumulate, void* stream);

/*
 * Synthetic inputs: out[i] = lo + (hi-lo) * u(seed, offset + i), u in [0,1)
 * from the counter-based SplitMix64 stream documented in tmv_amd/rng.py
 * (bit-identical to it).  normal: Box-Muller over pairs of that stream.
 * Not part of TMV; the benchmark's data generator.
 */
int tmvx_fill_uniform_f64(uint64_t seed, uint64_t offset, int64_t count,
                          double lo, double hi, double* out, void* stream);
int tmvx_fill_normal_f64(uint64_t seed, uint64_t offset, int64_t count,
                         double* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* TMVX_H */
