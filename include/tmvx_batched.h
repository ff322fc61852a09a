/*
 * tmvx_batched.h -- C ABI for many small independent systems on the MI355X.
 *
 * TMV has no batched entry points: a TMV caller factors and solves each
 * system with LU_Decompose + LU_LDivEq (src/TMV_LUDecompose.cpp:514-559,
 * src/TMV_LUDiv.cpp:185-226) or x = b/m with divideUsing(LU), and
 * multiplies with MultMM (src/TMV_MultMM.cpp:404-477), in a loop.  These
 * entry points do the whole loop in one launch with the same per-system
 * semantics: pivots are TMV's interchange list (0-based, first maximum in
 * the current row order, TMV's underflow rule), factors in place.
 *
 * Batch layout: arrays of device pointers (one per system) to column-major
 * matrices stored with leading dimension n (A: n x n, B: n x nrhs), and
 * device arrays of the per-system sizes.
 */
#ifndef TMVX_BATCHED_H
#define TMVX_BATCHED_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { TMVX_BATCH_F64 = 0, TMVX_BATCH_F32 = 1 };
enum { TMVX_BATCH_NMAX = 128, TMVX_BATCH_RHSMAX = 16 };

typedef struct tmvx_batch_plan_s* tmvx_batch_plan;

/*
 * A plan for a batch of `count` systems of orders n_host[i] (1..128) with
 * nrhs_host[i] (0..16; NULL = 0: factor only) right-hand sides, in FP64 or
 * FP32.  The plan orders the work list (largest systems first) and picks
 * the launch schedule: which systems run one per wave and which one per
 * 256-thread workgroup, and how many workgroups per CU.  tune != 0 times
 * every candidate schedule on a synthetic batch of the same sizes and
 * keeps the fastest; tune == 0 keeps the default (waves for n <= 32).
 * Allocates device memory and synchronises `stream`; not capturable.
 */
int tmvx_batch_plan_create(int dtype, int64_t count, const int32_t* n_host,
                           const int32_t* nrhs_host, int tune, tmvx_batch_plan* plan,
                           void* stream);
int tmvx_batch_plan_destroy(tmvx_batch_plan plan);
/* Text naming the plan's schedule (valid until the plan changes). */
const char* tmvx_batch_plan_describe(tmvx_batch_plan plan);
/* Force a schedule: index < tmvx_batch_schedule_count(), wave split,
 * workgroups per CU.  For tuning experiments and tests. */
int tmvx_batch_plan_set(tmvx_batch_plan plan, int schedule, int split, int wgs_per_cu,
                        void* stream);
int tmvx_batch_schedule_count(void);

/*
 * Factor every A[i] in place (P A = L U) and, when nrhs[i] > 0, overwrite
 * B[i] with the solution of A X = B.  piv[i] receives n 0-based
 * interchanges; info[i] (device) the 1-based index of the first zero pivot
 * or 0.  For a system with info[i] > 0 the factorisation is complete (TMV's
 * zero-column rule) but B[i] is left unspecified, as LAPACK gesv leaves it.
 * One kernel launch on `stream`: no allocation, no host synchronisation,
 * no memset -- safe to capture in a hipGraph and replay.
 * Per system: TMV LU_Decompose (src/TMV_LUDecompose.cpp:514-559) followed
 * by LU_LDivEq (src/TMV_LUDiv.cpp:185-226).
 */
int tmvx_gesv_batched_f64(tmvx_batch_plan plan, double* const* A, int64_t* const* piv,
                          double* const* B, int32_t* info, void* stream);
int tmvx_gesv_batched_f32(tmvx_batch_plan plan, float* const* A, int64_t* const* piv,
                          float* const* B, int32_t* info, void* stream);

/*
 * Solve with stored factors of tmvx_gesv_batched_* / tmvx_getrf_*:
 * trans == 0: A X = B (TMV LU_LDivEq, src/TMV_LUDiv.cpp:185-226);
 * trans != 0: A^T X = B (TMV LU_RDivEq on B^T, src/TMV_LUDiv.cpp:369-410).
 * n[i] <= 256 (systems with larger n are skipped); n, nrhs device arrays.
 * One launch, capturable.
 */
int tmvx_getrs_batched_f64(int trans, int64_t count, const int32_t* n, const int32_t* nrhs,
                           const double* const* LU, const int64_t* const* piv,
                           double* const* B, void* stream);
int tmvx_getrs_batched_f32(int trans, int64_t count, const int32_t* n, const int32_t* nrhs,
                           const float* const* LU, const int64_t* const* piv,
                           float* const* B, void* stream);

/*
 * C[i] = alpha A[i] B[i] + beta C[i], A[i] m x k, B[i] k x n, C[i] m x n,
 * column-major with leading dimensions m, k, m (m, n, k device arrays).
 * beta == 0: C is not read.  Per system TMV MultMM (src/TMV_MultMM.cpp:
 * 404-477).  One launch on MFMA, capturable.
 */
int tmvx_gemm_batched_f64(int64_t count, const int32_t* m, const int32_t* n, const int32_t* k,
                          double alpha, const double* const* A, const double* const* B,
                          double beta, double* const* C, void* stream);
int tmvx_gemm_batched_f32(int64_t count, const int32_t* m, const int32_t* n, const int32_t* k,
                          float alpha, const float* const* A, const float* const* B, float beta,
                          float* const* C, void* stream);

/*
 * Factor a batch that may hold systems above TMVX_BATCH_NMAX: n_host (host)
 * gives the orders; systems with n <= 128 run in one batched launch, those
 * with 128 < n <= 1024 in one launch of a workgroup per system, larger ones
 * through the blocked LU (tmvx_getrf_*), one after another on `stream`.
 * Pivots and info as tmvx_gesv_batched_*.  Allocates and synchronises; not
 * capturable.
 * Per system TMV LU_Decompose (src/TMV_LUDecompose.cpp:514-559).
 */
int tmvx_getrf_batched_f64(int64_t count, const int32_t* n_host, double* const* A,
                           int64_t* const* piv, int32_t* info, void* stream);
int tmvx_getrf_batched_f32(int64_t count, const int32_t* n_host, float* const* A,
                           int64_t* const* piv, int32_t* info, void* stream);

/*
 * Adjoint of X_i = A_i^-1 B_i for a batch (n_i <= 256), given stored
 * factors (LU_i, piv_i), the solutions X_i and the cotangents G_i = dL/dX_i
 * (all n_i x nrhs_i, leading dimension n_i):
 *   gradB_i = A_i^-T G_i,   gradA_i = -gradB_i X_i^T (+ gradA_i if accumulate)
 * One wave per system; the rank-nrhs update is fused into the kernel that
 * solves for gradB_i, so gradA_i is written once.  n, nrhs are device
 * arrays.  One launch, capturable.  No counterpart in TMV (no derivatives);
 * per system the transposed solve is LU_RDivEq (src/TMV_LUDiv.cpp:369-410).
 */
int tmvx_gesv_adjoint_batched_f64(int64_t count, const int32_t* n, const int32_t* nrhs,
                                  const double* const* LU, const int64_t* const* piv,
                                  const double* const* X, const double* const* G,
                                  double* const* gradB, double* const* gradA, int accumulate,
                                  void* stream);
int tmvx_gesv_adjoint_batched_f32(int64_t count, const int32_t* n, const int32_t* nrhs,
                                  const float* const* LU, const int64_t* const* piv,
                                  const float* const* X, const float* const* G,
                                  float* const* gradB, float* const* gradA, int accumulate,
                                  void* stream);
/*
 * The same adjoint without stored factors: A_i are the original matrices
 * (not modified); their factors are recomputed into a workspace, for
 * callers that drop the factors after the forward solve to save memory.
 * n_host, nrhs_host are host arrays.  Allocates and synchronises.
 */
int tmvx_gesv_adjoint_batched_recompute_f64(int64_t count, const int32_t* n_host,
                                            const int32_t* nrhs_host, const double* const* A,
                                            const double* const* X, const double* const* G,
                                            double* const* gradB, double* const* gradA,
                                            int accumulate, void* stream);
int tmvx_gesv_adjoint_batched_recompute_f32(int64_t count, const int32_t* n_host,
                                            const int32_t* nrhs_host, const float* const* A,
                                            const float* const* X, const float* const* G,
                                            float* const* gradB, float* const* gradA,
                                            int accumulate, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* TMVX_BATCHED_H */
