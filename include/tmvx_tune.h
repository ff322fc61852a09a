/*
 * tmvx_tune.h -- tuning hooks of the tmvx library (no counterpart in TMV;
 * TMV picks its GEMM blocking at compile time, src/TMV_MultMM_Block.cpp:
 * 119-164).  Used by the autotuner and the benchmark to time every GEMM
 * tile configuration; tmvx_gemm_* picks the default configuration itself.
 */
#ifndef TMVX_TUNE_H
#define TMVX_TUNE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Number of GEMM tile configurations for an element size (8 or 4 bytes). */
int tmvx_gemm_config_count(int elem_bytes);

/* Text description of configuration `cfg` ("BMxBNxBK waves WMxWN ..."). */
const char* tmvx_gemm_config_name(int elem_bytes, int cfg);

/* tmvx_gemm_f64 / _f32 with an explicit tile configuration. */
int tmvx_gemm_cfg_f64(int cfg, int layout_a, int layout_b, int64_t m, int64_t n,
                      int64_t k, double alpha, const double* A, int64_t lda,
                      const double* B, int64_t ldb, double beta, double* C,
                      int64_t ldc, void* stream);
int tmvx_gemm_cfg_f32(int cfg, int layout_a, int layout_b, int64_t m, int64_t n,
                      int64_t k, float alpha, const float* A, int64_t lda,
                      const float* B, int64_t ldb, float beta, float* C,
                      int64_t ldc, void* stream);

/*
 * PA = LU in FP32 with tournament pivoting (the mixed-precision solver's
 * factorisation; not TMV's pivot sequence).  Arguments as tmvx_getrf_f32.
 * No counterpart in TMV; exposed for tests and experiments.
 */
int tmvx_getrf_f32_tournament(int64_t n, float* A, int64_t lda, int64_t* piv,
                              int64_t* info, void* stream);

/*
 * A = LU without row interchanges (the mixed-precision solver's
 * factorisation after its random butterfly transformation; not TMV's
 * path).  piv receives the identity list; *info = the 1-based first zero or
 * non-finite pivot (the factors are then not usable), else 0.  Arguments as
 * tmvx_getrf_f64 / _f32.  No counterpart in TMV; exposed for tests and
 * experiments.
 */
int tmvx_getrf_nopiv_f64(int64_t n, double* A, int64_t lda, int64_t* piv,
                         int64_t* info, void* stream);
int tmvx_getrf_nopiv_f32(int64_t n, float* A, int64_t lda, int64_t* piv,
                         int64_t* info, void* stream);

/*
 * B := op(A)^-1 B for triangular A as the mixed-precision solver's
 * correction solves compute it: 64 x 64 diagonal blocks applied through
 * their inverses, formed off the solve's dependency chain (rounding differs
 * from TMV's substitution by the explicit inverses, whose error grows with
 * the blocks' condition).  Arguments as tmvx_trsm_f64 / _f32; a zero
 * diagonal is TMVX_ERR_SINGULAR whatever B holds.  No counterpart in TMV;
 * exposed for tests and experiments.
 */
int tmvx_trsm_blockinv_f64(int upper, int trans, int unit, int64_t n, int64_t nrhs,
                           const double* A, int64_t lda, double* B, int64_t ldb,
                           void* stream);
int tmvx_trsm_blockinv_f32(int upper, int trans, int unit, int64_t n, int64_t nrhs,
                           const float* A, int64_t lda, float* B, int64_t ldb,
                           void* stream);

/*
 * Panel timer (diagnostics; no counterpart in TMV).  While started, every
 * panel-kernel launch of the allocating factorisations (tmvx_getrf_*, and
 * through them the solvers) is bracketed by a pair of HIP events recorded
 * on the stream the panel runs on, for the first `cap` launches.  _stop
 * waits for the recorded events and returns the launch count and the summed
 * event-to-event time in milliseconds.  One timer per process: start/stop
 * pairs do not nest; TMVX_ERR_ARG for a second start or a stop without one.
 */
int tmvx_panel_timer_start(int cap);
int tmvx_panel_timer_stop(int64_t* launches, double* total_ms);

#ifdef __cplusplus
}
#endif

#endif /* TMVX_TUNE_H */
