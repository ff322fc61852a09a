// Internal interfaces of the LU factorisation and triangular solves, shared
// by the plain, mixed-precision and adjoint drivers.
#pragma once
#include "tmvx_common.h"

namespace tmvx {

// In-place PA = LU (pivots 0-based interchanges).  When perm != nullptr it
// also receives the composed row permutation: row i of PA is row perm[i] of
// A.  Synchronises; *info = 1-based first zero pivot or 0.
// lowprec (FP32 only): TMVX_LOWPREC_F16 / _BF16 (/ the split-operand
// _F16X2 / _BF16X2) run the trailing updates on FP16 / BF16 MFMA with
// per-block-row scaling (gemm_lowp_launch).
// pivoting: PIVOT_PARTIAL (TMV's rule), PIVOT_TOURNAMENT (FP32 only,
// tmvx_tslu.hip) or PIVOT_NONE (no interchanges: piv receives the identity
// list and *info the first zero or non-finite pivot) -- the last two for the
// mixed-precision solver, not for TMV parity.
//
// With na != nullptr (no allocation): the workspace is na->ws (at most
// na->bytes; larger n is TMVX_ERR_UNSUPPORTED), the look-ahead's side work
// runs on s itself (same kernels, same order, so the same results), and
// nothing is synchronised or copied to the host: na->d_info (device)
// receives the info, or -1 when a bounded wait expired.  FP32/FP64 partial
// pivoting only.
enum { PIVOT_PARTIAL = 0, PIVOT_TOURNAMENT = 1, PIVOT_NONE = 2 };
struct GetrfNoAlloc {
    void* ws;
    size_t bytes;
    i64* d_info;
};
template <typename T>
int getrf_impl(i64 n, T* A, i64 lda, i64* piv, i64* perm, i64* info, hipStream_t s,
               int lowprec = TMVX_LOWPREC_F32, int pivoting = PIVOT_PARTIAL,
               const GetrfNoAlloc* na = nullptr);
// The same for an m x n matrix: min(m, n) pivots (perm only for m == n).
template <typename T>
int getrf_mn_impl(i64 m, i64 n, T* A, i64 lda, i64* piv, i64* perm, i64* info, hipStream_t s,
                  int lowprec, int pivoting, const GetrfNoAlloc* na);

// Tournament pivoting (tmvx_tslu.hip): panel [j0, j0+jb) factored in place
// on its own columns (selection tree, interchange list piv[j0..j0+jb), the
// interchanges on the panel's columns, L and U); the other columns' share of
// the interchanges is left to the caller.  ws: tslu_ws_ints(n - j0) ints.
i64 tslu_ws_ints(i64 m);
template <typename T>
int tslu_panel(T* A, i64 lda, i64 n, i64 j0, int jb, i64* piv, i64* first_zero, int* ws,
               hipStream_t s);

// C := beta C + alpha A B on FP16 / BF16 MFMA, FP32 operands and result;
// with bmax_word != nullptr, B is scaled by a power of two fitted to
// max |B| (computed into *bmax_word first) and the result unscaled.
int gemm_lowp_launch(int lowprec, i64 m, i64 n, i64 k, float alpha, const float* A, i64 lda,
                     const float* B, i64 ldb, float beta, float* C, i64 ldc, unsigned* bmax_word,
                     hipStream_t s);

// Interchange list -> composed permutation (row i of PA = row perm[i] of A).
int piv_to_perm(i64 n, const i64* piv, i64* perm, hipStream_t s);

// op(A) X = B for triangular A (trsm semantics of tmvx_trsm_*).  When
// in_perm != nullptr the right-hand side is first gathered, B'[i] =
// B[in_perm[i]]; when out_perm != nullptr the solution is scattered,
// X row i -> B row out_perm[i].  d_singular (device int) is or-ed with 1 if
// a zero diagonal meets a non-zero entry.
template <typename T>
int trsm_impl(int upper, int trans, int unit, i64 n, i64 nrhs, const T* A, i64 lda,
              T* B, i64 ldb, const i64* in_perm, const i64* out_perm, int* d_singular,
              hipStream_t s);

// Single-launch-per-4-RHS blocked triangular solve (tmvx_trsv.hip):
// op(A) X = B with B rows gathered through in_perm and X rows scattered
// through out_perm.  X may alias B only when both permutations are null.
// TA is A's storage type, T the arithmetic and right-hand side type
// (instantiated for <double, double>, <float, float>, <double, float>).
// inv = 1: the diagonal blocks are applied through their inverses, off the
// dependency chain (for the mixed solver's correction solves; a zero
// diagonal then sets bit 0 of d_singular whatever the right-hand side).
template <typename T, typename TA = T>
int trsv_multi(int upper, int trans, int unit, i64 n, i64 nrhs, const TA* A, i64 lda,
               const T* B, i64 ldb, const i64* in_perm, T* X, i64 ldx, const i64* out_perm,
               int* d_singular, hipStream_t s, int inv = 0);

// Solve with the factors given the composed permutation (see getrf_impl);
// inv as for trsv_multi (used when nrhs <= 16).
template <typename T>
int getrs_perm(int trans, i64 n, i64 nrhs, const T* LU, i64 lda, const i64* perm, T* B,
               i64 ldb, int* d_singular, hipStream_t s, int inv = 0);

// A side stream with two events for the look-ahead factorisation, taken
// from a per-device pool (tmvx_util.hip) and returned after use; idle sets
// have no pending work.
struct SideSet {
    hipStream_t stream = nullptr;
    hipEvent_t ev_main = nullptr, ev_side = nullptr;
    int device = 0;
};
int side_acquire(SideSet** out);
void side_release(SideSet* ss);

// Fill a device buffer with zeros (a kernel, so it may be graph-captured).
int zero_bytes(void* p, size_t bytes, hipStream_t s);

} // namespace tmvx
