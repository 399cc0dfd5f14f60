// Host-side launchers shared between translation units (each kernel template
// is instantiated in exactly one .hip file).
#pragma once
#include "nrocm_internal.hpp"

// gemm.hip
hipError_t nrocm_gemm_d(hipStream_t st, int m, int n, int k, double alpha, const double *A, long rsa,
                        long csa, const double *B, long rsb, long csb, double beta, double *C, long rsc,
                        long csc);
hipError_t nrocm_gemm_s(hipStream_t st, int m, int n, int k, float alpha, const float *A, long rsa,
                        long csa, const float *B, long rsb, long csb, float beta, float *C, long rsc,
                        long csc);
// C (m x n) -= A (m x k) * B (k x n), column-major, k <= 64 (LU trailing update).
// mode: 1 = FP32 MFMA, 2 = operands rounded to BF16, 3 = three-way split BF16.
// max_groups > 0 caps the workgroups of the launch (0: one per 128x128 tile).
hipError_t nrocm_rank_update_d(hipStream_t st, int m, int n, int k, const double *A, long lda, const double *B,
                               long ldb, double *C, long ldc, int max_groups);
hipError_t nrocm_rank_update_s(hipStream_t st, int m, int n, int k, const float *A, long lda, const float *B,
                               long ldb, float *C, long ldc, int mode, int max_groups);

// lu.hip: trailing-update flavours of the blocked LU.
enum NrocmTrailing { TRAIL_EXACT = 0, TRAIL_F64 = 1, TRAIL_F32 = 2, TRAIL_BF16 = 3, TRAIL_BF16X3 = 4 };
int nrocm_getrf_d(nrocm_handle h, hipStream_t st, int m, int n, double *A, int lda, int *ipiv, int *info,
                  NrocmTrailing trail);
int nrocm_getrf_s(nrocm_handle h, hipStream_t st, int m, int n, float *A, int lda, int *ipiv, int *info,
                  NrocmTrailing trail);
// LU with the pivot search of each 64-column block restricted to the next 128
// rows (square n), for the mixed-precision solver.
// perm (optional) receives the composed row permutation: row i of the factors
// is row perm[i] of A.
int nrocm_getrf_window_d(nrocm_handle h, int n, double *A, int lda, int *ipiv, int *perm, int *info,
                         NrocmTrailing trail);
int nrocm_getrf_window_s(nrocm_handle h, int n, float *A, int lda, int *ipiv, int *perm, int *info,
                         NrocmTrailing trail);

// chol_batched.hip: n <= 64 single-matrix path.
// substitute > 0: Cholesky::new_with_substitute's substitute for pivots that are not > 0.
int nrocm_potrf_small(nrocm_handle h, int n, double *a, int lda, double *b, int ldb, int nrhs, int *info,
                      bool exact, double substitute = 0.0);

// solve_lookback.hip: LU::solve_mut for any n, one 64-row block per workgroup.
// Scratch: nrocm_getrs_scratch_bytes(n) bytes of device memory; allocates nothing.
constexpr int NROCM_GETRS_LOOKBACK_MIN = 1024;  // nrocm_dgetrs uses it from this n on
size_t nrocm_getrs_scratch_bytes(int n);
int nrocm_getrs_lookback(nrocm_handle h, hipStream_t st, int n, int nrhs, const double *lu, int lda, const int *ipiv,
                         double *b, int ldb, int *info, bool exact, void *scratch);
// Cholesky::solve_mut for large n (nrocm_dpotrs from NROCM_GETRS_LOOKBACK_MIN on),
// same scratch size.
int nrocm_potrs_lookback(nrocm_handle h, hipStream_t st, int n, int nrhs, const double *L, int lda, double *b,
                         int ldb, bool exact, void *scratch);
