/*
 * nalgebra_rocm.h -- C ABI of the MI355X (gfx950) implementation of
 * nalgebra's dense hot path: gemm, LU with partial pivoting + solve,
 * Cholesky + solve, Householder QR (+ tall least squares), plus mixed
 * precision LU with iterative refinement, batched uneven Cholesky and the
 * adjoint of the Cholesky solve.
 *
 * Conventions
 *  - Matrices are nalgebra DMatrix storage: column-major, element (i, j) at
 *    ptr[i + j*ld] (VecStorage, src/base/vec_storage.rs:26-32,237-239).
 *  - Every matrix, vector, pivot and info pointer is a DEVICE pointer, unless
 *    a parameter's name ends in _host.  Work is enqueued on the handle's
 *    stream and is asynchronous with respect to the host.
 *  - Return value: NROCM_OK (0), a negative value -k when argument k
 *    (1-based, the handle being argument 1) is invalid -- nalgebra panics in
 *    those cases -- or a positive NROCM_ERR_* for runtime failures.
 *    Numerical outcomes (singular U, non-positive-definite pivot) are
 *    reported through device-side info words, as nalgebra returns them as
 *    values (bool / Option) and LAPACK through info.
 *  - mode NROCM_EXACT reproduces nalgebra's arithmetic bit for bit (same
 *    per-element operation order, no fused multiply-add); NROCM_FAST allows
 *    fused multiply-adds and MFMA blocking (results within 1e-12 relative,
 *    LU pivots may differ only at near-ties).
 */
#ifndef NALGEBRA_ROCM_H
#define NALGEBRA_ROCM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NROCM_OK 0
#define NROCM_ERR_HIP 1          /* a HIP runtime call failed */
#define NROCM_ERR_ALLOC 2        /* device allocation failed */
#define NROCM_ERR_UNSUPPORTED 3  /* valid request this build does not implement */
#define NROCM_ERR_NOT_CONVERGED 4 /* iterative refinement missed its tolerance */

#define NROCM_EXACT 0
#define NROCM_FAST 1

#define NROCM_PREC_FP64 0
#define NROCM_PREC_FP32 1
#define NROCM_PREC_BF16 2
#define NROCM_PREC_BF16X3 3  /* BF16 MFMA on three-way split operands (~FP32 accuracy) */
/* OR-ed onto a precision of nrocm_dgesv_ir: refine with flexible GMRES in FP64
 * preconditioned by the low-precision factors instead of classic refinement. */
#define NROCM_PREC_GMRES 32
/* OR-ed onto a precision of nrocm_dgesv_ir: the pivot
 * of every column is searched among the next 128 rows of its 64-column block
 * only (windowed partial pivoting), instead of among all remaining rows. */
#define NROCM_PREC_WINDOW 64

typedef struct nrocm_handle_s *nrocm_handle;
typedef struct nrocm_batch_s *nrocm_batch;

/* ---------------------------------------------------------------- runtime
 * No reference counterpart (nalgebra has no device state); the handle owns
 * the stream binding, workspaces, graph cache and tuning table. */
int nrocm_create(nrocm_handle *handle);
int nrocm_destroy(nrocm_handle handle);
int nrocm_set_stream(nrocm_handle handle, void *hip_stream);
int nrocm_get_stream(nrocm_handle handle, void **hip_stream);
const char *nrocm_status_string(int status);
const char *nrocm_last_error(nrocm_handle handle);
int nrocm_version(void);

/* ---------------------------------------------------------------- gemm (a1)
 * C = alpha*A*B + beta*C with A m x k, B k x n, C m x n and arbitrary
 * element strides (row stride rs*, column stride cs*).  beta == 0 => C is not
 * read.  Replaces matrixmultiply::dgemm / sgemm as called from
 * src/base/blas_uninit.rs:276-313 (same argument order), i.e. Matrix::gemm
 * (src/base/blas.rs:729) and &A * &B (src/base/ops.rs:565-571). */
int nrocm_dgemm(nrocm_handle handle, size_t m, size_t k, size_t n, double alpha,
                const double *a, ptrdiff_t rsa, ptrdiff_t csa,
                const double *b, ptrdiff_t rsb, ptrdiff_t csb, double beta,
                double *c, ptrdiff_t rsc, ptrdiff_t csc);
int nrocm_sgemm(nrocm_handle handle, size_t m, size_t k, size_t n, float alpha,
                const float *a, ptrdiff_t rsa, ptrdiff_t csa,
                const float *b, ptrdiff_t rsb, ptrdiff_t csb, float beta,
                float *c, ptrdiff_t rsc, ptrdiff_t csc);

/* ---------------------------------------------------------------- LU (a4-a7)
 * In-place LU with partial pivoting, packed unit-L \ U exactly as LU::new
 * (src/linalg/lu.rs:93-122, gauss_step_swap :362-389) leaves it.
 * ipiv[i] (int32, 0-based) = pivot row chosen at step i (i when no swap;
 * LU's PermutationSequence holds the pairs (i, ipiv[i]) with ipiv[i] != i,
 * src/linalg/permutation_sequence.rs:84-93).  info[0] = 0, or 1 + the first
 * step whose column was entirely zero (nalgebra skips it, lu.rs:107-110).
 * Shape of nalgebra-lapack's LUScalar::xgetrf (nalgebra-lapack/src/lu.rs:351-380). */
int nrocm_dgetrf(nrocm_handle handle, int m, int n, double *a, int lda, int *ipiv,
                 int *info, int mode);

/* LU::solve_mut (lu.rs:242-260): B (n x nrhs) <- A^{-1} B.  info[0] = 0, or
 * 1 + the index of an exact zero on U's diagonal (solve_mut returns false and
 * B holds garbage); the n >= 1024 path also reports -1 if an internal wait
 * timed out and -2 for a pivot outside [0, n) (not applied).  Any n (n >= 1024: one 64-row block per workgroup, pipelined).
 * Shape of LUScalar::xgetrs. */
int nrocm_dgetrs(nrocm_handle handle, int n, int nrhs, const double *lu, int lda,
                 const int *ipiv, double *b, int ldb, int *info, int mode);

/* LU::solve_mut as nrocm_dgetrs, allocating nothing: no device allocation, no
 * graph, only kernel launches on the handle's stream, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid allocation).  No counterpart in the reference.  Its scratch is
 * a static device array of the library, one per GPU, held until the process
 * ends; n above NROCM_NOALLOC_MAX_N returns NROCM_ERR_UNSUPPORTED.  Calls
 * on the same GPU share that array: calls that overlap in time on one GPU
 * (two streams, two threads) race and must be serialised by the caller.
 * Same arguments, results and info as nrocm_dgetrs (info = -1: an internal
 * wait timed out). */
#define NROCM_NOALLOC_MAX_N 16384
int nrocm_dgetrs_noalloc(nrocm_handle handle, int n, int nrhs, const double *lu, int lda,
                         const int *ipiv, double *b, int ldb, int *info, int mode);

/* ---------------------------------------------------------------- Cholesky (a8-a9)
 * Cholesky::new (src/linalg/cholesky.rs:196-198,221-272): reads the lower
 * triangle, writes L there, leaves the strict upper triangle untouched.
 * info[0] = 0, or j+1 for the first column whose pivot is not > 0 (nalgebra
 * returns None).  Shape of CholeskyScalar::xpotrf (nalgebra-lapack/src/cholesky.rs:181-197). */
int nrocm_dpotrf(nrocm_handle handle, int n, double *a, int lda, int *info, int mode);
/* Cholesky::new_with_substitute (cholesky.rs:216-218, 237-264): a pivot that
 * is not > 0 (NaN included) is replaced by `substitute` before its square
 * root when substitute > 0; substitute <= 0 or NaN substitutes nothing (as
 * Some(s) with no positive root does there), so 0 is Cholesky::new. */
int nrocm_dpotrf_substitute(nrocm_handle handle, int n, double *a, int lda, double substitute,
                            int *info, int mode);
/* Determinant pieces from a factor's diagonal, written to the device array
 * out[3]: out[0] = 1 * a_00 * a_11 * ... in index order (LU::determinant's
 * product, lu.rs:301-314; the caller multiplies by the permutation's sign),
 * out[1] = out[0]^2 (Cholesky::determinant, cholesky.rs:155-164), out[2] =
 * sum of ln(a_ii^2) in index order (Cholesky::ln_determinant,
 * cholesky.rs:166-184; ln is the device's, within 1 ulp of libm's). */
int nrocm_ddiag_det(nrocm_handle handle, int n, const double *a, int lda, double *out);
/* Cholesky::solve_mut (cholesky.rs:122-129): B <- (L L^T)^{-1} B, any n.
 * n >= 1024: the forward sweep runs one 64-row block per workgroup; the
 * EXACT transposed solve keeps nalgebra's dot order, one dependent chain of
 * about n^2/16 additions on one wavefront. */
int nrocm_dpotrs(nrocm_handle handle, int n, int nrhs, const double *l, int lda,
                 double *b, int ldb, int mode);

/* ---------------------------------------------------------------- batched, uneven sizes
 * A batch describes `count` independent systems whose sizes are given on the
 * host (n_host[j], or m_host[j] for least squares).  Creating it sorts the
 * members by size, bins them (padded sizes 8, 16, ..., 64) and gives each
 * bin a default lanes-per-matrix schedule (nrocm_batch_autotune replaces it
 * by a measured one); calls that use it may be captured in a hipGraph (done
 * internally on repeated calls with the same pointers).
 * No reference counterpart: the reference solves a batch as a loop of
 * Cholesky::new + solve_mut. */
int nrocm_batch_create(nrocm_handle handle, int count, const int *n_host, int nrhs,
                       nrocm_batch *batch);
int nrocm_batch_destroy(nrocm_batch batch);
/* Use (1) or stop using (0) hipGraph capture + replay for this batch. */
int nrocm_batch_set_graph(nrocm_batch batch, int enable);
/* Autotuner of the batched Cholesky schedule (no reference counterpart).
 * For each size bin of the batch, times every compiled lanes-per-matrix
 * schedule (8, 16, 32 or 64 lanes, i.e. 8 to 1 matrices per wavefront) of
 * the fused factor+solve kernel in `mode` on up to 4096 generated SPD members
 * of that bin's sizes, and keeps the fastest for later
 * nrocm_dpotrf_potrs_batched calls in that mode.  The caller's matrices are
 * not touched.  Synchronises the handle's stream.  lanes (host, 8 entries or
 * NULL) receives the chosen lanes per padded size 8, 16, ..., 64 (0 where
 * the batch has no member of that size class). */
int nrocm_batch_autotune(nrocm_handle handle, nrocm_batch batch, int mode, int *lanes);
/* Schedule value of the element-tile kernel: one wavefront per matrix with the
 * lower triangle spread over an 8 x 8 grid of lanes (padded sizes 24..64 and
 * at most 4 right-hand sides); the autotuner times it beside 8..64 lanes. */
#define NROCM_SCHED_TILE 100
/* Schedule value of the element-tile kernel with two matrices per wavefront,
 * each over an 8 x 4 grid of 32 lanes (padded sizes 24 and 32, at most 4
 * right-hand sides); bit-identical to NROCM_SCHED_TILE's results, and timed
 * by the autotuner beside the other schedules. */
#define NROCM_SCHED_TILE2 101
/* Current lanes-per-matrix schedule of `mode` in the same format. */
int nrocm_batch_schedule(nrocm_batch batch, int mode, int *lanes);
/* Sets the schedule of `mode` (same format; 0 keeps a size class's current
 * choice).  Each entry must be 8, 16, 32 or 64 and at least the padded size,
 * or NROCM_SCHED_TILE / NROCM_SCHED_TILE2 where that kernel applies; otherwise
 * -3 and nothing changes. */
int nrocm_batch_set_schedule(nrocm_batch batch, int mode, const int *lanes);

/* For each j: Cholesky::new(A_j) then solve_mut(B_j), in one persistent
 * launch.  a[j] -> n_j x n_j (lda n_j), b[j] -> n_j x nrhs (ldb n_j); a and b
 * are device arrays of device pointers.  info[j] as for nrocm_dpotrf; when
 * info[j] != 0, b[j] is left unchanged. */
int nrocm_dpotrf_potrs_batched(nrocm_handle handle, nrocm_batch batch, double *const *a,
                               double *const *b, int *info, int mode);

/* For each j: LU::new(A_j) (src/linalg/lu.rs:93-122), then, when the batch has
 * right-hand sides, LU::solve_mut(B_j) (lu.rs:242-260), one launch per size
 * class.  a[j] -> n_j x n_j (lda n_j), overwritten by the packed L \ U exactly
 * as LU::new leaves it; ipiv (NULL or a device array of per-member device
 * pointers, n_j ints each) receives the 0-based pivot row of every step (i
 * when nothing was swapped); b[j] -> n_j x nrhs, overwritten by the solution.
 * info[j] = 0, or 1 + the largest i with U_ii == 0 (solve_mut returns false;
 * b[j] is then left unchanged).  n_j <= 64.  No reference counterpart as a
 * batch: the reference runs a loop of LU::new + solve_mut. */
int nrocm_dgetrf_getrs_batched(nrocm_handle handle, nrocm_batch batch, double *const *a,
                               int *const *ipiv, double *const *b, int *info, int mode);

/* Adjoint of X_j = A_j^{-1} B_j given LU::new's packed factors lu[j] and
 * pivots ipiv[j] (as nrocm_dgetrf_getrs_batched leaves them), the solution
 * X_j and the upstream gradient dX_j (n_j x nrhs):
 *   Lambda = A^{-T} dX = P^T L^{-T} U^{-T} dX, dB = Lambda, dA = -Lambda X^T
 * (n_j x n_j; products summed over right-hand sides in order).  The solves
 * are nalgebra's tr_solve_upper_triangular / tr_solve_lower_triangular
 * (src/linalg/solve.rs:232-276,379-440, dotx order) and the permutation its
 * inv_permute_rows (permutation_sequence.rs:108-116), bit for bit.  U must be
 * nonsingular.  No reference counterpart (nalgebra has no autodiff). */
int nrocm_dgetrs_backward_batched(nrocm_handle handle, nrocm_batch batch,
                                  const double *const *lu, const int *const *ipiv,
                                  const double *const *x, const double *const *dx,
                                  double *const *da, double *const *db);

/* Adjoint of X_j = A_j^{-1} B_j given the factor L_j, the solution X_j and the
 * upstream gradient dX_j (n_j x nrhs):  Lambda = A^{-1} dX,
 * dB = Lambda, dA = -1/2 (Lambda X^T + X Lambda^T) (symmetrised, full n x n).
 * No reference counterpart (nalgebra has no autodiff); built on
 * Cholesky::solve_mut semantics. */
int nrocm_dpotrs_backward_batched(nrocm_handle handle, nrocm_batch batch,
                                  const double *const *l, const double *const *x,
                                  const double *const *dx, double *const *da,
                                  double *const *db);

/* ---------------------------------------------------------------- QR (a10-a12)
 * QR::new (src/linalg/qr.rs:55-76): column i of A holds the unit Householder
 * axis in rows i..m, diag[i] the signed norm (R_ii = |diag[i]|, qr.rs:81-89),
 * the strict upper triangle holds R.  Exact nalgebra arithmetic. */
int nrocm_dgeqrf(nrocm_handle handle, int m, int n, double *a, int lda, double *diag);
int nrocm_sgeqrf(nrocm_handle handle, int m, int n, float *a, int lda, float *diag);
/* QR::q_tr_mul (qr.rs:157-171): B (m x nrhs) <- Q^T B. */
int nrocm_dqr_qtmul(nrocm_handle handle, int m, int n, const double *qr, int lda,
                    const double *diag, int nrhs, double *b, int ldb);
/* Q B with QR::q's Q (qr.rs:108-129; the role of nalgebra-lapack's
 * QrDecomposition::q_mul_mut, nalgebra-lapack/src/qr/abstraction.rs:83-90):
 * B (m x nrhs) <- Q B, reflectors in reverse order, exact nalgebra arithmetic. */
int nrocm_dqr_qmul(nrocm_handle handle, int m, int n, const double *qr, int lda,
                   const double *diag, int nrhs, double *b, int ldb);
/* R B with R the n x n upper triangle of QR::r (qr.rs:81-89), m >= n (the
 * role of QrDecomposition::r_mul_mut, abstraction.rs:136-146): B (n x nrhs)
 * <- R B in the reference BLAS TRMM order.  Uses the handle's workspace. */
int nrocm_dqr_rmul(nrocm_handle handle, int m, int n, const double *qr, int lda,
                   const double *diag, int nrhs, double *b, int ldb);
/* QR::solve_mut's upper solve (qr.rs:224-256) on the leading n x n block:
 * B (n x nrhs) <- R^{-1} B.  info[0] = 0 or 1 + index of a zero diag. */
int nrocm_dqr_rsolve(nrocm_handle handle, int n, const double *qr, int lda,
                     const double *diag, int nrhs, double *b, int ldb, int *info);

/* Batched tall least squares, min ||A_j x - b_j||, A_j m_j x n (f32, lda m_j,
 * n <= m_j <= 2048, n <= 64), batch created with the row counts m_j.  Blocked
 * compact-WY Householder QR of a workspace copy of [A_j | b_j] (LAPACK
 * sgeqrf/slarft conventions) with the WY products V^T C and C - V (T^T W) on
 * FP32 MFMA, x = R^{-1} (Q^T b)[0..n), then `refine` FP64 residual-refinement
 * steps x += R^{-1} (Q^T (b - A x))[0..n) with b - A x formed in FP64.  A_j and
 * b_j are not modified; x is n x count f64 (column j = system j); info[j] = 0 or
 * 1 + index of a zero R diagonal.  Replaces the tall solve nalgebra-lapack's
 * QrDecomposition::solve provides (nalgebra-lapack/src/qr/abstraction.rs:48-77,
 * qr_util.rs:167-230: ormqr then trtrs); core nalgebra's QR::solve is square-only
 * (src/linalg/qr.rs:214-217). */
int nrocm_sgels_batched(nrocm_handle handle, nrocm_batch batch, int n, float *const *a,
                        float *const *b, double *x, int refine, int *info);

/* ---------------------------------------------------------------- mixed precision (a13)
 * Solve A x = b (n x n, one right-hand side) by LU with partial pivoting whose
 * trailing updates run in `precision`: NROCM_PREC_FP32, NROCM_PREC_BF16 or
 * NROCM_PREC_BF16X3 on MFMA with FP32 accumulation (FP32 working copy, FP32
 * panels, pivot magnitudes compared in FP64), or NROCM_PREC_FP64 (FP64 MFMA),
 * followed by FP64 iterative refinement r = b - A x, A d = r, x += d until the
 * normwise backward error |b - A x|_inf / (|A|_inf |x|_inf + |b|_inf) <= tol
 * or maxit solves.  A and b are not modified.  iters_host / berr_host receive
 * the solves taken and the final backward error (the call synchronises the
 * stream once per refinement step).  Returns NROCM_ERR_NOT_CONVERGED if tol
 * was not reached.  With NROCM_PREC_WINDOW set, each 64-column block is factored
 * with its pivots chosen among the 128 rows starting at the block's diagonal
 * (one workgroup, no full-height pivot search); the rows below that window
 * are eliminated by a triangular solve with the block's U.  When those
 * factors are unusable (a zero pivot inside a window, a non-finite backward
 * error, or a refinement step from the third on that does not halve the
 * backward error, or tol missed after maxit solves) the call factors A again
 * with partial pivoting and refines from x = 0; iters_host then counts the
 * solves of both attempts and nrocm_ir_fallback reports it.  With
 * NROCM_PREC_GMRES set, each refinement cycle is FGMRES(20) on
 * A x = b in FP64, right-preconditioned by the factors (Carson & Higham 2017),
 * and iters_host counts preconditioner solves.  No reference counterpart:
 * replaces LU::new + solve_mut
 * (lu.rs:93,242) for large n. */
int nrocm_dgesv_ir(nrocm_handle handle, int n, const double *a, int lda, const double *b,
                   double *x, int precision, double tol, int maxit, int *iters_host,
                   double *berr_host);
/* 1 in *fell_back_host when the handle's last nrocm_dgesv_ir call replaced
 * windowed factors by partial-pivoting factors, else 0.  No reference
 * counterpart. */
int nrocm_ir_fallback(nrocm_handle handle, int *fell_back_host);

/* ---------------------------------------------------------------- synthetic inputs
 * Host-side seeded generators (xoshiro256** seeded through SplitMix64) so
 * that tests and benchmarks on every machine build identical matrices.
 * No reference counterpart (the reference draws from rand's thread RNG). */
void nrocm_gen_uniform(uint64_t seed, size_t count, double lo, double hi, double *out_host);
void nrocm_gen_normal(uint64_t seed, size_t count, double *out_host);
/* Batched SPD systems: member j (seed + j) draws G (n_j x n_j, N(0,1),
 * column-major) then B (n_j x nrhs, N(0,1)); A_j = G G^T + n_j I.  Members are
 * packed back to back: A at offA[j] (n_j^2 doubles), B at offB[j]. */
void nrocm_gen_spd_batch(uint64_t seed, int count, const int *n_host, int nrhs,
                         double *a_host, const int64_t *offa, double *b_host,
                         const int64_t *offb);

#ifdef __cplusplus
}
#endif

#endif /* NALGEBRA_ROCM_H */
