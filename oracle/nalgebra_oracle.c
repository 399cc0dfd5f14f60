/*
 * CPU oracle for the nalgebra dense hot path.
 *
 * TEST INFRASTRUCTURE ONLY: a plain-C restatement of nalgebra 0.35's gemm,
 * LU (+ solve), Cholesky (+ solve) and Householder QR (+ least squares), in
 * the reference's exact operation order (see nalgebra_oracle_body.h for the
 * per-function file:line citations).  It is the checker for the GPU library
 * and the timed CPU baseline; the product path never calls it.
 *
 * The reference is Rust and cannot be built in this image (no cargo/rustc),
 * so there is no oracle/_ref build; the restatement is pinned instead by the
 * reference's own known-answer tests, by LAPACK (numpy/scipy) on the same
 * inputs, by exact rational arithmetic on small cases and by hand-derived
 * rounding cases (tests/test_oracle.py).
 *
 * Build: make -C oracle   (gcc -O3 -ffp-contract=off, see Makefile)
 */
#include <math.h>
#include <stdint.h>
#include <stddef.h>

#define T double
#define F(name) nro_d_##name
#define SQRT sqrt
#define FABS fabs
#include "nalgebra_oracle_body.h"
#undef T
#undef F
#undef SQRT
#undef FABS

#define T float
#define F(name) nro_s_##name
#define SQRT sqrtf
#define FABS fabsf
#include "nalgebra_oracle_body.h"
#undef T
#undef F
#undef SQRT
#undef FABS

/* Adjoint of x = A^{-1} b for SPD A factored as L L^T (no reference
 * counterpart; it is built from Cholesky::solve_mut semantics).
 * Given g = dLoss/dx: lambda = A^{-1} g (one potrs with L), dLoss/db = lambda,
 * dLoss/dA = -1/2 (lambda x^T + x lambda^T) (symmetrised, full n x n). */
void nro_d_potrs_adjoint(long n, const double *L, long lda, const double *x,
                         const double *g, double *dA, long ldda, double *db)
{
    for (long i = 0; i < n; ++i)
        db[i] = g[i];
    nro_d_potrs(n, 1, L, lda, db, n);
    for (long c = 0; c < n; ++c)
        for (long r = 0; r < n; ++r) {
            double p = db[r] * x[c];
            double q = x[r] * db[c];
            dA[r + c * ldda] = -0.5 * (p + q);
        }
}

/* Adjoint of X = A^{-1} B for A = P^{-1} L U (LU::new's packed factors and
 * pivots), upstream gradient G (n x nrhs):
 *   Lambda = A^{-T} G = P^T L^{-T} U^{-T} G, dB = Lambda, dA = -Lambda X^T.
 * Composed from nalgebra's own routines as a user would write it:
 *   U^T y = g : tr_solve_upper_triangular_mut on U (solve.rs:255-276, :411-440:
 *               y_i = (g_i - dotx(U[0..i, i], y[0..i])) / U_ii, i ascending);
 *   L^T z = y : tr_solve_lower_triangular_mut on the unit lower L = lu.l()
 *               (solve.rs:232-253, :379-409: z_i = (y_i - dotx(L[i+1.., i], z[i+1..])) / 1,
 *               i descending);
 *   Lambda    : inv_permute_rows (permutation_sequence.rs:108-116: the swaps in
 *               reverse order).
 * dA[r, c] = -(sum over right-hand sides q, in order, of fl(Lambda[r, q] * X[c, q]))
 * (one product for one right-hand side).  Returns 0 when U has an exact zero
 * on its diagonal (tr_solve returns false), else 1. */
int nro_d_getrs_adjoint(long n, long nrhs, const double *LU, long lda, const int *ipiv, const double *X,
                        long ldx, const double *G, long ldg, double *dA, long ldda, double *dB, long lddb)
{
    for (long q = 0; q < nrhs; ++q) {
        double *y = dB + q * lddb;
        for (long i = 0; i < n; ++i) y[i] = G[i + q * ldg];
        for (long i = 0; i < n; ++i) {
            const double dot = nro_d_dot(i, LU + i * lda, 1, y, 1);
            const double d = LU[i + i * lda];
            if (d == 0.0) return 0;
            y[i] = (y[i] - dot) / d;
        }
        for (long i = n - 1; i >= 0; --i) {
            const double dot = nro_d_dot(n - i - 1, LU + (i + 1) + i * lda, 1, y + i + 1, 1);
            y[i] = (y[i] - dot) / 1.0;
        }
        for (long i = n - 1; i >= 0; --i) {
            const long p = ipiv[i];
            if (p != i) {
                const double t = y[i];
                y[i] = y[p];
                y[p] = t;
            }
        }
    }
    for (long c = 0; c < n; ++c)
        for (long r = 0; r < n; ++r) {
            double s = 0.0;
            for (long q = 0; q < nrhs; ++q) {
                const double pr = dB[r + q * lddb] * X[c + q * ldx];
                s = q == 0 ? pr : s + pr;
            }
            dA[r + c * ldda] = -s;
        }
    return 1;
}

/* Batched Cholesky factor + solve over matrices stored at the given element
 * offsets (column-major, lda = n, B is n x nrhs with ldb = n).  This is the
 * per-matrix loop a nalgebra user would write: Cholesky::new then solve_mut. */
void nro_d_potrf_potrs_batched(long count, const int *ns, const int64_t *offA, double *Abase,
                               const int64_t *offB, double *Bbase, long nrhs, int *info)
{
    for (long j = 0; j < count; ++j) {
        long n = ns[j];
        double *A = Abase + offA[j];
        double *B = Bbase + offB[j];
        info[j] = nro_d_potrf(n, A, n, 0, 0.0);
        if (info[j] == 0)
            nro_d_potrs(n, nrhs, A, n, B, n);
    }
}

/* Batched LU::new + LU::solve_mut over matrices at the given element offsets
 * (column-major, lda = n; B is n x nrhs, ldb = n): the per-matrix loop a
 * nalgebra user would write.  ipiv receives each member's pivots at offP[j];
 * info[j] = 0, or 1 + the largest i with U_ii == 0 (solve_mut returns false;
 * B_j is then left as it was). */
void nro_d_getrf_getrs_batched(long count, const int *ns, const int64_t *offA, double *Abase, const int64_t *offB,
                               double *Bbase, const int64_t *offP, int *ipiv, long nrhs, int *info)
{
    for (long j = 0; j < count; ++j) {
        const long n = ns[j];
        double *A = Abase + offA[j];
        int *piv = ipiv + offP[j];
        nro_d_getrf(n, n, A, n, piv);
        long z = 0;
        for (long i = 0; i < n; ++i)
            if (A[i + i * n] == 0.0) z = i + 1;
        info[j] = (int)z;
        if (z == 0 && nrhs > 0) nro_d_getrs(n, nrhs, A, n, piv, Bbase + offB[j], n);
    }
}

/* Batched tall least squares in f32 (QR::new + q_tr_mul + upper solve). */
void nro_s_qr_ls_batched(long count, const int *ms, long n, const int64_t *offA, float *Abase,
                         const int64_t *offb, float *bbase, float *diag, float *x, int *info)
{
    for (long j = 0; j < count; ++j) {
        long m = ms[j];
        info[j] = nro_s_qr_ls(m, n, Abase + offA[j], m, diag + j * n, bbase + offb[j], x + j * n) ? 0 : 1;
    }
}
