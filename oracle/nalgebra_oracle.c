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

/* Batched tall least squares in f32 (QR::new + q_tr_mul + upper solve). */
void nro_s_qr_ls_batched(long count, const int *ms, long n, const int64_t *offA, float *Abase,
                         const int64_t *offb, float *bbase, float *diag, float *x, int *info)
{
    for (long j = 0; j < count; ++j) {
        long m = ms[j];
        info[j] = nro_s_qr_ls(m, n, Abase + offA[j], m, diag + j * n, bbase + offb[j], x + j * n) ? 0 : 1;
    }
}
