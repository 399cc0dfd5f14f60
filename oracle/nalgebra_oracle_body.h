/*
 * Type-generic body of the CPU oracle.  Included twice by nalgebra_oracle.c,
 * once with T=double (suffix d) and once with T=float (suffix s), so that the
 * f64 and f32 restatements cannot drift apart.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of nalgebra 0.35's
 * dense hot path, used to check the GPU library.  Nothing in the product path
 * may call it.
 *
 * Every loop follows the reference's operation order exactly; every "fl(.)"
 * in the comments is one IEEE rounding.  The file must be compiled with
 * -ffp-contract=off so that no multiply-add is fused (Rust never contracts).
 *
 * Required macros: T (scalar type), F(name) (suffixing), SQRT, FABS.
 */

/* dotx: src/base/blas.rs:93-167.  Eight accumulators; accumulator t receives
 * the products of index i with i mod 8 == t while at least 8 elements remain;
 * then res = 0, res += acc0+acc4, res += acc1+acc5, res += acc2+acc6,
 * res += acc3+acc7, then the tail is added one product at a time.
 * Dyn-sized views never take the U2/U3/U4 shortcuts of blas.rs:48-86. */
T F(dot)(long len, const T *a, long inca, const T *b, long incb)
{
    T acc0 = 0, acc1 = 0, acc2 = 0, acc3 = 0, acc4 = 0, acc5 = 0, acc6 = 0, acc7 = 0;
    long i = 0;
    while (len - i >= 8) {
        acc0 += a[(i + 0) * inca] * b[(i + 0) * incb];
        acc1 += a[(i + 1) * inca] * b[(i + 1) * incb];
        acc2 += a[(i + 2) * inca] * b[(i + 2) * incb];
        acc3 += a[(i + 3) * inca] * b[(i + 3) * incb];
        acc4 += a[(i + 4) * inca] * b[(i + 4) * incb];
        acc5 += a[(i + 5) * inca] * b[(i + 5) * incb];
        acc6 += a[(i + 6) * inca] * b[(i + 6) * incb];
        acc7 += a[(i + 7) * inca] * b[(i + 7) * incb];
        i += 8;
    }
    T res = 0;
    res += acc0 + acc4;
    res += acc1 + acc5;
    res += acc2 + acc6;
    res += acc3 + acc7;
    for (long k = i; k < len; ++k)
        res += a[k * inca] * b[k * incb];
    return res;
}

/* axcpy: src/base/blas_uninit.rs:46-52 (b != 0) and :69-75 (b == 0).
 * y_r = fl(fl(fl(a*x_r)*c) + fl(b*y_r)); with b == 0, y is not read. */
void F(axcpy)(long len, T *y, long incy, T a, const T *x, long incx, T c, T b)
{
    if (b != 0) {
        for (long i = 0; i < len; ++i) {
            T p = a * x[i * incx];
            p = p * c;
            T q = b * y[i * incy];
            y[i * incy] = p + q;
        }
    } else {
        for (long i = 0; i < len; ++i) {
            T p = a * x[i * incx];
            y[i * incy] = p * c;
        }
    }
}

/* icamax: src/base/min_max.rs:221-241.  First index of the strict maximum of
 * |x_i| (norm1 of a real is its absolute value); ties keep the lowest index. */
long F(icamax)(long len, const T *x, long inc)
{
    T best = FABS(x[0]);
    long idx = 0;
    for (long i = 1; i < len; ++i) {
        T v = FABS(x[i * inc]);
        if (v > best) {
            best = v;
            idx = i;
        }
    }
    return idx;
}

/* gemm: src/base/blas_uninit.rs:318-331 -> gemv_uninit :127-177.
 * Column j1 of C is y = alpha*A*B[:,j1] + beta*y, evaluated as one axcpy per
 * column of A: y = fl(fl(alpha*A[:,0])*B[0,j1]) + fl(beta*y), then
 * y = fl(fl(alpha*A[:,j])*B[j,j1]) + y.  For Dyn sizes > 5 nalgebra instead
 * calls the third-party matrixmultiply crate, whose summation order is not
 * available here; parity against that path is tolerance-based. */
void F(gemm)(long m, long k, long n, T alpha,
             const T *A, long rsa, long csa,
             const T *B, long rsb, long csb,
             T beta, T *C, long rsc, long csc)
{
    for (long j1 = 0; j1 < n; ++j1) {
        T *y = C + j1 * csc;
        if (k == 0) {
            for (long r = 0; r < m; ++r)
                y[r * rsc] = (beta != 0) ? y[r * rsc] * beta : (T)0;
            continue;
        }
        F(axcpy)(m, y, rsc, alpha, A, rsa, B[j1 * csb], beta);
        for (long j = 1; j < k; ++j)
            F(axcpy)(m, y, rsc, alpha, A + j * csa, rsa, B[j * rsb + j1 * csb], (T)1);
    }
}

/* LU with partial pivoting: src/linalg/lu.rs:93-122 (LU::new),
 * :337-357 (gauss_step), :362-389 (gauss_step_swap).
 * ipiv[i] receives the pivot row of step i (0-based; i when no swap happened,
 * including skipped steps).  nalgebra never fails; the return value is the
 * LAPACK-style info: 1 + the first step whose pivot column was all zeros,
 * or 0. */
int F(getrf)(long m, long n, T *A, long lda, int *ipiv)
{
    long mn = m < n ? m : n;
    int info = 0;
    for (long i = 0; i < mn; ++i) {
        long piv = i + F(icamax)(m - i, A + i + i * lda, 1);
        T diag = A[piv + i * lda];
        ipiv[i] = (int)i;
        if (diag == 0) {            /* lu.rs:107-110: skip the step */
            if (!info)
                info = (int)(i + 1);
            continue;
        }
        T *coeffs = A + i + i * lda;       /* column i, rows i.. */
        if (piv != i) {
            ipiv[i] = (int)piv;
            /* lu.rs:114: swap rows i and piv in columns 0..i */
            for (long c = 0; c < i; ++c) {
                T t = A[i + c * lda];
                A[i + c * lda] = A[piv + c * lda];
                A[piv + c * lda] = t;
            }
            /* gauss_step_swap: lu.rs:377 coeffs.swap((0,0),(piv,0)) */
            T t = coeffs[0];
            coeffs[0] = coeffs[piv - i];
            coeffs[piv - i] = t;
        }
        T inv_diag = (T)1 / diag;                   /* lu.rs:346 / :374 */
        for (long r = 1; r < m - i; ++r)            /* coeffs *= inv_diag */
            coeffs[r] = coeffs[r] * inv_diag;
        for (long c = i + 1; c < n; ++c) {
            T *col = A + c * lda;
            if (piv != i) {                          /* lu.rs:385 */
                T t = col[i];
                col[i] = col[piv];
                col[piv] = t;
            }
            /* down.column_mut(k).axpy(-pivot_row[k], &coeffs, 1) */
            F(axcpy)(m - i - 1, col + i + 1, 1, -col[i], coeffs + 1, 1, (T)1, (T)1);
        }
    }
    return info;
}

/* PermutationSequence::permute_rows (permutation_sequence.rs:97-104) applied
 * to B, with the swaps recorded by getrf. */
void F(laswp)(long n, long nrhs, T *B, long ldb, const int *ipiv)
{
    for (long i = 0; i < n; ++i) {
        long p = ipiv[i];
        if (p != i)
            for (long c = 0; c < nrhs; ++c) {
                T t = B[i + c * ldb];
                B[i + c * ldb] = B[p + c * ldb];
                B[p + c * ldb] = t;
            }
    }
}

/* LU::solve_mut: lu.rs:242-260.  Row swaps, then
 * solve_lower_triangular_with_diag_mut(b, 1) (solve.rs:106-133), then
 * solve_upper_triangular_mut (solve.rs:137-182).  Returns 1 on success and
 * 0 when U has an exact zero on its diagonal (b then holds garbage). */
int F(getrs)(long n, long nrhs, const T *LU, long lda, const int *ipiv, T *B, long ldb)
{
    F(laswp)(n, nrhs, B, ldb, ipiv);
    for (long c = 0; c < nrhs; ++c) {
        T *b = B + c * ldb;
        for (long i = 0; i + 1 < n; ++i) {
            T coeff = b[i] / (T)1;
            F(axcpy)(n - i - 1, b + i + 1, 1, -coeff, LU + (i + 1) + i * lda, 1, (T)1, (T)1);
        }
    }
    for (long c = 0; c < nrhs; ++c) {
        T *b = B + c * ldb;
        for (long i = n - 1; i >= 0; --i) {
            T diag = LU[i + i * lda];
            if (diag == 0)
                return 0;
            T coeff = b[i] / diag;
            b[i] = coeff;
            F(axcpy)(i, b, 1, -coeff, LU + i * lda, 1, (T)1, (T)1);
        }
    }
    return 1;
}

/* Cholesky::new_internal: src/linalg/cholesky.rs:221-272.  Left-looking;
 * reads the lower triangle only and leaves the strict upper triangle as it
 * was.  Returns 0, or j+1 for the first column whose pivot is not > 0 (then
 * nalgebra returns None).  has_sub/sub give new_with_substitute's substitute. */
int F(potrf)(long n, T *A, long lda, int has_sub, T sub)
{
    for (long j = 0; j < n; ++j) {
        for (long k = 0; k < j; ++k) {
            T factor = -A[j + k * lda];
            F(axcpy)(n - j, A + j + j * lda, 1, factor, A + j + k * lda, 1, (T)1, (T)1);
        }
        T diag = A[j + j * lda];
        T denom;
        /* sqrt_denom: None if re <= 0, else try_sqrt (None unless >= 0) */
        if (diag > 0)
            denom = SQRT(diag);
        else if (has_sub && sub > 0)
            denom = SQRT(sub);
        else
            return (int)(j + 1);
        A[j + j * lda] = denom;
        for (long r = j + 1; r < n; ++r)        /* col /= denom */
            A[r + j * lda] = A[r + j * lda] / denom;
    }
    return 0;
}

/* Cholesky::solve_mut: cholesky.rs:122-129 ->
 * solve_lower_triangular_unchecked_mut (solve.rs:488-519) then
 * ad_solve_lower_triangular_unchecked_mut (solve.rs:697-755, dotc). */
void F(potrs)(long n, long nrhs, const T *L, long lda, T *B, long ldb)
{
    for (long c = 0; c < nrhs; ++c) {
        T *b = B + c * ldb;
        for (long i = 0; i < n; ++i) {
            T coeff = b[i] / L[i + i * lda];
            b[i] = coeff;
            F(axcpy)(n - i - 1, b + i + 1, 1, -coeff, L + (i + 1) + i * lda, 1, (T)1, (T)1);
        }
    }
    for (long c = 0; c < nrhs; ++c) {
        T *b = B + c * ldb;
        for (long i = n - 1; i >= 0; --i) {
            T d = F(dot)(n - i - 1, L + (i + 1) + i * lda, 1, b + i + 1, 1);
            b[i] = (b[i] - d) / L[i + i * lda];
        }
    }
}

/* simba's to_exp for reals: (|x|, +1) when x >= 0 (also for -0), else
 * (-x, -1).  The sign chosen at +-0 and NaN is not fixed by any reference
 * test; it only matters for zero columns. */
static inline void F(to_exp)(T x, T *mod, T *sgn)
{
    if (x >= 0) { *mod = x; *sgn = (T)1; }
    else        { *mod = -x; *sgn = (T)-1; }
}

/* f64::signum / f32::signum (Rust std): 1 for +0 and positives, -1 for -0 and
 * negatives, NaN for NaN. */
static inline T F(signum)(T x)
{
    if (x != x) return x;
    return signbit(x) ? (T)-1 : (T)1;
}

/* Reflection::reflect_with_sign (src/geometry/reflection.rs:70-83) of the
 * ncols columns of C (len rows each) through the unit axis x (bias 0). */
static void F(reflect_with_sign)(long len, const T *x, long ncols, T *C, long ldc, T sign)
{
    for (long c = 0; c < ncols; ++c) {
        T *col = C + c * ldc;
        T m_two = sign * (T)-2;
        T factor = (F(dot)(len, x, 1, col, 1) - (T)0) * m_two;
        F(axcpy)(len, col, 1, factor, x, 1, (T)1, sign);
    }
}

/* QR::new: src/linalg/qr.rs:55-76 -> householder::clear_column_unchecked
 * (householder.rs:61-85) -> reflection_axis_mut (householder.rs:19-53).
 * Column i keeps the unit axis in rows i..m; diag[i] is the signed norm
 * returned by reflection_axis_mut (R_ii = |diag[i]|, qr.rs:81-89). */
void F(geqrf)(long m, long n, T *A, long lda, T *diag)
{
    long mn = m < n ? m : n;
    for (long i = 0; i < mn; ++i) {
        T *x = A + i + i * lda;
        long len = m - i;
        T sq = (T)0 + F(dot)(len, x, 1, x, 1);       /* norm_squared, norm.rs:238-250 */
        T nrm = SQRT(sq);
        T mod, sgn;
        F(to_exp)(x[0], &mod, &sgn);
        T signed_norm = sgn * nrm;
        T factor = (sq + mod * nrm) * (T)2;
        x[0] += signed_norm;
        T ret;
        int not_zero;
        if (factor != 0) {
            T s = SQRT(factor);
            for (long r = 0; r < len; ++r)           /* unscale_mut, matrix.rs:1517 */
                x[r] = x[r] / s;
            T nn = SQRT((T)0 + F(dot)(len, x, 1, x, 1));   /* normalize_mut, norm.rs:505-514 */
            for (long r = 0; r < len; ++r)
                x[r] = x[r] / nn;
            ret = -signed_norm;
            not_zero = 1;
        } else {
            ret = signed_norm;
            not_zero = 0;
        }
        if (not_zero)
            F(reflect_with_sign)(len, x, n - i - 1, A + i + (i + 1) * lda, lda, F(signum)(ret));
        diag[i] = ret;
    }
}

/* QR::q_tr_mul: qr.rs:157-171.  B is m x nrhs. */
void F(qr_qtmul)(long m, long n, const T *QR, long lda, const T *diag, long nrhs, T *B, long ldb)
{
    long mn = m < n ? m : n;
    for (long i = 0; i < mn; ++i)
        F(reflect_with_sign)(m - i, QR + i + i * lda, nrhs, B + i, ldb, F(signum)(diag[i]));
}

/* Q B with QR::q's Q (qr.rs:108-129: the reflectors in reverse order, each
 * with the signum of its diag entry) applied to B (m x nrhs) instead of the
 * identity. */
void F(qr_qmul)(long m, long n, const T *QR, long lda, const T *diag, long nrhs, T *B, long ldb)
{
    long mn = m < n ? m : n;
    for (long i = mn - 1; i >= 0; --i)
        F(reflect_with_sign)(m - i, QR + i + i * lda, nrhs, B + i, ldb, F(signum)(diag[i]));
}

/* R B in place, R = the n x n upper triangle of QR::r (qr.rs:81-89, R_kk =
 * |diag_k|), in the reference BLAS TRMM loop (side left, upper, no
 * transpose, non-unit, alpha 1): for each column, for k ascending, a nonzero
 * b_k adds b_k R_ik to every b_i above it and becomes b_k R_kk. */
void F(qr_rmul)(long n, const T *QR, long lda, const T *diag, long nrhs, T *B, long ldb)
{
    for (long j = 0; j < nrhs; ++j) {
        T *b = B + j * ldb;
        for (long k = 0; k < n; ++k) {
            if (b[k] != (T)0) {
                T temp = b[k];
                for (long i = 0; i < k; ++i) b[i] = b[i] + temp * QR[i + k * lda];
                temp = temp * FABS(diag[k]);
                b[k] = temp;
            }
        }
    }
}

/* QR's private solve_upper_triangular_mut: qr.rs:224-256, on the leading
 * n x n block (R_ii = |diag_i|, R_ri = QR[r,i] above the diagonal).
 * Returns 0 on an exact-zero diagonal. */
int F(qr_rsolve)(long n, const T *QR, long lda, const T *diag, long nrhs, T *B, long ldb)
{
    for (long c = 0; c < nrhs; ++c) {
        T *b = B + c * ldb;
        for (long i = n - 1; i >= 0; --i) {
            T d = FABS(diag[i]);
            if (d == 0)
                return 0;
            T coeff = b[i] / d;
            b[i] = coeff;
            F(axcpy)(i, b, 1, -coeff, QR + i * lda, 1, (T)1, (T)1);
        }
    }
    return 1;
}

/* Tall least squares (m >= n): QR::new + q_tr_mul + upper solve on rows 0..n,
 * the core-nalgebra composition equivalent to nalgebra-lapack's
 * QrDecomposition::solve (nalgebra-lapack/src/qr_util.rs:167-230).
 * A (m x n) is overwritten by the factorization, b (m) by Q^T b; x (n) gets
 * the solution.  Returns 0 when R is singular. */
int F(qr_ls)(long m, long n, T *A, long lda, T *diag, T *b, T *x)
{
    F(geqrf)(m, n, A, lda, diag);
    F(qr_qtmul)(m, n, A, lda, diag, 1, b, m);
    for (long i = 0; i < n; ++i)
        x[i] = b[i];
    return F(qr_rsolve)(n, A, lda, diag, 1, x, n);
}
