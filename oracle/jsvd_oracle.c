/*
 * jsvd_oracle.c -- CPU restatement of the reference one-sided Jacobi SVD.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product and the CPU
 * baseline timed beside it.  The product never links or calls this file.
 *
 * Algorithm (Hestenes one-sided Jacobi, Rutishauser rotation, circle-method
 * ordering) restated from the reference's behaviour:
 *   schedule        src/domain/jacobi_schedule.cuh:16-57,
 *                   src/domain/jacobi_rotation_kernels.cuh:18-42
 *   dots + rotation src/domain/jacobi_rotation_kernels.cuh:276-326
 *   column update   src/domain/jacobi_rotation_kernels.cuh:329-353 (A, then V)
 *   sweep loop      src/domain/jacobi_svd.cu:138-188 (flag per sweep; the
 *                   verifying sweep is counted; max_sweeps stops silently)
 *   U / Sigma       src/domain/jacobi_rotation_kernels.cuh:370-413
 *   validation      src/domain/jacobi_svd.cu:197-223
 * Dot products are summed sequentially over rows (the reference sums
 * thread-strided partials and a shared-memory tree), so results agree with
 * the reference to rounding, not bit for bit.
 */
#include "jsvd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

int32_t jsvd_oracle_column_at(int32_t position, int32_t round, int32_t n, int32_t even_n)
{
    /* Slot 0 always holds column 0; slots 1..E-1 rotate by one per round. */
    if (position == 0)
        return 0;
    const int32_t ring = even_n - 1;
    int32_t r = (position - 1 - round) % ring;
    if (r < 0)
        r += ring;
    const int32_t column = 1 + r;
    return column < n ? column : -1;
}

int32_t jsvd_oracle_schedule(int32_t n, int32_t *pairs_out, int32_t *counts_out)
{
    /* Explicit circle array, rotated after every round (the host-table form). */
    if (n < 2)
        return 0;
    const int32_t even_n = n + (n & 1);
    int32_t *circle = (int32_t *)malloc(sizeof(int32_t) * (size_t)even_n);
    for (int32_t i = 0; i < even_n; ++i)
        circle[i] = i < n ? i : -1;
    const int32_t half = even_n / 2;
    for (int32_t round = 0; round < even_n - 1; ++round)
    {
        int32_t count = 0;
        int32_t *dst = pairs_out + (size_t)round * (size_t)half * 2;
        for (int32_t i = 0; i < half; ++i)
        {
            const int32_t lhs = circle[i];
            const int32_t rhs = circle[even_n - 1 - i];
            if (lhs >= 0 && rhs >= 0)
            {
                dst[2 * count] = lhs;
                dst[2 * count + 1] = rhs;
                ++count;
            }
        }
        counts_out[round] = count;
        const int32_t last = circle[even_n - 1];
        for (int32_t i = even_n - 1; i > 1; --i)
            circle[i] = circle[i - 1];
        circle[1] = last;
    }
    free(circle);
    return even_n - 1;
}

int32_t jsvd_oracle_rotation(double pp, double qq, double pq, double epsilon, double *c, double *s)
{
    const double rhs = epsilon * sqrt(fmax(pp, 0.0) * fmax(qq, 0.0));
    if (fabs(pq) > rhs && fabs(pq) > 1.0e-300)
    {
        const double zeta = (qq - pp) / (2.0 * pq);
        const double t = (zeta >= 0.0) ? (1.0 / (zeta + sqrt(1.0 + zeta * zeta)))
                                       : (-1.0 / (-zeta + sqrt(1.0 + zeta * zeta)));
        *c = 1.0 / sqrt(1.0 + t * t);
        *s = t * *c;
        return 1;
    }
    *c = 1.0;
    *s = 0.0;
    return 0;
}

int32_t jsvd_oracle_validate(uint64_t rows, uint64_t columns, uint64_t input_size, double epsilon,
                             int32_t max_sweeps, const char **message)
{
    const char *msg = 0;
    if (rows == 0 || columns == 0)
        msg = "Input matrix shape must be non-zero.";
    else if (rows < columns)
        msg = "One-sided Jacobi SVD requires rows >= columns in this implementation.";
    else if (input_size != rows * columns)
        msg = "Input buffer size mismatch.";
    else if (epsilon <= 0.0) /* jacobi_svd.cu:211: a NaN epsilon passes, as there */
        msg = "epsilon must be positive.";
    else if (max_sweeps <= 0)
        msg = "max_sweeps must be positive.";
    if (message)
        *message = msg;
    return msg ? JSVD_ORACLE_EINVAL : JSVD_ORACLE_OK;
}

/* One pair: dots over m rows, rotation, update of A (m rows) then V (n rows).
 * a and v are column-major working copies. */
static int32_t process_pair(double *a, double *v, int64_t m, int64_t n, int32_t p, int32_t q, double epsilon)
{
    double *ap = a + (size_t)p * (size_t)m;
    double *aq = a + (size_t)q * (size_t)m;
    double pp = 0.0, qq = 0.0, pq = 0.0;
    for (int64_t i = 0; i < m; ++i)
    {
        pp += ap[i] * ap[i];
        qq += aq[i] * aq[i];
        pq += ap[i] * aq[i];
    }
    double c, s;
    const int32_t rotated = jsvd_oracle_rotation(pp, qq, pq, epsilon, &c, &s);
    if (s != 0.0)
    {
        for (int64_t i = 0; i < m; ++i)
        {
            const double x = ap[i], y = aq[i];
            ap[i] = c * x - s * y;
            aq[i] = s * x + c * y;
        }
        double *vp = v + (size_t)p * (size_t)n;
        double *vq = v + (size_t)q * (size_t)n;
        for (int64_t i = 0; i < n; ++i)
        {
            const double x = vp[i], y = vq[i];
            vp[i] = c * x - s * y;
            vq[i] = s * x + c * y;
        }
    }
    return rotated;
}

static void solve_core(const double *a_rm, int64_t m, int64_t n, double epsilon, int32_t max_sweeps,
                       int32_t threads, double *u_rm, double *sigma, double *v_rm, int32_t *sweeps)
{
    double *a = (double *)malloc(sizeof(double) * (size_t)(m * n));
    double *v = (double *)malloc(sizeof(double) * (size_t)(n * n));
    for (int64_t i = 0; i < m; ++i)
        for (int64_t j = 0; j < n; ++j)
            a[j * m + i] = a_rm[i * n + j];
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < n; ++i)
            v[j * n + i] = (i == j) ? 1.0 : 0.0;

    const int32_t even_n = (int32_t)(n + (n & 1));
    const int32_t rounds = n < 2 ? 0 : even_n - 1;
    const int32_t half = even_n / 2;
    int32_t executed = 0;
    for (int32_t sweep = 0; sweep < max_sweeps; ++sweep)
    {
        int32_t any = 0;
        for (int32_t round = 0; round < rounds; ++round)
        {
            if (threads > 1)
            {
#ifdef _OPENMP
#pragma omp parallel for num_threads(threads) reduction(| : any) schedule(static)
#endif
                for (int32_t slot = 0; slot < half; ++slot)
                {
                    const int32_t p = jsvd_oracle_column_at(slot, round, (int32_t)n, even_n);
                    const int32_t q = jsvd_oracle_column_at(even_n - 1 - slot, round, (int32_t)n, even_n);
                    if (p >= 0 && q >= 0)
                        any |= process_pair(a, v, m, n, p, q, epsilon);
                }
            }
            else
            {
                for (int32_t slot = 0; slot < half; ++slot)
                {
                    const int32_t p = jsvd_oracle_column_at(slot, round, (int32_t)n, even_n);
                    const int32_t q = jsvd_oracle_column_at(even_n - 1 - slot, round, (int32_t)n, even_n);
                    if (p >= 0 && q >= 0)
                        any |= process_pair(a, v, m, n, p, q, epsilon);
                }
            }
        }
        executed = sweep + 1;
        if (!any)
            break;
    }

    for (int64_t j = 0; j < n; ++j)
    {
        const double *col = a + (size_t)j * (size_t)m;
        double norm2 = 0.0;
        for (int64_t i = 0; i < m; ++i)
            norm2 += col[i] * col[i];
        const double sj = sqrt(fmax(norm2, 0.0));
        sigma[j] = sj;
        for (int64_t i = 0; i < m; ++i)
            u_rm[i * n + j] = (sj > epsilon) ? (col[i] / sj) : 0.0;
    }
    for (int64_t i = 0; i < n; ++i)
        for (int64_t j = 0; j < n; ++j)
            v_rm[i * n + j] = v[j * n + i];
    *sweeps = executed;
    free(a);
    free(v);
}

int32_t jsvd_oracle_solve_f64(const double *a_rm, uint64_t m, uint64_t n, double epsilon, int32_t max_sweeps,
                              int32_t threads, double *u_rm, double *sigma, double *v_rm, int32_t *sweeps)
{
    if (jsvd_oracle_validate(m, n, m * n, epsilon, max_sweeps, 0) != JSVD_ORACLE_OK)
        return JSVD_ORACLE_EINVAL;
    solve_core(a_rm, (int64_t)m, (int64_t)n, epsilon, max_sweeps, threads, u_rm, sigma, v_rm, sweeps);
    return JSVD_ORACLE_OK;
}

int32_t jsvd_oracle_solve_batch_f64(int64_t count, const double *a, const int64_t *a_off, const int64_t *m,
                                    const int64_t *n, double epsilon, int32_t max_sweeps, int32_t threads,
                                    double *u, const int64_t *u_off, double *s, const int64_t *s_off, double *v,
                                    const int64_t *v_off, int32_t *sweeps)
{
    for (int64_t k = 0; k < count; ++k)
        if (jsvd_oracle_validate((uint64_t)m[k], (uint64_t)n[k], (uint64_t)(m[k] * n[k]), epsilon, max_sweeps, 0))
            return JSVD_ORACLE_EINVAL;
#ifdef _OPENMP
#pragma omp parallel for num_threads(threads > 0 ? threads : 1) schedule(dynamic, 1)
#endif
    for (int64_t k = 0; k < count; ++k)
        solve_core(a + a_off[k], m[k], n[k], epsilon, max_sweeps, 1, u + u_off[k], s + s_off[k], v + v_off[k],
                   sweeps + k);
    (void)threads;
    return JSVD_ORACLE_OK;
}
