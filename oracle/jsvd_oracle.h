/*
 * jsvd_oracle.h -- CPU restatement of the reference one-sided Jacobi SVD.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the HIP product is
 * compared against, and the CPU baseline timed beside it.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product path never calls it.
 *
 * Every function states the reference file:line it restates (paths are
 * relative to the root of kleedaisuki/jacobi-svd-cuda-opt).
 */
#ifndef JSVD_ORACLE_H
#define JSVD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes, the same meaning as the product's jsvd_status. */
#define JSVD_ORACLE_OK 0
#define JSVD_ORACLE_EINVAL 1

/* Circle-method column at a schedule position.
 * Reference: src/domain/jacobi_rotation_kernels.cuh:18-42.  Returns -1 for
 * the dummy column of an odd n. */
int32_t jsvd_oracle_column_at(int32_t position, int32_t round, int32_t n, int32_t even_n);

/* Host round-robin table.  Reference: src/domain/jacobi_schedule.cuh:16-57.
 * pairs_out holds (E-1)*(E/2) int32 pairs (p,q) laid out round-major; slots
 * with a dummy column are omitted, counts_out[r] is the pair count of round r.
 * Returns the number of rounds (E-1, or 0 for n<2). */
int32_t jsvd_oracle_schedule(int32_t n, int32_t *pairs_out, int32_t *counts_out);

/* Rutishauser rotation with the reference's threshold test.
 * Reference: src/domain/jacobi_rotation_kernels.cuh:305-326 (same as :175-196).
 * Returns 1 when the pair counts as rotated (sets the sweep flag). */
int32_t jsvd_oracle_rotation(double pp, double qq, double pq, double epsilon, double *c, double *s);

/* Input validation.  Reference: src/domain/jacobi_svd.cu:197-223 and :36-54
 * (file lines of validate_one_sided_inputs / validate_layout_transpose_config).
 * Returns JSVD_ORACLE_OK or JSVD_ORACLE_EINVAL and the reference's message. */
int32_t jsvd_oracle_validate(uint64_t rows, uint64_t columns, uint64_t input_size, double epsilon,
                             int32_t max_sweeps, const char **message);

/* Full FP64 solve, single thread.  Reference: src/domain/jacobi_svd.cu:234-397
 * (sweep loop :138-188 / :286-373, U/Sigma :370-413 of the kernels file).
 * a_rm: m x n row-major.  u_rm: m x n row-major, sigma: n, v_rm: n x n
 * row-major.  threads > 1 runs the disjoint pairs of each round on that many
 * OpenMP threads (same arithmetic, same result). */
int32_t jsvd_oracle_solve_f64(const double *a_rm, uint64_t m, uint64_t n, double epsilon, int32_t max_sweeps,
                              int32_t threads, double *u_rm, double *sigma, double *v_rm, int32_t *sweeps);

/* Many independent solves, one matrix per OpenMP thread.  Matrix k reads
 * a + a_off[k] (m[k] x n[k] row-major) and writes u + u_off[k], s + s_off[k],
 * v + v_off[k]; offsets are in elements. */
int32_t jsvd_oracle_solve_batch_f64(int64_t count, const double *a, const int64_t *a_off, const int64_t *m,
                                    const int64_t *n, double epsilon, int32_t max_sweeps, int32_t threads,
                                    double *u, const int64_t *u_off, double *s, const int64_t *s_off, double *v,
                                    const int64_t *v_off, int32_t *sweeps);

#ifdef __cplusplus
}
#endif

#endif /* JSVD_ORACLE_H */
