/*
 * Oracle of the GLU numeric phase — TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, single-threaded restatement of the reference's arithmetic,
 * used by tests/, __graft_entry__.smoke() and bench.py's CPU baseline as the
 * checker.  The product (libglu_hip.so) never links or calls it.
 *
 * Pattern arguments are the Symbolic_Matrix fields (include/symbolic.h:12-23
 * of the reference).
 */
#ifndef GLU_ORACLE_H
#define GLU_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_pattern {
    uint32_t n, nnz;
    int32_t num_lev;
    const uint32_t *sym_c_ptr, *sym_r_idx, *l_col_ptr;
    const uint32_t *csr_r_ptr, *csr_c_idx, *csr_diag_ptr;
    const int32_t *level_ptr, *level_idx;
} oracle_pattern;

/* FP32 right-looking factorization in place, src/numeric.cu semantics.
 * Returns the GESP value used (0 when perturb == 0). */
float oracle_factor_f32(const oracle_pattern *P, float *val, int perturb);

/* The same factorization visiting only the rows of L(:,k) in each updated
 * column (the dense scatter's other updates add a signed zero).  Equal to
 * oracle_factor_f32 bit for bit unless an entry holds -0.0 or a factor is
 * non-finite; its cost is the scalar update count (for large orders). */
float oracle_factor_f32_sparse(const oracle_pattern *P, float *val, int perturb);

/* Number of level-local updates each entry receives from more than one
 * column of the same level (the entries whose FP32 summation order the
 * reference leaves to its atomics).  mark[p] = 1 for such entries. */
uint32_t oracle_conflicts(const oracle_pattern *P, uint8_t *mark);

/* Symbolic_Matrix::solve, src/symbolic.cc:200-251 (FP32, FP64 scales). */
void oracle_solve_f32(const oracle_pattern *P, const float *lu, const uint32_t *col_perm,
                      const uint32_t *row_perm_inv, const double *col_scale_perm,
                      const double *row_scale, const float *rhs, float *x);

/* FP64 LU of the same permuted matrix on the same pattern (no pivoting,
 * column order).  val: FP64 values in the pattern, in place. */
void oracle_factor_f64(const oracle_pattern *P, double *val);

/* FP64 solve of A x = b (transpose = 0) or A' x = b with FP64 factors and the
 * same permutation/scaling as the FP32 solve. */
void oracle_solve_f64(const oracle_pattern *P, const double *lu, const uint32_t *row_perm,
                      const uint32_t *row_perm_inv, const uint32_t *col_perm,
                      const uint32_t *col_perm_inv, const double *col_scale_perm,
                      const double *row_scale, const double *b, double *x, int transpose);

/* ABFT of src/symbolic.cc:163-197 (Symbolic_Matrix::ABFTCalculateCCA and
 * ABFTCheckResult, FP32, sequential sums in the reference's order).
 * oracle_abft_cca: cca[i] = sum of column i of val (values before the
 * factorization).  oracle_abft_check: ccl[r] = 1 + sum of L(:,r);
 * abft[i] = sum_{j in U(:,i) incl. diagonal} ccl[row j] * lu[j]; returns the
 * first column with |abft - cca| > tol (the reference's test, so a NaN
 * difference passes), or -1; *max_diff = the largest |abft - cca| (NaN if
 * any difference is NaN). */
void oracle_abft_cca(const oracle_pattern *P, const float *val, float *cca);
int32_t oracle_abft_check(const oracle_pattern *P, const float *lu, const float *cca, float tol, float *abft,
                          float *max_diff);

/* Values of a refactorization with the frozen analysis: NICSLU's
 * _I_NicsLU_MC64ScaleForRefact (nicslu_mc64scale.c:58-85) on the caller's
 * CSC values ax0 (nnz_a entries, ap/ai the caller's CSC), my_DumpA
 * (preprocess.c:63-78) and predictLU (symbolic.cc:82-103).  order 0 is
 * NICSLU's refactor product (ax0 * rs) * csp, order 1 the analysis product
 * ax0 * (rs * csp).  dump_ax (nnz_a, may be NULL) receives the dumped FP64
 * values, val (nnz_lu) the FP32 values in the LU pattern. */
void oracle_refresh(const oracle_pattern *P, uint32_t nnz_a, const uint32_t *ap, const uint32_t *ai,
                    const double *ax0, const uint32_t *row_perm, const uint32_t *col_perm_inv,
                    const double *row_scale, const double *col_scale_perm, int order, double *dump_ax,
                    float *val);

#ifdef __cplusplus
}
#endif

#endif
