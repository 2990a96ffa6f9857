/*
 * Oracle of the GLU numeric phase — TEST INFRASTRUCTURE ONLY (the checker
 * for tests/, __graft_entry__.smoke() and bench.py's CPU baseline; never
 * linked by the product).  Compiled with -ffp-contract=off so that every
 * product is rounded before it is added, as in the reference.
 *
 * Each function restates the reference's arithmetic, sequentially:
 *
 *   oracle_factor_f32  src/numeric.cu — GESP value :490-506; per level
 *                      (:551-617) and per column of the level, step 1
 *                      :37-47 (RL) / :133-146 (RL_perturb: the pivot test
 *                      runs only where L(:,k) is non-empty), step 2 :51-89
 *                      (dense scatter of L(:,k) into tmpMem, every entry of
 *                      sub-column j below row k receives += -(tmp*U(k,j)),
 *                      zero for rows outside L(:,k)), step 3 :91-102.  The
 *                      stream-mode trio :204-345 computes the same values.
 *                      Columns of one level are taken in level_idx order;
 *                      the reference's atomics may order the updates an
 *                      entry receives within one level differently.
 *   oracle_solve_f32   src/symbolic.cc:200-251 (pivot = identity after
 *                      my_DumpA, MC64 scaling on).
 *   oracle_factor_f64  the same right-looking elimination in FP64 — the
 *                      FP64 LU of the same permuted matrix.
 *   oracle_solve_f64   FP64 permute/scale + triangular solves, A or A'.
 *   oracle_abft_*      src/symbolic.cc:163-197 (ABFT column checksums).
 */
#include "glu_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

float oracle_factor_f32(const oracle_pattern *P, float *val, int perturb)
{
    const uint32_t n = P->n;
    float pert = 0.0f;
    if (perturb) {
        float norm_a = 0.0f;
        for (uint32_t i = 0; i < n; ++i) {
            float tmp = 0.0f;
            for (uint32_t j = P->sym_c_ptr[i]; j < P->sym_c_ptr[i + 1]; ++j) tmp += fabsf(val[j]);
            if (norm_a < tmp) norm_a = tmp;
        }
        pert = 3.45e-4f * norm_a;
    }
    float *tmp = (float *)calloc(n ? n : 1, sizeof(float));
    for (int lv = 0; lv < P->num_lev; ++lv) {
        for (int t = P->level_ptr[lv]; t < P->level_ptr[lv + 1]; ++t) {
            const uint32_t k = (uint32_t)P->level_idx[t];
            const uint32_t dpos = P->l_col_ptr[k];
            const uint32_t lb = dpos + 1, le = P->sym_c_ptr[k + 1];
            for (uint32_t p = lb; p < le; ++p) {
                if (perturb && fabsf(val[dpos]) < pert) val[dpos] = pert;
                val[p] = val[p] / val[dpos];
                tmp[P->sym_r_idx[p]] = val[p];
            }
            for (uint32_t q = P->csr_diag_ptr[k] + 1; q < P->csr_r_ptr[k + 1]; ++q) {
                const uint32_t j = P->csr_c_idx[q];
                float s = 0.0f;
                for (uint32_t e = P->sym_c_ptr[j]; e < P->sym_c_ptr[j + 1]; ++e) {
                    const uint32_t r = P->sym_r_idx[e];
                    if (r == k) s = val[e];
                    if (r > k) val[e] = val[e] + (-(tmp[r] * s));
                }
            }
            for (uint32_t p = lb; p < le; ++p) tmp[P->sym_r_idx[p]] = 0.0f;
        }
    }
    free(tmp);
    return pert;
}

uint32_t oracle_conflicts(const oracle_pattern *P, uint8_t *mark)
{
    const uint32_t n = P->n;
    uint32_t *cnt = (uint32_t *)calloc(P->nnz ? P->nnz : 1, sizeof(uint32_t));
    uint32_t *stamp = (uint32_t *)malloc(sizeof(uint32_t) * (P->nnz ? P->nnz : 1));
    uint8_t *inl = (uint8_t *)calloc(n ? n : 1, 1);
    uint32_t total = 0;
    for (uint32_t p = 0; p < P->nnz; ++p) stamp[p] = 0xffffffffu;
    memset(mark, 0, P->nnz);
    for (int lv = 0; lv < P->num_lev; ++lv) {
        for (int t = P->level_ptr[lv]; t < P->level_ptr[lv + 1]; ++t) {
            const uint32_t k = (uint32_t)P->level_idx[t];
            for (uint32_t p = P->l_col_ptr[k] + 1; p < P->sym_c_ptr[k + 1]; ++p) inl[P->sym_r_idx[p]] = 1;
            for (uint32_t q = P->csr_diag_ptr[k] + 1; q < P->csr_r_ptr[k + 1]; ++q) {
                const uint32_t j = P->csr_c_idx[q];
                for (uint32_t e = P->sym_c_ptr[j]; e < P->sym_c_ptr[j + 1]; ++e) {
                    if (!inl[P->sym_r_idx[e]]) continue;
                    if (stamp[e] != (uint32_t)lv) {
                        stamp[e] = (uint32_t)lv;
                        cnt[e] = 0;
                    }
                    if (++cnt[e] == 2) {
                        mark[e] = 1;
                        ++total;
                    }
                }
            }
            for (uint32_t p = P->l_col_ptr[k] + 1; p < P->sym_c_ptr[k + 1]; ++p) inl[P->sym_r_idx[p]] = 0;
        }
    }
    free(cnt);
    free(stamp);
    free(inl);
    return total;
}

void oracle_solve_f32(const oracle_pattern *P, const float *lu, const uint32_t *col_perm,
                      const uint32_t *row_perm_inv, const double *col_scale_perm,
                      const double *row_scale, const float *rhs, float *x)
{
    const uint32_t n = P->n;
    float *b = (float *)malloc(sizeof(float) * (n ? n : 1));
    for (uint32_t j = 0; j < n; ++j) b[j] = (float)(rhs[col_perm[j]] * col_scale_perm[j]);
    for (uint32_t j = 0; j < n; ++j)
        for (uint32_t p = P->l_col_ptr[j] + 1; p < P->sym_c_ptr[j + 1]; ++p)
            b[P->sym_r_idx[p]] -= lu[p] * b[j];
    for (int64_t jj = (int64_t)n - 1; jj >= 0; --jj) {
        b[jj] /= lu[P->l_col_ptr[jj]];
        for (uint32_t p = P->sym_c_ptr[jj]; p < P->l_col_ptr[jj]; ++p) b[P->sym_r_idx[p]] -= lu[p] * b[jj];
    }
    for (uint32_t j = 0; j < n; ++j) x[j] = (float)(b[row_perm_inv[j]] * row_scale[j]);
    free(b);
}

void oracle_factor_f64(const oracle_pattern *P, double *val)
{
    const uint32_t n = P->n;
    double *tmp = (double *)calloc(n ? n : 1, sizeof(double));
    for (int lv = 0; lv < P->num_lev; ++lv) {
        for (int t = P->level_ptr[lv]; t < P->level_ptr[lv + 1]; ++t) {
            const uint32_t k = (uint32_t)P->level_idx[t];
            const uint32_t dpos = P->l_col_ptr[k];
            for (uint32_t p = dpos + 1; p < P->sym_c_ptr[k + 1]; ++p) {
                val[p] = val[p] / val[dpos];
                tmp[P->sym_r_idx[p]] = val[p];
            }
            for (uint32_t q = P->csr_diag_ptr[k] + 1; q < P->csr_r_ptr[k + 1]; ++q) {
                const uint32_t j = P->csr_c_idx[q];
                double s = 0.0;
                for (uint32_t e = P->sym_c_ptr[j]; e < P->sym_c_ptr[j + 1]; ++e) {
                    const uint32_t r = P->sym_r_idx[e];
                    if (r == k) s = val[e];
                    if (r > k) val[e] -= tmp[r] * s;
                }
            }
            for (uint32_t p = dpos + 1; p < P->sym_c_ptr[k + 1]; ++p) tmp[P->sym_r_idx[p]] = 0.0;
        }
    }
    free(tmp);
}

void oracle_solve_f64(const oracle_pattern *P, const double *lu, const uint32_t *row_perm,
                      const uint32_t *row_perm_inv, const uint32_t *col_perm,
                      const uint32_t *col_perm_inv, const double *col_scale_perm,
                      const double *row_scale, const double *b, double *x, int transpose)
{
    const uint32_t n = P->n;
    double *y = (double *)malloc(sizeof(double) * (n ? n : 1));
    if (!transpose) {
        for (uint32_t j = 0; j < n; ++j) y[j] = b[col_perm[j]] * col_scale_perm[j];
        for (uint32_t j = 0; j < n; ++j)
            for (uint32_t p = P->l_col_ptr[j] + 1; p < P->sym_c_ptr[j + 1]; ++p) y[P->sym_r_idx[p]] -= lu[p] * y[j];
        for (int64_t jj = (int64_t)n - 1; jj >= 0; --jj) {
            y[jj] /= lu[P->l_col_ptr[jj]];
            for (uint32_t p = P->sym_c_ptr[jj]; p < P->l_col_ptr[jj]; ++p) y[P->sym_r_idx[p]] -= lu[p] * y[jj];
        }
        for (uint32_t j = 0; j < n; ++j) x[j] = y[row_perm_inv[j]] * row_scale[j];
    } else {
        /* M = Dc P A Q Dr  =>  A' x = b  <=>  M' z = w, w[c] = rs[rp[c]] b[rp[c]],
         * x[cp[r]] = csp[r] z[r]. */
        for (uint32_t c = 0; c < n; ++c) y[c] = b[row_perm[c]] * row_scale[row_perm[c]];
        for (uint32_t r = 0; r < n; ++r) { /* U' forward: column r of U */
            double acc = y[r];
            for (uint32_t p = P->sym_c_ptr[r]; p < P->l_col_ptr[r]; ++p) acc -= lu[p] * y[P->sym_r_idx[p]];
            y[r] = acc / lu[P->l_col_ptr[r]];
        }
        for (int64_t r = (int64_t)n - 1; r >= 0; --r) { /* L' backward: column r of L */
            double acc = y[r];
            for (uint32_t p = P->l_col_ptr[r] + 1; p < P->sym_c_ptr[r + 1]; ++p) acc -= lu[p] * y[P->sym_r_idx[p]];
            y[r] = acc;
        }
        for (uint32_t i = 0; i < n; ++i) x[i] = y[col_perm_inv[i]] * col_scale_perm[col_perm_inv[i]];
    }
    free(y);
}

/* src/symbolic.cc:163-170 */
void oracle_abft_cca(const oracle_pattern *P, const float *val, float *cca)
{
    for (uint32_t i = 0; i < P->n; ++i) {
        float s = 0.f;
        for (uint32_t j = P->sym_c_ptr[i]; j < P->sym_c_ptr[i + 1]; ++j) s += val[j];
        cca[i] = s;
    }
}

/* src/symbolic.cc:172-197 */
int32_t oracle_abft_check(const oracle_pattern *P, const float *lu, const float *cca, float tol, float *abft,
                          float *max_diff)
{
    const uint32_t n = P->n;
    float *ccl = (float *)malloc(sizeof(float) * (n ? n : 1));
    for (uint32_t i = 0; i < n; ++i) {
        float s = 1.f;
        for (uint32_t j = P->l_col_ptr[i] + 1; j < P->sym_c_ptr[i + 1]; ++j) s += lu[j];
        ccl[i] = s;
    }
    int32_t first = -1;
    float mx = 0.f;
    for (uint32_t i = 0; i < n; ++i) {
        float s = 0.f;
        for (uint32_t j = P->sym_c_ptr[i]; j <= P->l_col_ptr[i]; ++j) s += ccl[P->sym_r_idx[j]] * lu[j];
        abft[i] = s;
        const float d = fabsf(s - cca[i]);
        if (d > tol && first < 0) first = (int32_t)i;
        if (d != d || mx != mx) mx = NAN;
        else if (d > mx) mx = d;
    }
    free(ccl);
    *max_diff = mx;
    return first;
}
