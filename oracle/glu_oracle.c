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
 *   oracle_refresh     NICSLU's refactor scaling (nicslu_mc64scale.c:58-85)
 *                      followed by my_DumpA (preprocess.c:63-78) and
 *                      predictLU (symbolic.cc:82-103).
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

/* oracle_factor_f32 with the dense scatter's zero updates left out: for
 * column k and each j in U(k,:), only the rows of L(:,k) are updated in
 * column j (found by a galloping merge of the two sorted row lists), in the
 * same order and with the same rounding.  A skipped update adds -(0 * u),
 * a signed zero, which leaves every finite value other than -0.0 unchanged;
 * so this equals oracle_factor_f32 bit for bit unless an entry holds -0.0
 * or a factor is infinite or NaN.  Its cost is the number of scalar updates
 * instead of the column lengths they fall in (n = 500k MNA: seconds instead
 * of minutes). */
float oracle_factor_f32_sparse(const oracle_pattern *P, float *val, int perturb)
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
    const uint32_t *rows = P->sym_r_idx;
    for (int lv = 0; lv < P->num_lev; ++lv) {
        for (int t = P->level_ptr[lv]; t < P->level_ptr[lv + 1]; ++t) {
            const uint32_t k = (uint32_t)P->level_idx[t];
            const uint32_t dpos = P->l_col_ptr[k];
            const uint32_t lb = dpos + 1, le = P->sym_c_ptr[k + 1];
            for (uint32_t p = lb; p < le; ++p) {
                if (perturb && fabsf(val[dpos]) < pert) val[dpos] = pert;
                val[p] = val[p] / val[dpos];
            }
            if (lb == le) continue;
            for (uint32_t q = P->csr_diag_ptr[k] + 1; q < P->csr_r_ptr[k + 1]; ++q) {
                const uint32_t j = P->csr_c_idx[q];
                const uint32_t end = P->sym_c_ptr[j + 1];
                uint32_t lo = P->sym_c_ptr[j], hi = end;
                while (lo < hi) { /* position of row k in column j */
                    const uint32_t mid = lo + (hi - lo) / 2;
                    if (rows[mid] < k) lo = mid + 1;
                    else hi = mid;
                }
                const float s = val[lo];
                uint32_t e = lo + 1;
                for (uint32_t p = lb; p < le; ++p) {
                    const uint32_t r = rows[p];
                    uint32_t a = e, b = e, step = 1;
                    while (b < end && rows[b] < r) {
                        a = b + 1;
                        b += step;
                        step <<= 1;
                    }
                    if (b > end) b = end;
                    while (a < b) {
                        const uint32_t mid = a + (b - a) / 2;
                        if (rows[mid] < r) a = mid + 1;
                        else b = mid;
                    }
                    val[a] = val[a] + (-(val[p] * s));
                    e = a + 1;
                }
            }
        }
    }
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

/* The FP64 elimination walks, in each updated column j, only the rows of
 * L(:,k) (a galloping merge of the two sorted row lists), as
 * oracle_factor_f32_sparse does: the updates of the other rows would
 * subtract zero. */
void oracle_factor_f64(const oracle_pattern *P, double *val)
{
    const uint32_t *rows = P->sym_r_idx;
    for (int lv = 0; lv < P->num_lev; ++lv) {
        for (int t = P->level_ptr[lv]; t < P->level_ptr[lv + 1]; ++t) {
            const uint32_t k = (uint32_t)P->level_idx[t];
            const uint32_t dpos = P->l_col_ptr[k];
            const uint32_t lb = dpos + 1, le = P->sym_c_ptr[k + 1];
            for (uint32_t p = lb; p < le; ++p) val[p] = val[p] / val[dpos];
            if (lb == le) continue;
            for (uint32_t q = P->csr_diag_ptr[k] + 1; q < P->csr_r_ptr[k + 1]; ++q) {
                const uint32_t j = P->csr_c_idx[q];
                const uint32_t end = P->sym_c_ptr[j + 1];
                uint32_t lo = P->sym_c_ptr[j], hi = end;
                while (lo < hi) {
                    const uint32_t mid = lo + (hi - lo) / 2;
                    if (rows[mid] < k) lo = mid + 1;
                    else hi = mid;
                }
                const double s = val[lo];
                uint32_t e = lo + 1;
                for (uint32_t p = lb; p < le; ++p) {
                    const uint32_t r = rows[p];
                    uint32_t a = e, b = e, step = 1;
                    while (b < end && rows[b] < r) {
                        a = b + 1;
                        b += step;
                        step <<= 1;
                    }
                    if (b > end) b = end;
                    while (a < b) {
                        const uint32_t mid = a + (b - a) / 2;
                        if (rows[mid] < r) a = mid + 1;
                        else b = mid;
                    }
                    val[a] -= val[p] * s;
                    e = a + 1;
                }
            }
        }
    }
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

/* Refactor values of a new value set ax0 with the frozen analysis:
 *   1. _I_NicsLU_MC64ScaleForRefact (src/nicslu/source/nicslu_mc64scale.c:
 *      58-85, called by NicsLU_ReFactorize, nicslu_refact.c:74): NICSLU row
 *      i is the caller's CSC column i, its column index is col_perm_inv of
 *      the caller's row (_I_NicsLU_Permute, nicslu_internal.c:68-97), and
 *      ax[j] = ax0[j] * s * csp[ai[j]], evaluated left to right;
 *      order = 1 evaluates ax0[j] * (s * csp[ai[j]]) instead, the analysis
 *      product of _I_NicsLU_MC64Scale (:44), for tests that must tell the
 *      two apart;
 *   2. my_DumpA (src/preprocess/preprocess.c:63-78): GLU column c holds
 *      NICSLU row row_perm[c] with its entries in order (pivots identity);
 *   3. predictLU (src/symbolic.cc:82-103): each pattern entry takes the
 *      first dumped entry of its column with the same row, cast to FP32,
 *      and 0 where there is none.
 * dump_ax (may be NULL) receives the dumped FP64 values of step 2. */
void oracle_refresh(const oracle_pattern *P, uint32_t nnz_a, const uint32_t *ap, const uint32_t *ai,
                    const double *ax0, const uint32_t *row_perm, const uint32_t *col_perm_inv,
                    const double *row_scale, const double *col_scale_perm, int order, double *dump_ax,
                    float *val)
{
    const uint32_t n = P->n;
    double *sc = (double *)malloc(sizeof(double) * (nnz_a ? nnz_a : 1));
    for (uint32_t i = 0; i < n; ++i) {
        const double s = row_scale[i];
        for (uint32_t j = ap[i]; j < ap[i + 1]; ++j) {
            const double c = col_scale_perm[col_perm_inv[ai[j]]];
            sc[j] = order ? ax0[j] * (s * c) : ax0[j] * s * c;
        }
    }
    double *w = (double *)malloc(sizeof(double) * (n ? n : 1));
    uint8_t *has = (uint8_t *)calloc(n ? n : 1, 1);
    uint32_t q = 0;
    for (uint32_t c = 0; c < n; ++c) {
        const uint32_t old = row_perm[c], b = ap[old], e = ap[old + 1];
        if (dump_ax)
            for (uint32_t j = b; j < e; ++j) dump_ax[q++] = sc[j];
        for (uint32_t j = e; j-- > b;) { /* reverse: the first match wins, as find() */
            const uint32_t r = col_perm_inv[ai[j]];
            w[r] = sc[j];
            has[r] = 1;
        }
        for (uint32_t p = P->sym_c_ptr[c]; p < P->sym_c_ptr[c + 1]; ++p) {
            const uint32_t r = P->sym_r_idx[p];
            val[p] = has[r] ? (float)w[r] : 0.0f;
        }
        for (uint32_t j = b; j < e; ++j) has[col_perm_inv[ai[j]]] = 0;
    }
    free(has);
    free(w);
    free(sc);
}
