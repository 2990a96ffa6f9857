// Symbolic analysis: fill-in pattern, CSR transpose, value placement and
// level sets.  Outputs equal src/symbolic.cc's field for field:
//   fill_in   (:11-53)  column i's LU pattern is the smallest superset of
//             A(:,i) closed under "row r < i present => L-pattern of column
//             r (from its diagonal down) present", plus a forced diagonal.
//             The reference builds it by repeated sorted set_union in
//             increasing r; the closure is unique, so a marker-array
//             worklist followed by one sort yields the same sorted set.
//   csr       (:55-79)  row-major copy of the pattern, columns ascending,
//             csr_diag_ptr = position of the diagonal in each row.
//   predictLU (:82-103) FP64 -> FP32 cast of A into the pattern, 0 for fill
//             (the reference's linear find picks the first match; a
//             first-writer-wins scatter picks the same one).
//   leveling  (:106-151) level(i) = 1 + max level over (a) rows r < i with
//             U(r,i) != 0 whose column r has a non-empty L, and (b) columns
//             c < i with L(i,c) != 0; level_idx lists columns level by level
//             in increasing column order.
#include "glu_host.h"

#include <algorithm>

namespace glu {

void symbolic(uint32_t n, const uint32_t *ai, const uint32_t *ap, Symbolic &s)
{
    symbolic_fill_in(n, ai, ap, s);
    symbolic_csr(s);
    symbolic_leveling(s);
}

void symbolic_fill_in(uint32_t n, const uint32_t *ai, const uint32_t *ap, Symbolic &s)
{
    s.n = n;
    s.sym_c_ptr.assign(1, 0);
    s.sym_r_idx.clear();
    s.l_col_ptr.clear();
    s.l_col_ptr.reserve(n);
    s.sym_c_ptr.reserve(n + 1);
    s.sym_r_idx.reserve((size_t)ap[n] * 3);

    std::vector<uint32_t> mark(n, UINT32_MAX);
    std::vector<uint32_t> col, work;
    for (uint32_t i = 0; i < n; ++i) {
        col.clear();
        work.clear();
        for (uint32_t k = ap[i]; k < ap[i + 1]; ++k) {
            uint32_t r = ai[k];
            if (mark[r] == i) continue;
            mark[r] = i;
            col.push_back(r);
            if (r < i) work.push_back(r);
        }
        while (!work.empty()) {
            uint32_t r = work.back();
            work.pop_back();
            for (uint32_t p = s.l_col_ptr[r]; p < s.sym_c_ptr[r + 1]; ++p) {
                uint32_t q = s.sym_r_idx[p];
                if (mark[q] == i) continue;
                mark[q] = i;
                col.push_back(q);
                if (q < i) work.push_back(q);
            }
        }
        if (mark[i] != i) {
            mark[i] = i;
            col.push_back(i);
        }
        std::sort(col.begin(), col.end());
        uint32_t base = s.sym_c_ptr.back();
        uint32_t dpos = (uint32_t)(std::lower_bound(col.begin(), col.end(), i) - col.begin());
        s.l_col_ptr.push_back(base + dpos);
        s.sym_c_ptr.push_back(base + (uint32_t)col.size());
        s.sym_r_idx.insert(s.sym_r_idx.end(), col.begin(), col.end());
    }
    s.nnz = s.sym_c_ptr.back();
}

void symbolic_csr(Symbolic &s)
{
    const uint32_t n = s.n;
    s.csr_r_ptr.assign(n + 1, 0);
    for (uint32_t p = 0; p < s.nnz; ++p) ++s.csr_r_ptr[s.sym_r_idx[p] + 1];
    for (uint32_t i = 0; i < n; ++i) s.csr_r_ptr[i + 1] += s.csr_r_ptr[i];
    s.csr_c_idx.assign(s.nnz, 0);
    s.csr_diag_ptr.assign(n, 0);
    {
        std::vector<uint32_t> fillpos(s.csr_r_ptr.begin(), s.csr_r_ptr.end() - 1);
        for (uint32_t c = 0; c < n; ++c)
            for (uint32_t p = s.sym_c_ptr[c]; p < s.sym_c_ptr[c + 1]; ++p) {
                uint32_t r = s.sym_r_idx[p];
                uint32_t q = fillpos[r]++;
                s.csr_c_idx[q] = c;
                if (r == c) s.csr_diag_ptr[r] = q;
            }
    }

}

void symbolic_leveling(Symbolic &s)
{
    const uint32_t n = s.n;
    std::vector<int> inlevel(n, 0), level_size(n, 0);
    int num_lev = 0;
    for (uint32_t i = 0; i < n; ++i) {
        int max_lv = -1;
        for (uint32_t p = s.sym_c_ptr[i]; p < s.l_col_ptr[i]; ++p) {
            uint32_t r = s.sym_r_idx[p];
            if (s.l_col_ptr[r] + 1 != s.sym_c_ptr[r + 1] && inlevel[r] > max_lv) max_lv = inlevel[r];
        }
        for (uint32_t q = s.csr_r_ptr[i]; q < s.csr_diag_ptr[i]; ++q) {
            uint32_t c = s.csr_c_idx[q];
            if (inlevel[c] > max_lv) max_lv = inlevel[c];
        }
        int lv = max_lv + 1;
        inlevel[i] = lv;
        ++level_size[lv];
        if (lv > num_lev) num_lev = lv;
    }
    ++num_lev;
    s.num_lev = num_lev;
    s.level_ptr.assign(num_lev + 1, 0);
    for (int l = 0; l < num_lev; ++l) s.level_ptr[l + 1] = s.level_ptr[l] + level_size[l];
    s.level_idx.assign(n, 0);
    std::vector<int> cursor(s.level_ptr.begin(), s.level_ptr.end() - 1);
    for (uint32_t i = 0; i < n; ++i) s.level_idx[cursor[inlevel[i]]++] = (int)i;
}

void predict_lu(const uint32_t *ai, const uint32_t *ap, const double *ax, Symbolic &s,
                std::vector<uint32_t> *lu_from_dump)
{
    uint32_t n = s.n;
    s.val.assign(s.nnz, 0.0f);
    if (lu_from_dump) lu_from_dump->assign(s.nnz, UINT32_MAX);
    std::vector<uint32_t> where(n, UINT32_MAX), stamp(n, UINT32_MAX);
    for (uint32_t i = 0; i < n; ++i) {
        for (uint32_t k = ap[i]; k < ap[i + 1]; ++k) {
            uint32_t r = ai[k];
            if (stamp[r] == i) continue;  // first occurrence wins, as find()
            stamp[r] = i;
            where[r] = k;
        }
        for (uint32_t p = s.sym_c_ptr[i]; p < s.sym_c_ptr[i + 1]; ++p) {
            uint32_t r = s.sym_r_idx[p];
            if (stamp[r] == i) {
                s.val[p] = (float)ax[where[r]];
                if (lu_from_dump) (*lu_from_dump)[p] = where[r];
            }
        }
    }
}

}  // namespace glu
