// Host preprocess: validation, MC64 static pivoting, AMD ordering,
// permutation, MC64 scaling and the permuted dump the symbolic phase reads.
//
// Sequence and semantics follow the reference's path:
//   preprocess_from_arrays      src/preprocess/preprocess_arrays.c:18-98
//   NicsLU_CreateMatrix checks  src/nicslu/source/nicslu_check.c:8-24
//   NicsLU_Analyze              src/nicslu/source/nicslu_analyze.c:12-364
//     _I_NicsLU_Check           nicslu_check.c:26-77
//     _I_NicsLU_ConstructCSR    nicslu_internal.c:39-66
//     _I_NicsLU_AMDSort         nicslu_amd.c:27-57
//     _I_NicsLU_Permute         nicslu_internal.c:68-97
//     scale vectors             nicslu_analyze.c:328-335
//     _I_NicsLU_MC64Scale       nicslu_mc64scale.c:8-56
//   my_DumpA                    src/preprocess/preprocess.c:8-100
// with NICSLU's defaults (nicslu_initialize.c:35-48, nicslu_default.h):
// MC64 on, no extra scaling, AMD alpha = 10, aggressive absorption on.
#include "glu_host.h"

#include <cmath>
#include <cstring>

namespace glu {

namespace {
constexpr double kAmdAlpha = 10.0;  // NICSLU_AMD_FLAG2
constexpr int kAmdAggressive = 1;   // NICSLU_AMD_FLAG1
}  // namespace

int analyze(uint32_t n, uint32_t nnz, const double *ax, const uint32_t *ai,
            const uint32_t *ap, Analysis &out, std::string &err)
{
    // Structure checks of NicsLU_CreateMatrix (nicslu_check.c:8-24).  The
    // monotonicity test starts at i=1, as the reference's does.
    if (n == 0 || nnz == 0 || !ax || !ai || !ap) {
        err = "empty or null matrix";
        return kErrInvalid;
    }
    if (ap[0] != 0 || ap[n] != nnz) {
        err = "indptr must start at 0 and end at nnz";
        return kErrInvalid;
    }
    for (uint32_t i = 1; i < n; ++i)
        if (ap[i] > ap[i + 1]) {
            err = "indptr is not monotone";
            return kErrInvalid;
        }
    for (uint32_t k = 0; k < nnz; ++k)
        if (ai[k] >= n) {
            err = "row index out of range";
            return kErrInvalid;
        }

    // Static pivoting.
    std::vector<int32_t> match;
    std::vector<double> u, v;
    int64_t rank = mc64_match(n, ai, ap, ax, match, u, v);
    if (rank < 0 || (uint32_t)rank < n) {
        err = "matrix is structurally singular";
        return kErrStructSingular;
    }

    // _I_NicsLU_Check: adjacent duplicates are an error.  Its sortedness
    // test compares the first index of a row against (unsigned)-1, so any
    // non-empty matrix reports "not sorted" and the transpose path below is
    // the one always taken.
    bool need_sort = false;
    for (uint32_t r = 0; r < n; ++r) {
        int64_t jlast = -1;
        for (uint32_t p = ap[r]; p < ap[r + 1]; ++p) {
            int64_t j = ai[p];
            if (j == jlast) {
                err = "duplicate entry";
                return kErrDuplicate;
            }
            if ((uint32_t)j < (uint32_t)jlast) need_sort = true;
            jlast = j;
        }
    }

    // Rows reordered by the matching: row i of the result is row match[i].
    std::vector<uint32_t> cp0(n + 1), ci0(nnz);
    {
        uint32_t ct = 0;
        cp0[0] = 0;
        for (uint32_t i = 0; i < n; ++i) {
            uint32_t row = (uint32_t)match[i];
            for (uint32_t k = ap[row]; k < ap[row + 1]; ++k) ci0[ct++] = ai[k];
            cp0[i + 1] = ct;
        }
    }
    std::vector<uint32_t> cp, ci;
    if (need_sort) {
        // Counting-sort transpose (sorted by construction).
        cp.assign(n + 1, 0);
        ci.assign(nnz, 0);
        std::vector<uint32_t> cnt(n, 0);
        for (uint32_t k = 0; k < nnz; ++k) ++cnt[ci0[k]];
        for (uint32_t i = 0; i < n; ++i) cp[i + 1] = cp[i] + cnt[i];
        std::memcpy(cnt.data(), cp.data(), sizeof(uint32_t) * n);
        for (uint32_t j = 0; j < n; ++j)
            for (uint32_t k = cp0[j]; k < cp0[j + 1]; ++k) ci[cnt[ci0[k]]++] = j;
    } else {
        cp.swap(cp0);
        ci.swap(ci0);
    }

    std::vector<int32_t> perm, pinv;
    amd_order(n, ci.data(), cp.data(), kAmdAlpha, kAmdAggressive, perm, pinv);

    out.n = n;
    out.nnz = nnz;
    out.row_perm.assign(n, 0);
    out.row_perm_inv.assign(n, 0);
    out.col_perm.assign(n, 0);
    out.col_perm_inv.assign(n, 0);
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t row = (uint32_t)match[perm[i]];
        out.row_perm_inv[row] = i;
        out.row_perm[i] = row;
        out.col_perm[i] = (uint32_t)perm[i];
        out.col_perm_inv[i] = (uint32_t)pinv[i];
    }

    // Scale vectors from the duals: exp(v) per row, exp(u) per permuted col.
    out.row_scale.assign(n, 1.0);
    out.col_scale_perm.assign(n, 1.0);
    for (uint32_t j = 0; j < n; ++j) {
        out.row_scale[j] = std::exp(v[j]);
        out.col_scale_perm[out.col_perm_inv[j]] = std::exp(u[j]);
    }

    // Relabelled, scaled values in the caller's storage order; note the
    // product order ax * (rs * csp) of the analysis pass.
    std::vector<uint32_t> ai_perm(nnz);
    std::vector<double> ax_scaled(nnz);
    for (uint32_t r = 0; r < n; ++r) {
        double s = out.row_scale[r];
        for (uint32_t k = ap[r]; k < ap[r + 1]; ++k) {
            uint32_t c = (uint32_t)pinv[ai[k]];
            ai_perm[k] = c;
            ax_scaled[k] = ax[k] * (s * out.col_scale_perm[c]);
        }
    }

    // my_DumpA: GLU column i is row row_perm[i], entries in storage order.
    out.ap.assign(n + 1, 0);
    out.ai.assign(nnz, 0);
    out.ax.assign(nnz, 0.0);
    out.src_pos.assign(nnz, 0);
    uint32_t p = 0;
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t old = out.row_perm[i];
        for (uint32_t k = ap[old]; k < ap[old + 1]; ++k) {
            out.ax[p] = ax_scaled[k];
            out.ai[p] = ai_perm[k];
            out.src_pos[p] = k;
            ++p;
        }
        out.ap[i + 1] = p;
    }
    out.in_ap.assign(ap, ap + n + 1);
    out.in_ai.assign(ai, ai + nnz);
    return kOk;
}

}  // namespace glu
