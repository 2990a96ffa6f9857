// Host-side check of the C++ mirror (no GPU needed): preprocess_from_arrays
// + Symbolic_Matrix::{fill_in, csr, predictLU, leveling} reproduce the arrays
// of glu_analyze on the same input, and the console lines of the reference.
// Input: a column-sorted triplet file (argv[1]).  Prints "OK" or a reason.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <sstream>
#include <vector>

#include "preprocess.h"
#include "symbolic.h"

template <typename T>
static bool same(const std::vector<T> &v, const glu_analysis *a, glu_field f)
{
    std::vector<T> w(v.size());
    if (glu_analysis_get(a, f, w.data(), sizeof(T) * w.size()) != GLU_OK) return false;
    return std::memcmp(v.data(), w.data(), sizeof(T) * v.size()) == 0;
}

int main(int argc, char **argv)
{
    if (argc < 2) return 2;
    unsigned n = 0, nnz = 0;
    double *ax = nullptr;
    unsigned *ai = nullptr, *ap = nullptr;
    if (glu_read_triplet_column(argv[1], &n, &nnz, &ax, &ai, &ap) != 0) {
        std::puts("read failed");
        return 1;
    }
    glu_analysis *an = nullptr;
    double *dx = nullptr;
    unsigned *di = nullptr, *dp = nullptr;
    if (preprocess_from_arrays(n, nnz, ax, ai, ap, &an, &dx, &di, &dp) != 0) {
        std::puts("preprocess failed");
        return 1;
    }
    std::ostringstream out, err;
    Symbolic_Matrix s(n, out, err);
    s.fill_in(di, dp);
    s.csr();
    s.predictLU(di, dp, dx);
    s.leveling();
    const bool ok = same(s.sym_c_ptr, an, GLU_F_SYM_C_PTR) && same(s.sym_r_idx, an, GLU_F_SYM_R_IDX) &&
                    same(s.l_col_ptr, an, GLU_F_L_COL_PTR) && same(s.csr_r_ptr, an, GLU_F_CSR_R_PTR) &&
                    same(s.csr_c_idx, an, GLU_F_CSR_C_IDX) && same(s.csr_diag_ptr, an, GLU_F_CSR_DIAG_PTR) &&
                    same(s.val, an, GLU_F_VAL) && same(s.level_ptr, an, GLU_F_LEVEL_PTR) &&
                    same(s.level_idx, an, GLU_F_LEVEL_IDX);
    const std::string o = out.str();
    const bool lines = o.find("Symbolic nonzero: " + std::to_string(s.nnz)) != std::string::npos &&
                       o.find("Number of levels: " + std::to_string(s.num_lev)) != std::string::npos;
    std::printf("%s n=%u nnz=%u nnz_lu=%u levels=%d\n", ok && lines ? "OK" : "MISMATCH", n, nnz, s.nnz, s.num_lev);
    glu_analysis_destroy(an);
    std::free(ax);
    std::free(ai);
    std::free(ap);
    std::free(dx);
    std::free(di);
    std::free(dp);
    return ok && lines ? 0 : 1;
}
