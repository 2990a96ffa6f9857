// Host-side analysis for the GLU numeric phase: static pivoting (MC64),
// fill-reducing ordering (AMD on A+A'), permutation + scaling, and the
// symbolic structures the device kernels consume.
//
// Every routine here reproduces the reference's results bit for bit:
// NICSLU's NicsLU_Analyze (src/nicslu/source/nicslu_analyze.c:12-364) as
// driven by preprocess_from_arrays (src/preprocess/preprocess_arrays.c:18-98)
// and my_DumpA (src/preprocess/preprocess.c:8-100), then Symbolic_Matrix
// (src/symbolic.cc:11-151).  Tie-breaks in matching and ordering decide the
// pattern, so the algorithms follow the same steps in the same order.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace glu {

// Status codes shared with the C-ABI (include/glu_hip.h).  The negative
// values -5..-8 keep NICSLU's meanings (nicslu.h:11-28).
enum Status : int {
    kOk = 0,
    kErrArg = -2,
    kErrMemory = -3,
    kErrStructSingular = -5,
    kErrNumericSingular = -6,
    kErrInvalid = -7,
    kErrDuplicate = -8,
};

// Result of the host preprocess.  "NICSLU orientation": the caller's CSC
// column pointers are NICSLU's row pointers (the reference hands its CSC
// arrays to NICSLU unchanged with cfgi[0]=1), so a "row" below is a column of
// the caller's matrix.  Field meanings follow SNicsLU (nicslu.h:44-70).
struct Analysis {
    uint32_t n = 0, nnz = 0;
    std::vector<uint32_t> row_perm, row_perm_inv, col_perm, col_perm_inv;
    std::vector<double> row_scale;       // indexed by original row
    std::vector<double> col_scale_perm;  // indexed by permuted column
    // Permuted + scaled matrix as my_DumpA emits it: GLU column c holds the
    // entries of caller column row_perm[c]; row indices are permuted labels,
    // in the caller's storage order within the column (not sorted).
    std::vector<double> ax;
    std::vector<uint32_t> ai, ap;
    // For each dumped entry, its position in the caller's ax (value refresh).
    std::vector<uint32_t> src_pos;
    // The caller's CSC structure, kept for FP64 residuals on the original A.
    std::vector<uint32_t> in_ap, in_ai;
};

// ax/ai/ap: caller CSC (n columns, ap[n] == nnz).  Returns kOk or a code.
int analyze(uint32_t n, uint32_t nnz, const double *ax, const uint32_t *ai,
            const uint32_t *ap, Analysis &out, std::string &err);

// Symbolic_Matrix fields (include/symbolic.h:12-23), same meaning.
struct Symbolic {
    uint32_t n = 0, nnz = 0;
    int num_lev = 0;
    std::vector<uint32_t> sym_c_ptr, sym_r_idx, l_col_ptr;
    std::vector<uint32_t> csr_r_ptr, csr_c_idx, csr_diag_ptr;
    std::vector<float> val;
    std::vector<int> level_ptr, level_idx;
};

// fill_in + csr + leveling (src/symbolic.cc:11-79,106-151).
void symbolic(uint32_t n, const uint32_t *ai, const uint32_t *ap, Symbolic &s);
// The three steps separately (s.n set by the caller for fill_in).
void symbolic_fill_in(uint32_t n, const uint32_t *ai, const uint32_t *ap, Symbolic &s);
void symbolic_csr(Symbolic &s);
void symbolic_leveling(Symbolic &s);
// predictLU (src/symbolic.cc:82-103): FP64 -> FP32 placement of A into the
// LU pattern, fill entries zero.  Also returns, for every LU position, the
// dumped-entry index it came from or UINT32_MAX for fill.
void predict_lu(const uint32_t *ai, const uint32_t *ap, const double *ax, Symbolic &s,
                std::vector<uint32_t> *lu_from_dump);

// MC64 maximum-product matching with dual variables (nicslu_mc64.c).
// Returns the matching size (n on success) or a negative status.
// match[col] = row matched to col; u = column duals; v = row duals.
int64_t mc64_match(uint32_t n, const uint32_t *ai, const uint32_t *ap, const double *ax,
                   std::vector<int32_t> &match, std::vector<double> &u,
                   std::vector<double> &v);

// Approximate minimum degree on the pattern of A+A' (nicslu_amd.c).
// ai/ap: row structure (rows sorted ascending, no duplicates).
// perm[k] = k-th pivot, pinv[i] = position of i.
void amd_order(uint32_t n, const uint32_t *ai, const uint32_t *ap, double alpha, int aggressive,
               std::vector<int32_t> &perm, std::vector<int32_t> &pinv);

}  // namespace glu
