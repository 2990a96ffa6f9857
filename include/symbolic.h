// C++ host mirror of the reference's Symbolic_Matrix (include/symbolic.h:9-51):
// same public fields, same methods, same console lines.  The methods run this
// build's host code (bit-exact with the reference's fill_in, csr, predictLU
// and leveling); solve() runs the FP32 triangular solves on the GPU.
//
// One difference of signature: solve() takes the glu_analysis* that
// preprocess_from_arrays() returns where the reference passes its SNicsLU*
// (both carry the permutations and MC64 scale vectors of the preprocess).
#ifndef GLU_MIRROR_SYMBOLIC_H
#define GLU_MIRROR_SYMBOLIC_H

#include <iostream>
#include <memory>
#include <vector>

#include "glu_hip.h"
#include "type.h"

class Symbolic_Matrix
{
public:
    unsigned int n;
    unsigned int nnz;
    int num_lev;
    std::vector<unsigned> sym_c_ptr;
    std::vector<unsigned> sym_r_idx;
    std::vector<unsigned> csr_r_ptr;
    std::vector<unsigned> csr_c_idx;
    std::vector<unsigned> csr_diag_ptr;
    std::vector<REAL> val;
    std::vector<unsigned> l_col_ptr;  // indices of diagonal elements
    std::vector<int> level_idx;
    std::vector<int> level_ptr;

    void predictLU(unsigned *ai, unsigned *ap, double *ax);  // src/symbolic.cc:82-103
    void csr();                                              // src/symbolic.cc:55-79
    void leveling();                                         // src/symbolic.cc:106-151
    void fill_in(unsigned *ai, unsigned *ap);                // src/symbolic.cc:11-53
    // src/symbolic.cc:200-251 (FP32, on the GPU); an empty vector on error
    // (the message goes to the error stream).
    std::vector<REAL> solve(const glu_analysis *an, const std::vector<REAL> &rhs);

    // ABFT (src/symbolic.cc:163-197; the reference compiles it under
    // GLU_DEBUG, this build always): CCA = column sums of val before the
    // factorization, then the checksum rebuilt from the factors in val,
    // compared column by column with tolerance 1e-5.  Both run on the GPU
    // (glu_abft_sums / glu_abft_check) and print the reference's lines.
    std::vector<REAL> CCA;
    void ABFTCalculateCCA();
    void ABFTCheckResult();

    Symbolic_Matrix(unsigned n, std::ostream &out, std::ostream &err)
        : n(n), nnz(0), num_lev(0), m_out(out), m_err(err) {}
    Symbolic_Matrix(std::ostream &out, std::ostream &err)
        : n(0), nnz(0), num_lev(0), m_out(out), m_err(err) {}

private:
    std::ostream &m_out;
    std::ostream &m_err;
};

#endif
