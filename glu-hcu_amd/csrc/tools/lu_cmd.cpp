// lu_cmd: the command-line front end of the reference (src/lu_cmd.cpp:25-138)
// on this build's C++ interface.  ./lu_cmd -i file.mtx [-p] [-a]
// Reads a column-sorted triplet file, preprocesses (MC64 + AMD), runs the
// symbolic steps with their timings, factors on the GPU (LUonDevice),
// solves A x = 1 and writes x to x.dat, one value per line.  -a runs the
// ABFT checksums around the factorization, as the reference's GLU_DEBUG
// build does (src/lu_cmd.cpp:110-120).
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <vector>

#include "numeric.h"
#include "preprocess.h"
#include "symbolic.h"

using Clock = std::chrono::steady_clock;

static void help_message()
{
    std::cout << "\nGLU program (MI355X numeric phase)\n"
              << "Usage: ./lu_cmd -i inputfile\n"
              << "Additional usage: ./lu_cmd -i inputfile -p\n"
              << "-p to enable perturbation\n"
              << "-a to check the factors with ABFT column checksums" << std::endl;
}

static double ms_since(Clock::time_point t)
{
    return std::chrono::duration<double, std::milli>(Clock::now() - t).count();
}

int main(int argc, char **argv)
{
    const char *matrixName = nullptr;
    bool PERTURB = false, ABFT = false;
    if (argc < 3) {
        help_message();
        return -1;
    }
    for (int i = 1; i < argc;) {
        if (std::strcmp(argv[i], "-i") == 0) {
            if (i + 1 >= argc) {
                help_message();
                return -1;
            }
            matrixName = argv[i + 1];
            i += 2;
        } else if (std::strcmp(argv[i], "-p") == 0) {
            PERTURB = true;
            i += 1;
        } else if (std::strcmp(argv[i], "-a") == 0) {
            ABFT = true;
            i += 1;
        } else {
            help_message();
            return -1;
        }
    }
    if (!matrixName) {
        help_message();
        return -1;
    }
    glu_analysis *an = nullptr;
    double *ax = nullptr;
    unsigned *ai = nullptr, *ap = nullptr;
    if (preprocess(matrixName, &an, &ax, &ai, &ap)) {
        std::cerr << "Matrix preprocessing failed." << std::endl;
        return -1;
    }
    int64_t info[8];
    glu_analysis_info(an, info);
    const unsigned n = (unsigned)info[0];
    std::cout << "Matrix Row: " << n << std::endl;
    std::cout << "Original nonzero: " << info[1] << std::endl;

    Symbolic_Matrix A_sym(n, std::cout, std::cerr);
    auto t = Clock::now();
    A_sym.fill_in(ai, ap);
    std::cout << "Symbolic time: " << ms_since(t) << " ms" << std::endl;
    t = Clock::now();
    A_sym.csr();
    std::cout << "CSR time: " << ms_since(t) << " ms" << std::endl;
    t = Clock::now();
    A_sym.predictLU(ai, ap, ax);
    std::cout << "PredictLU time: " << ms_since(t) << " ms" << std::endl;
    t = Clock::now();
    A_sym.leveling();
    std::cout << "Leveling time: " << ms_since(t) << " ms" << std::endl;

    if (ABFT) A_sym.ABFTCalculateCCA();
    LUonDevice(A_sym, std::cout, std::cerr, PERTURB);
    if (ABFT) A_sym.ABFTCheckResult();

    std::vector<REAL> b(n, 1.f);
    std::vector<REAL> x = A_sym.solve(an, b);
    {
        std::ofstream x_f("x.dat");
        for (double xx : x) x_f << xx << '\n';
    }
    glu_analysis_destroy(an);
    std::free(ax);
    std::free(ai);
    std::free(ap);
    return x.size() == n ? 0 : 1;
}
