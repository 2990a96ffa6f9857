// Test infrastructure only: a C-ABI driver around the reference's own host
// code (NICSLU analysis, preprocess.c's my_DumpA, symbolic.cc), compiled
// from the reference sources into oracle/_ref/ by oracle/Makefile.
//
// It is the checker for the product's host pipeline (MC64, AMD, permute,
// scale, fill-in, CSR, predictLU, leveling, FP32 solve), and the NICSLU
// CPU baseline timer.  Nothing in the product links or loads it.
//
// Call sequence mirrored: src/preprocess/preprocess_arrays.c:18-98 (with
// the 6-argument NicsLU_CreateMatrix of nicslu.h:132-139), then
// src/glu_binding.cpp:105-109 (fill_in, csr, predictLU, leveling), then
// Symbolic_Matrix::solve (src/symbolic.cc:200-251).

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <sstream>
#include <string>
#include <vector>

extern "C" {
#include "nicslu.h"
#include "nicslu_internal.h"
#include "preprocess.h"
}
#include "symbolic.h"

struct RefHandle {
    SNicsLU nicslu;
    bool nicslu_live = false;
    std::ostringstream out, err;
    Symbolic_Matrix *sym = nullptr;
    std::vector<double> dump_ax;
    std::vector<unsigned> dump_ai, dump_ap;
};

extern "C" {

// Returns 0 on success, otherwise the NICSLU / preprocess error code.
int ref_analyze(unsigned n, unsigned nnz, const double *ax, const unsigned *ai,
                const unsigned *ap, void **out_handle)
{
    *out_handle = nullptr;
    RefHandle *h = new RefHandle();
    int ret = NicsLU_Initialize(&h->nicslu);
    if (ret != NICS_OK) { delete h; return ret; }
    h->nicslu_live = true;
    // NicsLU_CreateMatrix copies its inputs (nicslu_matrix.c:44-46).
    ret = NicsLU_CreateMatrix(&h->nicslu, n, nnz, const_cast<double *>(ax),
                              const_cast<unsigned *>(ai), const_cast<unsigned *>(ap));
    if (ret != NICS_OK) { NicsLU_Destroy(&h->nicslu); delete h; return ret; }
    h->nicslu.cfgi[0] = 1;
    h->nicslu.cfgf[1] = 0;
    ret = NicsLU_Analyze(&h->nicslu);
    if (ret != NICS_OK) { NicsLU_Destroy(&h->nicslu); delete h; return ret; }

    double *dax = nullptr;
    unsigned *dai = nullptr, *dap = nullptr;
    ret = my_DumpA(&h->nicslu, &dax, &dai, &dap);
    if (ret != 0) { NicsLU_Destroy(&h->nicslu); delete h; return ret; }
    h->dump_ax.assign(dax, dax + nnz);
    h->dump_ai.assign(dai, dai + nnz);
    h->dump_ap.assign(dap, dap + n + 1);

    h->sym = new Symbolic_Matrix(n, h->out, h->err);
    h->sym->fill_in(dai, dap);
    h->sym->csr();
    h->sym->predictLU(dai, dap, dax);
    h->sym->leveling();
    free(dax);
    free(dai);
    free(dap);
    *out_handle = h;
    return 0;
}

void ref_free(void *hp)
{
    RefHandle *h = static_cast<RefHandle *>(hp);
    if (!h) return;
    delete h->sym;
    if (h->nicslu_live) NicsLU_Destroy(&h->nicslu);
    delete h;
}

// sizes: [0]=n [1]=nnz(A) [2]=nnz(L+U) [3]=num_lev
void ref_sizes(void *hp, long long *sizes)
{
    RefHandle *h = static_cast<RefHandle *>(hp);
    sizes[0] = h->sym->n;
    sizes[1] = h->nicslu.nnz;
    sizes[2] = h->sym->nnz;
    sizes[3] = h->sym->num_lev;
}

// Perms/scales: row_perm, row_perm_inv, col_perm, col_perm_inv (n each),
// row_scale, col_scale_perm (n each).
void ref_perms(void *hp, unsigned *rp, unsigned *rpi, unsigned *cp, unsigned *cpi,
               double *row_scale, double *col_scale_perm)
{
    RefHandle *h = static_cast<RefHandle *>(hp);
    unsigned n = h->nicslu.n;
    memcpy(rp, h->nicslu.row_perm, n * sizeof(unsigned));
    memcpy(rpi, h->nicslu.row_perm_inv, n * sizeof(unsigned));
    memcpy(cp, h->nicslu.col_perm, n * sizeof(unsigned));
    memcpy(cpi, h->nicslu.col_perm_inv, n * sizeof(unsigned));
    memcpy(row_scale, h->nicslu.row_scale, n * sizeof(double));
    memcpy(col_scale_perm, h->nicslu.col_scale_perm, n * sizeof(double));
}

void ref_dump(void *hp, double *ax, unsigned *ai, unsigned *ap)
{
    RefHandle *h = static_cast<RefHandle *>(hp);
    memcpy(ax, h->dump_ax.data(), h->dump_ax.size() * sizeof(double));
    memcpy(ai, h->dump_ai.data(), h->dump_ai.size() * sizeof(unsigned));
    memcpy(ap, h->dump_ap.data(), h->dump_ap.size() * sizeof(unsigned));
}

void ref_symbolic(void *hp, unsigned *sym_c_ptr, unsigned *sym_r_idx, unsigned *l_col_ptr,
                  unsigned *csr_r_ptr, unsigned *csr_c_idx, unsigned *csr_diag_ptr,
                  float *val, int *level_ptr, int *level_idx)
{
    RefHandle *h = static_cast<RefHandle *>(hp);
    Symbolic_Matrix *s = h->sym;
    std::copy(s->sym_c_ptr.begin(), s->sym_c_ptr.end(), sym_c_ptr);
    std::copy(s->sym_r_idx.begin(), s->sym_r_idx.end(), sym_r_idx);
    std::copy(s->l_col_ptr.begin(), s->l_col_ptr.end(), l_col_ptr);
    std::copy(s->csr_r_ptr.begin(), s->csr_r_ptr.end(), csr_r_ptr);
    std::copy(s->csr_c_idx.begin(), s->csr_c_idx.end(), csr_c_idx);
    std::copy(s->csr_diag_ptr.begin(), s->csr_diag_ptr.end(), csr_diag_ptr);
    std::copy(s->val.begin(), s->val.end(), val);
    std::copy(s->level_ptr.begin(), s->level_ptr.end(), level_ptr);
    std::copy(s->level_idx.begin(), s->level_idx.end(), level_idx);
}

// Reference FP32 host solve (src/symbolic.cc:200-251) with caller-supplied
// factors `lu` (L\U in the symbolic pattern).
void ref_solve(void *hp, const float *lu, const float *rhs, float *x)
{
    RefHandle *h = static_cast<RefHandle *>(hp);
    Symbolic_Matrix *s = h->sym;
    std::vector<float> saved(s->val);
    std::copy(lu, lu + s->nnz, s->val.begin());
    std::vector<float> r(rhs, rhs + s->n);
    std::vector<float> xv = s->solve(&h->nicslu, r);
    std::copy(xv.begin(), xv.end(), x);
    s->val.swap(saved);
}

// The reference's ABFT (src/symbolic.cc:163-197, compiled with GLU_DEBUG=1):
// ABFTCalculateCCA on val0 (the values before the factorization), then
// ABFTCheckResult on the factors lu.  cca[n] receives CCA.  Returns the
// first failing column the check reports, -1 if it reports a pass, -2 if
// its message is neither.
int ref_abft(void *hp, const float *val0, const float *lu, float *cca)
{
    RefHandle *h = static_cast<RefHandle *>(hp);
    Symbolic_Matrix *s = h->sym;
    std::vector<float> saved(s->val);
    std::copy(val0, val0 + s->nnz, s->val.begin());
    s->ABFTCalculateCCA();
    std::copy(s->CCA.begin(), s->CCA.end(), cca);
    std::copy(lu, lu + s->nnz, s->val.begin());
    h->out.str("");
    h->err.str("");
    s->ABFTCheckResult();
    s->val.swap(saved);
    const std::string e = h->err.str(), o = h->out.str();
    unsigned col = 0;
    if (sscanf(e.c_str(), "Column %u:", &col) == 1) return (int)col;
    return o.find("passed ABFT check") != std::string::npos ? -1 : -2;
}

// The reference's refactor values for a new value set ax0 (caller CSC
// order): NICSLU's own _I_NicsLU_MC64ScaleForRefact + _I_NicsLU_Scale, as
// NicsLU_ReFactorize applies them (nicslu_refact.c:74-77), then my_DumpA
// and Symbolic_Matrix::predictLU on the frozen pattern.  dump_ax receives
// the dumped FP64 values (nnz(A)), val the FP32 values (nnz(L+U)).
// Returns 0 or the error code of the step that failed.
int ref_refact_values(void *hp, const double *ax0, double *dump_ax, float *val)
{
    RefHandle *h = static_cast<RefHandle *>(hp);
    const unsigned nnz = (unsigned)h->nicslu.nnz;
    std::vector<double> a0(ax0, ax0 + nnz);
    _I_NicsLU_MC64ScaleForRefact(&h->nicslu, a0.data());
    int ret = _I_NicsLU_Scale(&h->nicslu);
    if (ret != NICS_OK) return ret;
    double *dax = nullptr;
    unsigned *dai = nullptr, *dap = nullptr;
    ret = my_DumpA(&h->nicslu, &dax, &dai, &dap);
    if (ret != 0) return ret;
    std::copy(dax, dax + nnz, dump_ax);
    Symbolic_Matrix *s = h->sym;
    std::vector<float> saved;
    saved.swap(s->val);   // predictLU appends to val
    s->predictLU(dai, dap, dax);
    std::copy(s->val.begin(), s->val.end(), val);
    s->val.swap(saved);
    free(dax);
    free(dai);
    free(dap);
    return 0;
}

static double now_s()
{
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

// NICSLU CPU baseline (FP64, its own pivoting): Analyze, Factorize, then
// `reps` ReFactorize calls (multi-threaded when threads > 1) on the value
// sets vals[r*nnz ...] and one Solve.  times[0..3] = analyze s, factorize s,
// mean refactor s (wall clock), solve s.  Returns 0 or a NICSLU code.
int ref_nicslu_time(unsigned n, unsigned nnz, const double *ax, const unsigned *ai,
                    const unsigned *ap, const double *vals, int reps, int threads,
                    double *times, double *x_out)
{
    SNicsLU nl;
    int ret = NicsLU_Initialize(&nl);
    if (ret != NICS_OK) return ret;
    ret = NicsLU_CreateMatrix(&nl, n, nnz, const_cast<double *>(ax),
                              const_cast<unsigned *>(ai), const_cast<unsigned *>(ap));
    if (ret != NICS_OK) { NicsLU_Destroy(&nl); return ret; }
    nl.cfgi[0] = 1;  // CSC input, as the GLU path sets it
    double t0 = now_s();
    ret = NicsLU_Analyze(&nl);
    double t1 = now_s();
    if (ret != NICS_OK) { NicsLU_Destroy(&nl); return ret; }
    ret = NicsLU_Factorize(&nl);
    double t2 = now_s();
    if (ret != NICS_OK) { NicsLU_Destroy(&nl); return ret; }
    bool mt = threads > 1;
    if (mt) {
        ret = NicsLU_CreateScheduler(&nl);
        if (ret < 0) { NicsLU_Destroy(&nl); return ret; }
        ret = NicsLU_CreateThreads(&nl, threads, TRUE);
        if (ret != NICS_OK) { NicsLU_Destroy(&nl); return ret; }
        NicsLU_BindThreads(&nl, FALSE);
    }
    double t3 = now_s();
    for (int r = 0; r < reps; ++r) {
        double *v = const_cast<double *>(vals + (size_t)r * nnz);
        ret = mt ? NicsLU_ReFactorize_MT(&nl, v) : NicsLU_ReFactorize(&nl, v);
        if (ret != NICS_OK) { NicsLU_Destroy(&nl); return ret; }
    }
    double t4 = now_s();
    std::vector<double> b(n, 1.0);
    NicsLU_Solve(&nl, b.data());
    double t5 = now_s();
    if (x_out) memcpy(x_out, b.data(), n * sizeof(double));
    times[0] = t1 - t0;
    times[1] = t2 - t1;
    times[2] = reps > 0 ? (t4 - t3) / reps : 0.0;
    times[3] = t5 - t4;
    NicsLU_Destroy(&nl);
    return 0;
}

}  // extern "C"
