// C++ host mirror of the reference's GLU interface: Symbolic_Matrix methods,
// LUonDevice, preprocess/preprocess_from_arrays and the triplet reader,
// implemented over this library's host analysis and C-ABI.
//
// Reference counterparts: src/symbolic.cc:11-151,200-251 (Symbolic_Matrix),
// src/numeric.cu:347-639 (LUonDevice), src/preprocess/preprocess.c:102-192
// and preprocess_arrays.c:18-98 (preprocess*), and
// src/nicslu/util/nicslu_readfile.c:98-228 (NicsLU_ReadTripletColumnToSparse).
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "host/glu_host.h"
#include "numeric.h"
#include "preprocess.h"
#include "symbolic.h"

namespace {

glu::Symbolic view_of(const Symbolic_Matrix &m)
{
    glu::Symbolic s;
    s.n = m.n;
    s.nnz = m.nnz;
    s.sym_c_ptr = m.sym_c_ptr;
    s.sym_r_idx = m.sym_r_idx;
    s.l_col_ptr = m.l_col_ptr;
    s.csr_r_ptr = m.csr_r_ptr;
    s.csr_c_idx = m.csr_c_idx;
    s.csr_diag_ptr = m.csr_diag_ptr;
    return s;
}

glu_pattern pattern_of(const Symbolic_Matrix &m)
{
    glu_pattern p{};
    p.n = m.n;
    p.nnz = m.nnz;
    p.num_lev = m.num_lev;
    p.sym_c_ptr = m.sym_c_ptr.data();
    p.sym_r_idx = m.sym_r_idx.data();
    p.l_col_ptr = m.l_col_ptr.data();
    p.csr_r_ptr = m.csr_r_ptr.data();
    p.csr_c_idx = m.csr_c_idx.data();
    p.csr_diag_ptr = m.csr_diag_ptr.data();
    p.level_ptr = m.level_ptr.data();
    p.level_idx = m.level_idx.data();
    return p;
}

}  // namespace

void Symbolic_Matrix::fill_in(unsigned *ai, unsigned *ap)
{
    glu::Symbolic s;
    glu::symbolic_fill_in(n, ai, ap, s);
    sym_c_ptr.swap(s.sym_c_ptr);
    sym_r_idx.swap(s.sym_r_idx);
    l_col_ptr.swap(s.l_col_ptr);
    nnz = s.nnz;
    m_out << "Symbolic nonzero: " << nnz << std::endl;
}

void Symbolic_Matrix::csr()
{
    glu::Symbolic s = view_of(*this);
    glu::symbolic_csr(s);
    csr_r_ptr.swap(s.csr_r_ptr);
    csr_c_idx.swap(s.csr_c_idx);
    csr_diag_ptr.swap(s.csr_diag_ptr);
}

void Symbolic_Matrix::predictLU(unsigned *ai, unsigned *ap, double *ax)
{
    glu::Symbolic s = view_of(*this);
    glu::predict_lu(ai, ap, ax, s, nullptr);
    val.swap(s.val);
}

void Symbolic_Matrix::leveling()
{
    glu::Symbolic s = view_of(*this);
    glu::symbolic_leveling(s);
    num_lev = s.num_lev;
    level_ptr.swap(s.level_ptr);
    level_idx.swap(s.level_idx);
    m_out << "Number of levels: " << num_lev << std::endl;
}

std::vector<REAL> Symbolic_Matrix::solve(const glu_analysis *an, const std::vector<REAL> &rhs)
{
    std::vector<REAL> x(n);
    if (n == 0) return x;
    if (!an || rhs.size() != n || val.size() != nnz) {
        m_err << "solve: analysis missing or size mismatch" << std::endl;
        return {};
    }
    glu_plan *plan = nullptr;
    if (glu_plan_create_from_analysis(an, nullptr, &plan) != GLU_OK) {
        m_err << "solve: " << glu_last_error() << std::endl;
        return {};
    }
    float *d_lu = nullptr, *d_b = nullptr;
    bool ok = hipMalloc(&d_lu, sizeof(float) * nnz) == hipSuccess && hipMalloc(&d_b, sizeof(float) * n) == hipSuccess;
    ok = ok && hipMemcpy(d_lu, val.data(), sizeof(float) * nnz, hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpy(d_b, rhs.data(), sizeof(float) * n, hipMemcpyHostToDevice) == hipSuccess;
    if (ok && glu_solve(plan, d_lu, d_b, d_b, 0, nullptr) != GLU_OK) {
        m_err << "solve: " << glu_last_error() << std::endl;
        ok = false;
    }
    ok = ok && hipMemcpy(x.data(), d_b, sizeof(float) * n, hipMemcpyDeviceToHost) == hipSuccess;
    if (d_lu) (void)hipFree(d_lu);
    if (d_b) (void)hipFree(d_b);
    glu_plan_destroy(plan);
    if (!ok) {
        m_err << "solve: device error" << std::endl;
        return {};
    }
    return x;
}

// Device ABFT of one value set: CCA of val (want_cca) or the check of the
// factors in val against CCA (first failing column, or -1).  Returns false
// on a device error (the message goes to err).
static bool abft_on_device(Symbolic_Matrix &A, bool want_cca, std::vector<REAL> &cca, int32_t &first,
                           std::ostream &err)
{
    glu_pattern pat = pattern_of(A);
    glu_plan *plan = nullptr;
    if (glu_plan_create(&pat, nullptr, &plan) != GLU_OK) {
        err << "ABFT: " << glu_last_error() << std::endl;
        return false;
    }
    float *d_val = nullptr, *d_cca = nullptr, *d_mx = nullptr;
    int32_t *d_first = nullptr;
    const size_t n = A.n, nnz = A.nnz;
    bool ok = hipMalloc(&d_val, sizeof(float) * nnz) == hipSuccess &&
              hipMalloc(&d_cca, sizeof(float) * n) == hipSuccess && hipMalloc(&d_mx, sizeof(float)) == hipSuccess &&
              hipMalloc(&d_first, sizeof(int32_t)) == hipSuccess;
    ok = ok && hipMemcpy(d_val, A.val.data(), sizeof(float) * nnz, hipMemcpyHostToDevice) == hipSuccess;
    if (ok && want_cca) {
        ok = glu_abft_sums(plan, 1, d_val, d_cca, nullptr) == GLU_OK &&
             hipMemcpy(cca.data(), d_cca, sizeof(float) * n, hipMemcpyDeviceToHost) == hipSuccess;
    } else if (ok) {
        ok = hipMemcpy(d_cca, cca.data(), sizeof(float) * n, hipMemcpyHostToDevice) == hipSuccess &&
             glu_abft_check(plan, 1, d_val, d_cca, 1e-5f, d_first, d_mx, nullptr) == GLU_OK &&
             hipMemcpy(&first, d_first, sizeof(int32_t), hipMemcpyDeviceToHost) == hipSuccess;
    }
    if (!ok) err << "ABFT: " << glu_last_error() << std::endl;
    (void)hipFree(d_val);
    (void)hipFree(d_cca);
    (void)hipFree(d_mx);
    (void)hipFree(d_first);
    glu_plan_destroy(plan);
    return ok;
}

void Symbolic_Matrix::ABFTCalculateCCA()
{
    CCA.assign(n, 0.f);
    int32_t first = -1;
    if (n && nnz && val.size() == nnz) (void)abft_on_device(*this, true, CCA, first, m_err);
}

void Symbolic_Matrix::ABFTCheckResult()
{
    if (n == 0 || nnz == 0 || val.size() != nnz || CCA.size() != n) {
        m_err << "ABFT: no checksums (call ABFTCalculateCCA before the factorization)" << std::endl;
        return;
    }
    int32_t first = -1;
    if (!abft_on_device(*this, false, CCA, first, m_err)) return;
    if (first >= 0) {
        // the failing column's rebuilt checksum, for the message
        const unsigned i = (unsigned)first;
        float s = 0.f;
        for (unsigned j = sym_c_ptr[i]; j <= l_col_ptr[i]; ++j) {
            const unsigned r = sym_r_idx[j];
            float cl = 1.f;
            for (unsigned q = l_col_ptr[r] + 1; q < sym_c_ptr[r + 1]; ++q) cl += val[q];
            s += cl * val[j];
        }
        m_err << "Column " << i << ": CCA = " << CCA[i] << ", CCA_ABFT[i] = " << s << std::endl;
        m_err << "More to come..." << std::endl;
        return;
    }
    m_out << "Results passed ABFT check." << std::endl;
}

void LUonDevice(Symbolic_Matrix &A_sym, std::ostream &out, std::ostream &err, bool PERTURB)
{
    if (A_sym.n == 0 || A_sym.nnz == 0) {
        err << "Matrix is empty; skipping GPU factorization." << std::endl;
        return;
    }
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        err << "No HIP-capable GPU detected." << std::endl;
        return;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess || hipSetDevice(0) != hipSuccess) {
        err << "hipSetDevice failed" << std::endl;
        return;
    }
    out << "Device 0: " << prop.name << " has been selected." << std::endl;
    if (PERTURB) {
        float norm_a = 0;
        for (unsigned i = 0; i < A_sym.n; ++i) {
            float tmp = 0;
            for (unsigned j = A_sym.sym_c_ptr[i]; j < A_sym.sym_c_ptr[i + 1]; ++j) tmp += std::fabs(A_sym.val[j]);
            if (norm_a < tmp) norm_a = tmp;
        }
        out << "Gaussian elimination with static pivoting (GESP)..." << std::endl;
        out << "1-Norm of A matrix is " << norm_a << ", Perturbation value is " << 3.45e-4f * norm_a << std::endl;
    }
    hipEvent_t t0, t1;
    (void)hipEventCreate(&t0);
    (void)hipEventCreate(&t1);
    (void)hipEventRecord(t0, nullptr);
    glu_pattern pat = pattern_of(A_sym);
    glu_plan *plan = nullptr;
    if (glu_plan_create(&pat, nullptr, &plan) != GLU_OK) {
        err << "glu_plan_create failed: " << glu_last_error() << std::endl;
    } else {
        if (glu_factor_host(plan, A_sym.val.data(), PERTURB ? 1 : 0) != GLU_OK)
            err << "glu_factor_host failed: " << glu_last_error() << std::endl;
        glu_plan_destroy(plan);
    }
    (void)hipEventRecord(t1, nullptr);
    (void)hipEventSynchronize(t1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, t0, t1);
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    out << "Total GPU time: " << ms << " ms" << std::endl;
}

extern "C" int preprocess_from_arrays(unsigned int n, unsigned int nnz, const double *ax_in,
                                      const unsigned int *ai_in, const unsigned int *ap_in, glu_analysis **an,
                                      double **ax_out, unsigned int **ai_out, unsigned int **ap_out)
{
    if (!ax_in || !ai_in || !ap_in || !an || !ax_out || !ai_out || !ap_out) return -1;
    if (n == 0 || nnz == 0) return -1;
    *an = nullptr;
    *ax_out = nullptr;
    *ai_out = nullptr;
    *ap_out = nullptr;
    glu_analysis *a = nullptr;
    if (glu_analyze(n, nnz, ax_in, ai_in, ap_in, &a) != GLU_OK) {
        std::fprintf(stderr, "preprocess_from_arrays: %s\n", glu_last_error());
        return -1;
    }
    double *ax = static_cast<double *>(std::malloc(sizeof(double) * nnz));
    unsigned *ai = static_cast<unsigned *>(std::malloc(sizeof(unsigned) * nnz));
    unsigned *ap = static_cast<unsigned *>(std::malloc(sizeof(unsigned) * (n + 1)));
    bool ok = ax && ai && ap && glu_analysis_get(a, GLU_F_DUMP_AX, ax, sizeof(double) * nnz) == GLU_OK &&
              glu_analysis_get(a, GLU_F_DUMP_AI, ai, sizeof(unsigned) * nnz) == GLU_OK &&
              glu_analysis_get(a, GLU_F_DUMP_AP, ap, sizeof(unsigned) * (n + 1)) == GLU_OK;
    if (!ok) {
        std::free(ax);
        std::free(ai);
        std::free(ap);
        glu_analysis_destroy(a);
        return -1;
    }
    *an = a;
    *ax_out = ax;
    *ai_out = ai;
    *ap_out = ap;
    return 0;
}

extern "C" int glu_read_triplet_column(const char *file, unsigned int *n, unsigned int *nnz, double **ax,
                                       unsigned int **ai, unsigned int **ap)
{
    if (!file || !n || !nnz || !ax || !ai || !ap) return -2;
    *ax = nullptr;
    *ai = nullptr;
    *ap = nullptr;
    *n = *nnz = 0;
    FILE *f = std::fopen(file, "r");
    if (!f) return -4;
    char line[4096];
    unsigned m = 0, cols = 0, cnt = 0;
    bool header = false;
    while (std::fgets(line, sizeof line, f)) {
        if (line[0] == '%') continue;
        if (std::sscanf(line, "%u %u %u", &m, &cols, &cnt) == 3) header = true;
        break;
    }
    if (!header || m != cols) {
        std::fclose(f);
        return -7;
    }
    std::vector<unsigned> col(cnt), row(cnt);
    std::vector<double> v(cnt);
    for (unsigned k = 0; k < cnt; ++k) {
        if (std::fscanf(f, "%u %u %lf", &row[k], &col[k], &v[k]) != 3 || row[k] == 0 || col[k] == 0 ||
            row[k] > m || col[k] > m || (k > 0 && col[k] - 1 < col[k - 1])) {
            std::fclose(f);
            return -7;  // also refuses triplets that are not column-sorted
        }
        --row[k];
        --col[k];
    }
    std::fclose(f);
    *ax = static_cast<double *>(std::malloc(sizeof(double) * (cnt ? cnt : 1)));
    *ai = static_cast<unsigned *>(std::malloc(sizeof(unsigned) * (cnt ? cnt : 1)));
    *ap = static_cast<unsigned *>(std::calloc(m + 1, sizeof(unsigned)));
    if (!*ax || !*ai || !*ap) {
        std::free(*ax);
        std::free(*ai);
        std::free(*ap);
        *ax = nullptr;
        *ai = nullptr;
        *ap = nullptr;
        return -3;
    }
    for (unsigned k = 0; k < cnt; ++k) {
        (*ax)[k] = v[k];
        (*ai)[k] = row[k];
        ++(*ap)[col[k] + 1];
    }
    for (unsigned c = 0; c < m; ++c) (*ap)[c + 1] += (*ap)[c];
    *n = m;
    *nnz = cnt;
    return 0;
}

extern "C" int preprocess(const char *matrixName, glu_analysis **an, double **ax, unsigned int **ai,
                          unsigned int **ap)
{
    if (!matrixName || !an || !ax || !ai || !ap) return -1;
    unsigned n = 0, nnz = 0;
    double *rx = nullptr;
    unsigned *ri = nullptr, *rp = nullptr;
    std::printf("Reading matrix...\n");
    const int rc = glu_read_triplet_column(matrixName, &n, &nnz, &rx, &ri, &rp);
    if (rc == -4) {
        std::printf("File cannot open\n");
        return -1;
    }
    if (rc != 0) {
        std::printf("Read invalid matrix\n");
        return -1;
    }
    std::printf("Preprocessing matrix...\n");
    const int ret = preprocess_from_arrays(n, nnz, rx, ri, rp, an, ax, ai, ap);
    std::free(rx);
    std::free(ri);
    std::free(rp);
    return ret;
}
