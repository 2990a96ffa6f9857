// Plan construction (host) and the C-ABI of include/glu_hip.h.
//
// A plan uploads one sparsity pattern once: the Symbolic_Matrix arrays that
// LUonDevice copies on every call (src/numeric.cu:435-458), plus the tables
// this build adds — the CSR->CSC map, the per-update scatter map, work
// units and level schedules for the factorization and the four triangular
// sweeps, and (from an analysis) the permutation/scales and the structure of
// the original A for value refresh and FP64 refinement.
//
// Every index array is validated on the host before upload (pointer
// monotonicity, ranges, sorted unique rows, diagonal positions, CSR/CSC
// agreement, level permutation, fill closure); the kernels index without
// bounds checks and rely on it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <exception>
#include <memory>
#include <mutex>
#include <string>
#include <atomic>
#include <system_error>
#include <thread>
#include <vector>

#include "../../include/glu_hip.h"
#include "glu_dev.h"
#include "host/glu_host.h"

namespace glu {
namespace launch {
void pert(const FactorDev &F, const float *val, unsigned int *norm_bits, float *d_pert, hipStream_t st);
void factor_grid(const FactorDev &F, float *val, const float *d_pert, int lvl, uint32_t items, int perturb,
                 int threads, hipStream_t st);
void factor_fused(const FactorDev &F, float *val, const float *d_pert, int l0, int l1, int perturb,
                  int threads, hipStream_t st);
void factor_divide(const FactorDev &F, float *val, const float *d_pert, uint32_t ndiv, int perturb, hipStream_t st);
void tri_grid(const TriDev &T, const float *lu, float *x, int lvl, uint32_t rows, hipStream_t st);
void tri_fused(const TriDev &T, const float *lu, float *x, int l0, int l1, int threads, hipStream_t st);
void perm_scale_f32(const float *in, float *out, const uint32_t *P, const double *S, int by_perm, uint32_t n,
                    hipStream_t st);
void perm_scale_f64(const double *in, float *out, const uint32_t *P, const double *S, int by_perm,
                    uint32_t n, hipStream_t st);
void update_x(const float *t, double *x, const uint32_t *P, const double *S, int by_perm, uint32_t n,
              hipStream_t st);
void refresh(const RefineDev &R, const double *ax, float *lu, uint32_t nnz_lu, int batch, hipStream_t st);
void residual(const RefineDev &R, const double *ax, const double *b, const double *x, double *r, uint32_t n,
              int transpose, int with_anorm, unsigned long long *norms, uint32_t lanes, hipStream_t st);
size_t batch_lds_bytes(int mode, uint32_t n);
void batch_same(const MatDesc *desc, int batch, const double *ax, const double *b, double *x, float *lu,
                float *lucsr, float *t32, double *r64, double *norms, double *kry, int32_t *iters, double *eta,
                double tol, int maxit, int perturb, int threads, int mode, size_t fact_lds, size_t ref_lds,
                int l0, hipStream_t st);
void batch_factor(const MatDesc *desc, int batch, const double *ax, float *lu, double *norms, int perturb,
                  int threads, size_t fact_lds, int l0, hipStream_t st);
void il_values(const RefineDev &R, const FactorDev &F, const double *ax, float *S, unsigned int *bits,
               double *pertd, uint32_t nnz, uint32_t n, int batch, int perturb, hipStream_t st);
void factor_grid_il(const FactorDev &F, float *S, const double *pertd, int lvl, uint32_t waves, int batch,
                    int perturb, hipStream_t st);
void il_to_rows(const float *S, float *lu, uint32_t nnz, int batch, hipStream_t st);
uint32_t il_items_per_wave();
uint32_t il_group();
void abft_cca(const FactorDev &F, const float *val, float *cca, int batch, hipStream_t st);
void abft_check(const FactorDev &F, const uint32_t *row_idx, const uint32_t *diag, const float *lu,
                const float *cca, float *ccl, float tol, int32_t *first_bad, float *max_diff, int batch,
                hipStream_t st);
void batch_group(const MatDesc *descs, const void *ios, const uint32_t *order, int count, double tol,
                 int maxit, int perturb, int threads, int mode, size_t fact_lds, size_t ref_lds, hipStream_t st);
void tri_df(const TriDev &T, const float *lu, float *x, int l0, int l1, uint32_t k0, uint32_t k1, uint32_t n,
            int threads, hipStream_t st);
size_t enable_lds();
void factor_window(const FactorDev &F, int w, uint32_t nslot, float *val, const float *d_pert, int perturb,
                   int threads, hipStream_t st);
void zero_u64(unsigned long long *p, int count, hipStream_t st);
void solve_one(const MatDesc *desc, uint32_t nval, const float *lu, float *lucsr, const float *b, float *x,
               float *t32, uint32_t r0, int threads, int mode, size_t lds, hipStream_t st);
void refine_one(const MatDesc *desc, uint32_t nval, const double *ax, const double *b, double *x,
                const float *lu, float *lucsr, float *t32, double *r64, double *norms, double *kry, int32_t *iters,
                double *eta, double tol, int maxit, int threads, size_t ref_lds, int transpose, hipStream_t st);
void copy_df(const MatDesc *desc, uint32_t nval, const float *lu, float *lucsr, int transpose, hipStream_t st);
void sweeps_df(const MatDesc *desc, const float *lucsr, float *t32, uint32_t r0, int threads, size_t lds,
               int transpose, hipStream_t st);
void adjoint_values(const RefineDev &R, const double *lam, const double *x, double *grad, uint32_t n,
                    hipStream_t st);
size_t matio_size();
}  // namespace launch
}  // namespace glu

using namespace glu;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------

static thread_local std::string g_err;

static glu_status fail(glu_status s, const std::string &msg)
{
    g_err = msg;
    return s;
}

#define HIP_TRY(call)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) return fail(GLU_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

extern "C" const char *glu_last_error(void) { return g_err.c_str(); }

// ---------------------------------------------------------------------------
// analysis object
// ---------------------------------------------------------------------------

struct glu_analysis {
    Analysis an;
    Symbolic sym;
    std::vector<uint32_t> lu_from_dump;
    int64_t updates = 0;
};

static int64_t count_updates(const Symbolic &s)
{
    int64_t u = 0;
    for (uint32_t k = 0; k < s.n; ++k) {
        int64_t L = (int64_t)s.sym_c_ptr[k + 1] - s.l_col_ptr[k] - 1;
        int64_t U = (int64_t)s.csr_r_ptr[k + 1] - s.csr_diag_ptr[k] - 1;
        u += L * U;
    }
    return u;
}

extern "C" glu_status glu_analyze(uint32_t n, uint32_t nnz, const double *ax, const uint32_t *ai,
                                  const uint32_t *ap, glu_analysis **out)
{
    if (!out) return fail(GLU_ERR_ARGUMENT, "out is null");
    *out = nullptr;
    if (n == 0 || nnz == 0) return fail(GLU_ERR_EMPTY, "matrix is empty");
    if (!ax || !ai || !ap) return fail(GLU_ERR_ARGUMENT, "null array");
    std::unique_ptr<glu_analysis> a(new (std::nothrow) glu_analysis());
    if (!a) return fail(GLU_ERR_MEMORY, "out of host memory");
    std::string err;
    int rc;
    try {
        rc = analyze(n, nnz, ax, ai, ap, a->an, err);
        if (rc == kOk) {
            // Duplicates the reference's adjacency check lets through
            // (non-adjacent repeats in one column) are refused here.
            std::vector<uint32_t> stamp(n, UINT32_MAX);
            for (uint32_t c = 0; c < n && rc == kOk; ++c)
                for (uint32_t k = ap[c]; k < ap[c + 1]; ++k) {
                    if (stamp[ai[k]] == c) {
                        rc = kErrDuplicate;
                        err = "duplicate entry";
                        break;
                    }
                    stamp[ai[k]] = c;
                }
        }
        if (rc == kOk) {
            symbolic(n, a->an.ai.data(), a->an.ap.data(), a->sym);
            predict_lu(a->an.ai.data(), a->an.ap.data(), a->an.ax.data(), a->sym, &a->lu_from_dump);
            a->updates = count_updates(a->sym);
        }
    } catch (const std::bad_alloc &) {
        return fail(GLU_ERR_MEMORY, "out of host memory during analysis");
    }
    if (rc != kOk) return fail((glu_status)rc, err);
    *out = a.release();
    return GLU_OK;
}

// Independent patterns analysed on host threads (a work queue over the
// matrices); each item as glu_analyze.
extern "C" glu_status glu_analyze_many(int count, const uint32_t *n, const uint32_t *nnz,
                                       const double *const *ax, const uint32_t *const *ai,
                                       const uint32_t *const *ap, int threads, glu_analysis **out,
                                       glu_status *status)
{
    if (count <= 0 || !n || !nnz || !ax || !ai || !ap || !out || !status)
        return fail(GLU_ERR_ARGUMENT, "bad argument");
    for (int i = 0; i < count; ++i) out[i] = nullptr;
    int hw = (int)std::thread::hardware_concurrency();
    int nt = threads > 0 ? threads : (hw > 0 ? hw : 1);
    nt = std::max(1, std::min(nt, count));
    std::vector<std::string> msg(count);
    std::atomic<int> next(0);
    auto work = [&]() {
        for (int i = next.fetch_add(1); i < count; i = next.fetch_add(1)) {
            status[i] = glu_analyze(n[i], nnz[i], ax[i], ai[i], ap[i], out + i);
            if (status[i] != GLU_OK) msg[i] = glu_last_error();
        }
    };
    std::vector<std::thread> pool;
    try {
        pool.reserve(nt - 1);
        for (int t = 1; t < nt; ++t) pool.emplace_back(work);
    } catch (const std::exception &) {
        // fewer threads than asked for (none left, or no memory): the ones
        // that started share the queue with this thread
    }
    work();
    for (std::thread &t : pool) t.join();
    for (int i = 0; i < count; ++i)
        if (status[i] != GLU_OK) return fail(status[i], "matrix " + std::to_string(i) + ": " + msg[i]);
    return GLU_OK;
}

extern "C" void glu_analysis_destroy(glu_analysis *a) { delete a; }

extern "C" glu_status glu_analysis_info(const glu_analysis *a, int64_t info[8])
{
    if (!a || !info) return fail(GLU_ERR_ARGUMENT, "null argument");
    info[0] = a->sym.n;
    info[1] = a->an.nnz;
    info[2] = a->sym.nnz;
    info[3] = a->sym.num_lev;
    info[4] = a->updates;
    info[5] = info[6] = info[7] = 0;
    return GLU_OK;
}

template <typename T>
static glu_status copy_field(const std::vector<T> &v, void *dst, size_t bytes)
{
    if (bytes != v.size() * sizeof(T))
        return fail(GLU_ERR_ARGUMENT, "size mismatch: expected " + std::to_string(v.size() * sizeof(T)) +
                                          " bytes, got " + std::to_string(bytes));
    if (bytes) std::memcpy(dst, v.data(), bytes);
    return GLU_OK;
}

extern "C" glu_status glu_analysis_get(const glu_analysis *a, glu_field f, void *dst, size_t bytes)
{
    if (!a || (!dst && bytes)) return fail(GLU_ERR_ARGUMENT, "null argument");
    const Symbolic &s = a->sym;
    const Analysis &an = a->an;
    switch (f) {
    case GLU_F_SYM_C_PTR: return copy_field(s.sym_c_ptr, dst, bytes);
    case GLU_F_SYM_R_IDX: return copy_field(s.sym_r_idx, dst, bytes);
    case GLU_F_L_COL_PTR: return copy_field(s.l_col_ptr, dst, bytes);
    case GLU_F_CSR_R_PTR: return copy_field(s.csr_r_ptr, dst, bytes);
    case GLU_F_CSR_C_IDX: return copy_field(s.csr_c_idx, dst, bytes);
    case GLU_F_CSR_DIAG_PTR: return copy_field(s.csr_diag_ptr, dst, bytes);
    case GLU_F_VAL: return copy_field(s.val, dst, bytes);
    case GLU_F_LEVEL_PTR: return copy_field(s.level_ptr, dst, bytes);
    case GLU_F_LEVEL_IDX: return copy_field(s.level_idx, dst, bytes);
    case GLU_F_ROW_PERM: return copy_field(an.row_perm, dst, bytes);
    case GLU_F_ROW_PERM_INV: return copy_field(an.row_perm_inv, dst, bytes);
    case GLU_F_COL_PERM: return copy_field(an.col_perm, dst, bytes);
    case GLU_F_COL_PERM_INV: return copy_field(an.col_perm_inv, dst, bytes);
    case GLU_F_ROW_SCALE: return copy_field(an.row_scale, dst, bytes);
    case GLU_F_COL_SCALE_PERM: return copy_field(an.col_scale_perm, dst, bytes);
    case GLU_F_DUMP_AX: return copy_field(an.ax, dst, bytes);
    case GLU_F_DUMP_AI: return copy_field(an.ai, dst, bytes);
    case GLU_F_DUMP_AP: return copy_field(an.ap, dst, bytes);
    }
    return fail(GLU_ERR_ARGUMENT, "unknown field");
}

// ---------------------------------------------------------------------------
// plan
// ---------------------------------------------------------------------------

namespace {

constexpr int kGridThreads = 256;
constexpr int kDefaultFusedThreads = 1024;
constexpr int kDefaultKrylov = 2;         // FGMRES(2) cycles after the first solve
constexpr int kDefaultFuseUnits = 2048;   // lanes
constexpr uint32_t kCopyRange = 32768;    // slots (128 KB of LDS) per range of the batch value copy
constexpr int kMaxGroupSplit = 2;         // group members on the single-matrix path, at most

struct Segment {
    int l0, l1;      // levels [l0, l1)
    bool fused;
    uint32_t items;  // grid segments: items (factor) or rows (solve)
    uint32_t k0 = 0, k1 = 0;   // solve segments: dataflow tasks [k0, k1)
    int win = -1;              // factor segments: an LDS window
};

struct TriHost {
    std::vector<uint32_t> rows, lvl_ptr, lvl_long;
    std::vector<uint32_t> len;     // entries of each row
    std::vector<uint32_t> tasks;   // dataflow task starts (+ end sentinel)
    std::vector<Segment> segs;
};

// Device arena: one allocation, 256-byte aligned sub-buffers.
struct Arena {
    std::vector<std::pair<size_t, size_t>> parts;  // offset, bytes
    size_t total = 0;
    size_t add(size_t bytes)
    {
        size_t off = total;
        total += (bytes + 255) & ~size_t(255);
        parts.push_back({off, bytes});
        return off;
    }
};

}  // namespace

struct glu_plan {
    int device = 0;
    uint32_t n = 0, nnz = 0, nnz_a = 0;
    int num_lev = 0;
    uint32_t ndiv = 0;           // division items of the factorization
    int64_t updates = 0;
    int64_t units = 0;
    int fused_threads = kDefaultFusedThreads;
    int fuse_units = kDefaultFuseUnits;
    bool has_scaling = false, has_refine = false;
    uint32_t a_nlong = 0, a_nlong_t = 0;   // long rows / columns of A
    uint32_t a_kmax = 0;                   // most entries in a row or a column of A

    std::vector<Segment> fsegs;
    TriHost tri[4];          // lo, up, tup, tlo
    bool tri_dataflow = false;   // fused solve segments run barrier-free
    int solve_df_opt = 0;        // glu_plan_opts.solve_dataflow
    size_t lds_cap = 0;          // dynamic LDS per workgroup (bytes)
    uint32_t df_nval = 0;        // values of the per-matrix sweep copy
    uint32_t tdf_nval = 0;       // the same for A' x = b
    float *d_tlucsr = nullptr;   // glu_refine's A' x = b value copy
    // LDS windows allowed (from the last level); windows hold narrow steps
    // only, so small orders get one whatever the cap, and n=500k MNA gains
    // up to ~16 (refactor 14.3 -> 12.4 ms; no change beyond)
    int max_windows = 64;
    uint32_t win_values = 0;     // values one window may stage (0: what LDS holds)
    int krylov = kDefaultKrylov;  // batch refinement: FGMRES restart length (0: plain)
    int nwin = 0;                // LDS windows of the factorization
    int win_steps = 0;           // steps they cover
    uint32_t win_nslot = 0;      // the largest window's values
    uint32_t copy_range = 0;     // slots the batch kernels' value copy stages in LDS at a time
    std::vector<uint32_t> win_slots;   // values of each window
    // batched kernels: steps [0, batch_grid_lev) run as one grid launch per
    // step over an interleaved copy of the batch (grid_waves: each one's
    // waves for one group of interleaved matrices: long merges one each,
    // other items il_items_per_wave() a wave)
    int batch_grid_opt = 0;      // glu_plan_opts.batch_grid_levels
    int batch_grid_lev = 0;
    std::vector<uint32_t> grid_waves;

    // hipGraph replay of the launch sequences (use_graph): one executable
    // graph per entry point, keyed on the buffers it was captured with
    struct GraphSlot {
        hipGraphExec_t exec = nullptr;
        const void *k[3] = {nullptr, nullptr, nullptr};
        int mode = -1;
    };
    bool use_graph = false;
    hipStream_t cap_stream = nullptr;
    GraphSlot g_factor, g_solve[2];

    char *d_arena = nullptr;
    size_t arena_bytes = 0;
    FactorDev F{};
    SolveDev S{};
    RefineDev R{};
    MatDesc *d_desc = nullptr;   // lives in the arena

    // scalars and single-matrix workspaces (also in the arena)
    unsigned int *d_norm_bits = nullptr;
    float *d_pert = nullptr;
    unsigned long long *d_norms = nullptr;
    float *d_t32 = nullptr;      // n
    const uint32_t *d_rowidx = nullptr, *d_diag = nullptr;   // sym_r_idx, l_col_ptr (arena)
    float *d_abft = nullptr;     // ABFT column sums of L, n per matrix (grown on demand)
    int abft_cap = 0;
    double *d_r64 = nullptr;     // n
    float *d_val = nullptr;      // nnz (glu_factor_host)

    // batch workspaces, grown on demand
    float *d_bt32 = nullptr;
    double *d_br64 = nullptr;
    float *d_blucsr = nullptr;   // sweep value copies, df_nval per matrix
    double *d_bnorms = nullptr;  // |A|_inf, |b|_inf per matrix
    char *d_rout = nullptr;      // glu_refine's iteration count and backward error (device)
    double *d_bkry = nullptr;    // Arnoldi vectors V, Z = M V, W = A Z: (3 krylov + 1) n per matrix
    int batch_cap = 0;
    double *d_fnorms = nullptr;  // glu_batch_factor: GESP value per matrix (2 doubles each)
    float *d_il = nullptr;       // interleaved batch values (groups of 16 matrices)
    unsigned int *d_il_bits = nullptr;   // per-matrix GESP maxima, zero between calls
    int il_cap = 0;              // matrices d_il, d_il_bits and d_fnorms hold

    ~glu_plan()
    {
        if (d_arena) (void)hipFree(d_arena);
        if (d_val) (void)hipFree(d_val);
        if (d_bt32) (void)hipFree(d_bt32);
        if (d_br64) (void)hipFree(d_br64);
        if (d_blucsr) (void)hipFree(d_blucsr);
        if (d_bkry) (void)hipFree(d_bkry);
        if (d_rout) (void)hipFree(d_rout);
        if (d_tlucsr) (void)hipFree(d_tlucsr);
        if (d_abft) (void)hipFree(d_abft);
        if (d_bnorms) (void)hipFree(d_bnorms);
        if (d_fnorms) (void)hipFree(d_fnorms);
        if (d_il) (void)hipFree(d_il);
        if (d_il_bits) (void)hipFree(d_il_bits);
        for (GraphSlot *g : {&g_factor, &g_solve[0], &g_solve[1]})
            if (g->exec) (void)hipGraphExecDestroy(g->exec);
        if (cap_stream) (void)hipStreamDestroy(cap_stream);
    }
};

namespace {

// ------------------------- validation -----------------------------------

bool check_ptr(const uint32_t *p, uint32_t n, uint32_t total, std::string &err, const char *what)
{
    if (p[0] != 0 || p[n] != total) {
        err = std::string(what) + ": pointer array must start at 0 and end at the entry count";
        return false;
    }
    for (uint32_t i = 0; i < n; ++i)
        if (p[i] > p[i + 1]) {
            err = std::string(what) + ": pointer array not monotone";
            return false;
        }
    return true;
}

bool validate_pattern(const glu_pattern &P, std::string &err)
{
    const uint32_t n = P.n, nnz = P.nnz;
    if (!P.sym_c_ptr || !P.sym_r_idx || !P.l_col_ptr || !P.csr_r_ptr || !P.csr_c_idx || !P.csr_diag_ptr ||
        !P.level_ptr || !P.level_idx) {
        err = "null pattern array";
        return false;
    }
    if (nnz >= 0x80000000u) {
        err = "nnz(L+U) must be below 2^31";
        return false;
    }
    if (P.num_lev <= 0 || (uint32_t)P.num_lev > n) {
        err = "num_lev out of range";
        return false;
    }
    if (!check_ptr(P.sym_c_ptr, n, nnz, err, "sym_c_ptr") || !check_ptr(P.csr_r_ptr, n, nnz, err, "csr_r_ptr"))
        return false;
    for (uint32_t c = 0; c < n; ++c) {
        for (uint32_t p = P.sym_c_ptr[c]; p < P.sym_c_ptr[c + 1]; ++p) {
            if (P.sym_r_idx[p] >= n || (p > P.sym_c_ptr[c] && P.sym_r_idx[p] <= P.sym_r_idx[p - 1])) {
                err = "sym_r_idx out of range or not strictly ascending in column " + std::to_string(c);
                return false;
            }
        }
        const uint32_t d = P.l_col_ptr[c];
        if (d < P.sym_c_ptr[c] || d >= P.sym_c_ptr[c + 1] || P.sym_r_idx[d] != c) {
            err = "l_col_ptr does not point at the diagonal of column " + std::to_string(c);
            return false;
        }
    }
    // CSR must be the exact transpose of the CSC pattern.
    std::vector<uint32_t> cursor(P.csr_r_ptr, P.csr_r_ptr + n);
    for (uint32_t c = 0; c < n; ++c)
        for (uint32_t p = P.sym_c_ptr[c]; p < P.sym_c_ptr[c + 1]; ++p) {
            const uint32_t r = P.sym_r_idx[p];
            const uint32_t q = cursor[r]++;
            if (q >= P.csr_r_ptr[r + 1] || P.csr_c_idx[q] != c) {
                err = "CSR arrays are not the transpose of the CSC pattern";
                return false;
            }
        }
    for (uint32_t r = 0; r < n; ++r) {
        const uint32_t q = P.csr_diag_ptr[r];
        if (q < P.csr_r_ptr[r] || q >= P.csr_r_ptr[r + 1] || P.csr_c_idx[q] != r) {
            err = "csr_diag_ptr does not point at the diagonal of row " + std::to_string(r);
            return false;
        }
    }
    if (P.level_ptr[0] != 0 || (uint32_t)P.level_ptr[P.num_lev] != n) {
        err = "level_ptr must span 0..n";
        return false;
    }
    for (int l = 0; l < P.num_lev; ++l)
        if (P.level_ptr[l] > P.level_ptr[l + 1]) {
            err = "level_ptr not monotone";
            return false;
        }
    std::vector<int> level_of(n, -1);
    for (int l = 0; l < P.num_lev; ++l)
        for (int t = P.level_ptr[l]; t < P.level_ptr[l + 1]; ++t) {
            const int k = P.level_idx[t];
            if (k < 0 || (uint32_t)k >= n || level_of[k] != -1) {
                err = "level_idx is not a permutation";
                return false;
            }
            level_of[k] = l;
        }
    // Dependencies of the right-looking update must sit in earlier levels
    // (src/symbolic.cc:115-130): otherwise two columns of one level race.
    for (uint32_t i = 0; i < n; ++i) {
        for (uint32_t p = P.sym_c_ptr[i]; p < P.l_col_ptr[i]; ++p) {
            const uint32_t r = P.sym_r_idx[p];
            if (P.l_col_ptr[r] + 1 != P.sym_c_ptr[r + 1] && level_of[r] >= level_of[i]) {
                err = "levels violate a column dependency";
                return false;
            }
        }
        for (uint32_t q = P.csr_r_ptr[i]; q < P.csr_diag_ptr[i]; ++q)
            if (level_of[P.csr_c_idx[q]] >= level_of[i]) {
                err = "levels violate a row dependency";
                return false;
            }
    }
    return true;
}

// --------------------------- tables -------------------------------------

struct Tables {
    std::vector<uint32_t> upos;        // CSR position -> CSC position
    // column tasks, by level: task t has lanes 64 t .. 64 t + 63 in cln
    // ({L position, pivot position, update count <= 4, 0}); the lane's e-th
    // update reads U value cup[256 t + 64 e + lane] and targets ctgt[same]
    std::vector<uint4> cln;
    std::vector<uint32_t> ctgt, cup;
    std::vector<uint32_t> ctask_ptr;   // [num_lev + 2]
    std::vector<uint32_t> ment, ment_ptr, ment_long, msrc_ptr;
    std::vector<MergeSrc> msrc;
    std::vector<uint2> dvd;            // division items, in CSC position order
    std::vector<uint32_t> dvd_ptr;     // [num_lev + 1]: items of the columns of levels < l (built by level)
};

// Runs work(worker, i) for i in [0, count) on up to `threads` host threads
// (the calling one included), items taken in increasing order from a shared
// counter.  The first exception thrown in any worker is rethrown here after
// every thread has finished.
template <typename Work>
void parallel_items(size_t count, unsigned threads, Work work)
{
    threads = (unsigned)std::max<size_t>(1, std::min<size_t>(threads, count));
    std::atomic<size_t> next(0);
    std::exception_ptr failure;
    std::mutex fail_mu;
    auto run = [&](unsigned worker) {
        try {
            for (size_t i = next.fetch_add(1); i < count; i = next.fetch_add(1)) work(worker, i);
        } catch (...) {
            std::lock_guard<std::mutex> lock(fail_mu);
            if (!failure) failure = std::current_exception();
            next.store(count);
        }
    };
    std::vector<std::thread> pool;
    try {
        pool.reserve(threads - 1);
        for (unsigned t = 1; t < threads; ++t) pool.emplace_back(run, t);
    } catch (const std::exception &) {
        // fewer threads than asked for: the ones that started share the work
    }
    run(0);
    for (std::thread &t : pool) t.join();
    if (failure) std::rethrow_exception(failure);
}

// Host threads for building plan tables: the cores available, at most 16.
unsigned table_threads()
{
    const unsigned hw = std::thread::hardware_concurrency();
    return std::max(1u, std::min(16u, hw ? hw : 1u));
}

// One level's factorization items (the level's slice of Tables).
struct LevelItems {
    std::vector<uint4> cln;
    std::vector<uint32_t> ctgt, cup;
    std::vector<uint32_t> ment, msrc_cnt;
    uint32_t ment_long = 0;
    std::vector<MergeSrc> msrc;
    std::vector<uint2> dvd;
};

// Lane items of the factorization, level by level.  For pivot column k of
// level l, every (i in L_k, j in U_k) update targets the CSC position of
// (i,j) in column j; the fill closure of the pattern guarantees it exists
// (checked).  Targets that two or more columns of the level update become
// merge entries whose sources keep the level's column order (level_idx lists
// a level's columns in increasing order); all other updates are items.
// The update map is built per target column and the items per level, each
// on host threads; levels are then concatenated in order, so the tables do
// not depend on the thread count.
bool build_factor_tables(const glu_pattern &P, Tables &T, int64_t &updates, std::string &err)
{
    const uint32_t n = P.n, nnz = P.nnz;
    const int nl = P.num_lev;
    const unsigned threads = table_threads();
    T.upos.resize(nnz);
    std::vector<uint32_t> csc2csr(nnz);
    {
        std::vector<uint32_t> cursor(P.csr_r_ptr, P.csr_r_ptr + n);
        for (uint32_t c = 0; c < n; ++c)
            for (uint32_t p = P.sym_c_ptr[c]; p < P.sym_c_ptr[c + 1]; ++p) {
                const uint32_t q = cursor[P.sym_r_idx[p]]++;
                T.upos[q] = p;
                csc2csr[p] = q;
            }
    }
    // Scatter targets per (k, jj, ii), built per target column j (columns
    // write disjoint ranges of umap).
    std::vector<uint64_t> moff(n + 1);
    uint64_t off = 0;
    for (uint32_t k = 0; k < n; ++k) {
        moff[k] = off;
        const uint64_t L = P.sym_c_ptr[k + 1] - P.l_col_ptr[k] - 1;
        const uint64_t U = P.csr_r_ptr[k + 1] - P.csr_diag_ptr[k] - 1;
        off += L * U;
    }
    moff[n] = off;
    updates = (int64_t)off;
    std::vector<uint32_t> umap;
    try {
        umap.assign(off, 0);
    } catch (const std::bad_alloc &) {
        err = "update tables too large for host memory";
        return false;
    }
    {
        constexpr uint32_t kCols = 256;   // target columns per work item
        std::vector<std::vector<uint32_t>> stamps(threads), wheres(threads);
        std::atomic<uint32_t> bad_col(UINT32_MAX);
        parallel_items((n + kCols - 1) / kCols, threads, [&](unsigned w, size_t item) {
            std::vector<uint32_t> &stamp = stamps[w], &where = wheres[w];
            if (stamp.empty()) {
                stamp.assign(n, UINT32_MAX);
                where.assign(n, 0);
            }
            const uint32_t j1 = std::min<uint32_t>(n, (uint32_t)(item + 1) * kCols);
            for (uint32_t j = (uint32_t)item * kCols; j < j1; ++j) {
                for (uint32_t p = P.sym_c_ptr[j]; p < P.sym_c_ptr[j + 1]; ++p) {
                    stamp[P.sym_r_idx[p]] = j;
                    where[P.sym_r_idx[p]] = p;
                }
                for (uint32_t p = P.sym_c_ptr[j]; p < P.l_col_ptr[j]; ++p) {
                    const uint32_t k = P.sym_r_idx[p];
                    const uint32_t L = P.sym_c_ptr[k + 1] - P.l_col_ptr[k] - 1;
                    if (L == 0) continue;
                    const uint32_t jj = csc2csr[p] - P.csr_diag_ptr[k] - 1;
                    const uint64_t base = moff[k] + (uint64_t)jj * L;
                    const uint32_t l0 = P.l_col_ptr[k] + 1;
                    for (uint32_t ii = 0; ii < L; ++ii) {
                        const uint32_t i = P.sym_r_idx[l0 + ii];
                        if (stamp[i] != j) {
                            uint32_t none = UINT32_MAX;
                            bad_col.compare_exchange_strong(none, j);
                            return;
                        }
                        umap[base + ii] = where[i];
                    }
                }
            }
        });
        if (bad_col.load() != UINT32_MAX) {
            err = "pattern is not closed under fill (column " + std::to_string(bad_col.load()) + ")";
            return false;
        }
    }
    // Items per level, on host threads (each with its own marks over the
    // positions), heaviest levels first.
    std::vector<LevelItems> lv(nl);
    std::vector<uint64_t> lwork(nl, 0);
    for (int l = 0; l < nl; ++l)
        for (int t = P.level_ptr[l]; t < P.level_ptr[l + 1]; ++t) {
            const uint32_t k = (uint32_t)P.level_idx[t];
            lwork[l] += (moff[k + 1] - moff[k]) + 1;
        }
    std::vector<int> lorder(nl);
    for (int l = 0; l < nl; ++l) lorder[l] = l;
    std::stable_sort(lorder.begin(), lorder.end(), [&](int a, int b) { return lwork[a] > lwork[b]; });
    struct Marks {
        std::vector<uint32_t> stamp, cnt, slot, touched;
    };
    std::vector<Marks> marks(threads);
    parallel_items((size_t)nl, threads, [&](unsigned w, size_t item) {
        const int l = lorder[item];
        Marks &M = marks[w];
        if (M.stamp.empty()) {
            M.stamp.assign(nnz, UINT32_MAX);
            M.cnt.assign(nnz, 0);
            M.slot.assign(nnz, 0);
        }
        LevelItems &out = lv[l];
        auto each_update = [&](auto f) {
            for (int t = P.level_ptr[l]; t < P.level_ptr[l + 1]; ++t) {
                const uint32_t k = (uint32_t)P.level_idx[t];
                const uint32_t dpos = P.l_col_ptr[k];
                const uint32_t L = P.sym_c_ptr[k + 1] - dpos - 1;
                if (L == 0) continue;
                const uint32_t uf = P.csr_diag_ptr[k] + 1;
                const uint32_t U = P.csr_r_ptr[k + 1] - uf;
                for (uint32_t jj = 0; jj < U; ++jj) {
                    const uint32_t up = T.upos[uf + jj];
                    const uint64_t base = moff[k] + (uint64_t)jj * L;
                    for (uint32_t ii = 0; ii < L; ++ii) f(umap[base + ii], dpos + 1 + ii, up, dpos);
                }
            }
        };
        M.touched.clear();
        each_update([&](uint32_t tgt, uint32_t, uint32_t, uint32_t) {
            if (M.stamp[tgt] != (uint32_t)l) {
                M.stamp[tgt] = (uint32_t)l;
                M.cnt[tgt] = 0;
                M.touched.push_back(tgt);
            }
            ++M.cnt[tgt];
        });
        std::sort(M.touched.begin(), M.touched.end());
        // long merge entries (a wave each) first, then the rest
        std::vector<uint32_t> fillpos;
        for (int pass = 0; pass < 2; ++pass)
            for (uint32_t tgt : M.touched) {
                if (M.cnt[tgt] < 2 || (M.cnt[tgt] > kLongMerge) != (pass == 0)) continue;
                M.slot[tgt] = (uint32_t)out.ment.size();
                out.ment.push_back(tgt);
                out.msrc_cnt.push_back(M.cnt[tgt]);
                fillpos.push_back((uint32_t)out.msrc.size());
                out.msrc.resize(out.msrc.size() + M.cnt[tgt]);
                if (pass == 0) ++out.ment_long;
            }
        each_update([&](uint32_t tgt, uint32_t lpos, uint32_t up, uint32_t dpos) {
            if (M.cnt[tgt] >= 2) out.msrc[fillpos[M.slot[tgt]]++] = MergeSrc{lpos, up, dpos, 0};
        });
        // Column tasks: a lane takes one L entry A(i,k) of a pivot column
        // and up to four consecutive entries of its U list; lanes sharing a
        // U chunk are adjacent (they read the same U values), and tasks are
        // 64 lanes.  Merge targets are left to their merge entries (~0u).
        uint32_t used = 64;   // lanes used in the open task
        for (int t = P.level_ptr[l]; t < P.level_ptr[l + 1]; ++t) {
            const uint32_t k = (uint32_t)P.level_idx[t];
            const uint32_t dpos = P.l_col_ptr[k];
            const uint32_t L = P.sym_c_ptr[k + 1] - dpos - 1;
            const uint32_t U = P.csr_r_ptr[k + 1] - P.csr_diag_ptr[k] - 1;
            if (!L || !U) continue;
            const uint32_t uf = P.csr_diag_ptr[k] + 1;
            for (uint32_t c0 = 0; c0 < U; c0 += 4) {
                const uint32_t nu = std::min<uint32_t>(4, U - c0);
                for (uint32_t ii = 0; ii < L; ++ii) {
                    if (used == 64) {
                        out.cln.resize(out.cln.size() + 64, make_uint4(0u, 0u, 0u, 0u));
                        out.ctgt.resize(out.ctgt.size() + 256, UINT32_MAX);
                        out.cup.resize(out.cup.size() + 256, 0u);
                        used = 0;
                    }
                    const size_t lane = out.cln.size() - 64 + used, slot0 = out.ctgt.size() - 256 + used;
                    out.cln[lane] = make_uint4(dpos + 1 + ii, dpos, nu, 0u);
                    for (uint32_t e = 0; e < nu; ++e) {
                        const uint32_t tgt = umap[moff[k] + (uint64_t)(c0 + e) * L + ii];
                        out.ctgt[slot0 + 64 * e] = M.cnt[tgt] >= 2 ? UINT32_MAX : tgt;
                        out.cup[slot0 + 64 * e] = T.upos[uf + c0 + e];
                    }
                    ++used;
                }
            }
        }
        for (int t = P.level_ptr[l]; t < P.level_ptr[l + 1]; ++t) {
            const uint32_t k = (uint32_t)P.level_idx[t];
            const uint32_t dpos = P.l_col_ptr[k];
            for (uint32_t p = dpos + 1; p < P.sym_c_ptr[k + 1]; ++p)
                out.dvd.push_back(make_uint2(p, dpos | (p == dpos + 1 ? kFirstBit : 0u)));
        }
    });
    umap = std::vector<uint32_t>();
    // concatenation in level order
    T.ctask_ptr.assign(nl + 2, 0);
    T.ment_ptr.assign(nl + 2, 0);
    T.ment_long.assign(nl + 1, 0);
    T.dvd_ptr.assign(nl + 1, 0);
    T.msrc_ptr.assign(1, 0);
    size_t ncln = 0, nment = 0, nmsrc = 0, ndvd = 0;
    for (const LevelItems &o : lv) {
        ncln += o.cln.size();
        nment += o.ment.size();
        nmsrc += o.msrc.size();
        ndvd += o.dvd.size();
    }
    T.cln.reserve(ncln);
    T.ctgt.reserve(ncln * 4);
    T.cup.reserve(ncln * 4);
    T.ment.reserve(nment);
    T.msrc_ptr.reserve(nment + 1);
    T.msrc.reserve(nmsrc);
    T.dvd.reserve(ndvd);
    for (int l = 0; l < nl; ++l) {
        LevelItems &o = lv[l];
        T.ctask_ptr[l] = (uint32_t)(T.cln.size() / 64);
        T.ment_ptr[l] = (uint32_t)T.ment.size();
        T.dvd_ptr[l] = (uint32_t)T.dvd.size();
        T.ment_long[l] = o.ment_long;
        T.cln.insert(T.cln.end(), o.cln.begin(), o.cln.end());
        T.ctgt.insert(T.ctgt.end(), o.ctgt.begin(), o.ctgt.end());
        T.cup.insert(T.cup.end(), o.cup.begin(), o.cup.end());
        T.ment.insert(T.ment.end(), o.ment.begin(), o.ment.end());
        for (uint32_t c : o.msrc_cnt) T.msrc_ptr.push_back(T.msrc_ptr.back() + c);
        T.msrc.insert(T.msrc.end(), o.msrc.begin(), o.msrc.end());
        T.dvd.insert(T.dvd.end(), o.dvd.begin(), o.dvd.end());
        o = LevelItems();
    }
    T.ctask_ptr[nl] = T.ctask_ptr[nl + 1] = (uint32_t)(T.cln.size() / 64);
    T.ment_ptr[nl] = T.ment_ptr[nl + 1] = (uint32_t)T.ment.size();
    T.dvd_ptr[nl] = (uint32_t)T.dvd.size();
    if (T.ctgt.size() >= UINT32_MAX || T.msrc.size() >= UINT32_MAX) {
        err = "more than 2^32 update items";
        return false;
    }
    return true;
}

// Level sets of one triangular sweep.  dir > 0: rows ascending, dependencies
// on smaller indices; dir < 0: descending.  deps(r, f) calls f(c) for each
// dependency c of row r; len(r) is the row's entry count.  Within a level,
// rows longer than kLongRow come first (lvl_long of them), then the rest,
// each group in increasing row order.

template <typename Deps, typename Len>
void tri_levels(uint32_t n, int dir, Deps deps, Len len, TriHost &H)
{
    std::vector<int> lev(n, 0);
    int nlev = 0;
    for (uint32_t t = 0; t < n; ++t) {
        const uint32_t r = dir > 0 ? t : n - 1 - t;
        int m = -1;
        deps(r, [&](uint32_t c) { m = std::max(m, lev[c]); });
        lev[r] = m + 1;
        nlev = std::max(nlev, m + 2);
    }
    H.lvl_ptr.assign(nlev + 1, 0);
    H.lvl_long.assign(nlev, 0);
    H.len.resize(n);
    for (uint32_t r = 0; r < n; ++r) H.len[r] = len(r);
    for (uint32_t r = 0; r < n; ++r) {
        ++H.lvl_ptr[lev[r] + 1];
        if (len(r) > kLongRow) ++H.lvl_long[lev[r]];
    }
    for (int l = 0; l < nlev; ++l) H.lvl_ptr[l + 1] += H.lvl_ptr[l];
    H.rows.assign(n, 0);
    std::vector<uint32_t> cur_long(H.lvl_ptr.begin(), H.lvl_ptr.end() - 1), cur_short(nlev);
    for (int l = 0; l < nlev; ++l) cur_short[l] = H.lvl_ptr[l] + H.lvl_long[l];
    for (uint32_t r = 0; r < n; ++r) {
        if (len(r) > kLongRow) H.rows[cur_long[lev[r]]++] = r;
        else H.rows[cur_short[lev[r]]++] = r;
    }
}

void make_segments(const std::vector<uint32_t> &work, uint32_t narrow, std::vector<Segment> &segs)
{
    segs.clear();
    const int nl = (int)work.size();
    int l = 0;
    while (l < nl) {
        if (work[l] > narrow) {
            segs.push_back(Segment{l, l + 1, false, work[l]});
            ++l;
            continue;
        }
        int e = l;
        while (e < nl && work[e] <= narrow) ++e;
        segs.push_back(Segment{l, e, true, 0});
        l = e;
    }
}

// Lanes a level needs: column tasks and long merge entries a wave each.
uint32_t level_lanes(const Tables &T, int l)
{
    const uint32_t nc = T.ctask_ptr[l + 1] - T.ctask_ptr[l];
    const uint32_t nm = T.ment_ptr[l + 1] - T.ment_ptr[l];
    const uint32_t ml = T.ment_long[l];
    return std::max({nc * 64, nm - ml, ml * 64});   // (divisions run after the last step)
}

// LDS windows: runs of consecutive steps whose items touch at most `budget`
// distinct CSC positions run inside one workgroup with those values staged
// in LDS (slots, in increasing position order) and the run's items rewritten
// to address slots.  A run shorter than kMinWindow steps stays in global
// memory (staging would cost more than it saves).
constexpr int kMinWindow = 6;

struct Window {
    int l0 = 0, l1 = 0;            // steps [l0, l1)
    std::vector<uint32_t> pos;     // slot -> CSC position
    std::vector<uint2> cln;
    std::vector<uint32_t> ctgt, cup;
    std::vector<uint32_t> ment;
    std::vector<MergeSrc> msrc;
    std::vector<uint4> mrec;
    std::vector<uint2> dvd;
    uint32_t c0 = 0, g0 = 0, m0 = 0, s0 = 0, d0 = 0;   // first task / target / item of each kind
};

// The records {target, L, U, pivot of source 0}, {L, U, pivot of source 1,
// source count} of merge entries [e0, e1), positions through map.
template <typename Map>
void merge_records(const std::vector<uint32_t> &ment, const std::vector<uint32_t> &msrc_ptr,
                   const std::vector<MergeSrc> &msrc, uint32_t e0, uint32_t e1, Map map, std::vector<uint4> &out)
{
    for (uint32_t e = e0; e < e1; ++e) {
        const uint32_t s = msrc_ptr[e], cnt = msrc_ptr[e + 1] - s;
        const MergeSrc &a = msrc[s], &b = msrc[s + 1];
        out.push_back(make_uint4(map(ment[e]), map(a.lpos), map(a.upos), map(a.dpos)));
        out.push_back(make_uint4(map(b.lpos), map(b.upos), map(b.dpos), cnt));
    }
}

// The CSC positions the items of step l read or write, each once per call
// of f (duplicates possible).  Division items are not step items: they run
// after the last step, on the values in global memory.
template <typename F>
void step_positions(const Tables &T, int l, F f)
{
    for (uint32_t t = T.ctask_ptr[l]; t < T.ctask_ptr[l + 1]; ++t)
        for (uint32_t lane = 0; lane < 64; ++lane) {
            const uint4 &ln = T.cln[(size_t)t * 64 + lane];
            if (ln.z == 0) continue;
            f(ln.x), f(ln.y);
            for (uint32_t e = 0; e < ln.z; ++e) {
                const size_t q = 256 * (size_t)t + 64 * e + lane;
                f(T.cup[q]);
                if (T.ctgt[q] != UINT32_MAX) f(T.ctgt[q]);
            }
        }
    for (uint32_t e = T.ment_ptr[l]; e < T.ment_ptr[l + 1]; ++e) {
        f(T.ment[e]);
        for (uint32_t k = T.msrc_ptr[e]; k < T.msrc_ptr[e + 1]; ++k) {
            const MergeSrc &m = T.msrc[k];
            f(m.lpos), f(m.upos), f(m.dpos);
        }
    }
}

void make_window(const Tables &T, int a, int b, const std::vector<uint32_t> &slot_of_pos,
                 const std::vector<uint32_t> &pos, Window &w)
{
    w.l0 = a;
    w.l1 = b;
    w.pos = pos;
    auto slot = [&](uint32_t q) { return slot_of_pos[q]; };
    w.c0 = T.ctask_ptr[a];
    const uint32_t c1 = T.ctask_ptr[b];
    w.g0 = 256 * w.c0;
    // padding entries (no update): ~0u targets, U position slot 0
    for (size_t q = w.g0; q < 256 * (size_t)c1; ++q) {
        w.ctgt.push_back(T.ctgt[q] == UINT32_MAX ? UINT32_MAX : slot(T.ctgt[q]));
        w.cup.push_back(slot(T.cup[q]) != UINT32_MAX ? slot(T.cup[q]) : 0u);
    }
    for (size_t q = (size_t)w.c0 * 64; q < (size_t)c1 * 64; ++q) {
        const uint4 &ln = T.cln[q];
        w.cln.push_back(ln.z ? make_uint2(slot(ln.x), slot(ln.y)) : make_uint2(0u, 0u));
    }
    w.m0 = T.ment_ptr[a];
    const uint32_t m1 = T.ment_ptr[b];
    w.s0 = T.msrc_ptr[w.m0];
    for (uint32_t e = w.m0; e < m1; ++e) w.ment.push_back(slot(T.ment[e]));
    merge_records(T.ment, T.msrc_ptr, T.msrc, w.m0, m1, slot, w.mrec);
    for (uint32_t k = w.s0; k < T.msrc_ptr[m1]; ++k) {
        const MergeSrc &m = T.msrc[k];
        w.msrc.push_back(MergeSrc{slot(m.lpos), slot(m.upos), slot(m.dpos), 0});
    }
    w.d0 = 0;   // no division items inside a window (they run after the last step)
}

// Windows chosen backward from the last step, so that the final run (the
// longest chain of short steps) is one window.  Apart from that final
// window, windows hold narrow steps only (at most `narrow` lanes): a window
// runs in one workgroup, so a wide step inside it would give up the grid's
// parallelism.
void choose_windows(const Tables &T, int num_lev, uint32_t nnz, uint32_t budget, uint32_t narrow, int max_windows,
                    std::vector<Window> &wins)
{
    wins.clear();
    if (budget == 0 || max_windows <= 0) return;
    std::vector<uint32_t> stamp(nnz, UINT32_MAX), slot_of_pos(nnz, UINT32_MAX);
    std::vector<uint32_t> fresh;
    int b = num_lev + 1;
    uint32_t gen = 0;
    while (b > 0 && (int)wins.size() < max_windows) {
        const uint32_t lim = b == num_lev + 1 ? UINT32_MAX : narrow;
        if (level_lanes(T, b - 1) > lim) {
            --b;
            continue;
        }
        ++gen;
        uint32_t count = 0;
        int a = b;
        while (a > 0 && level_lanes(T, a - 1) <= lim) {
            fresh.clear();
            step_positions(T, a - 1, [&](uint32_t q) {
                if (stamp[q] != gen) {
                    stamp[q] = gen;
                    fresh.push_back(q);
                }
            });
            if (count + fresh.size() > budget) {
                for (uint32_t q : fresh) stamp[q] = UINT32_MAX;
                break;
            }
            count += (uint32_t)fresh.size();
            --a;
        }
        if (b - a < kMinWindow) {
            b = a < b ? a : b - 1;   // too short: leave these steps in global memory
            continue;
        }
        std::vector<uint32_t> pos;
        for (int l = a; l < b; ++l)
            step_positions(T, l, [&](uint32_t q) {
                if (slot_of_pos[q] == UINT32_MAX) {
                    slot_of_pos[q] = 0;
                    pos.push_back(q);
                }
            });
        std::sort(pos.begin(), pos.end());
        for (uint32_t i = 0; i < pos.size(); ++i) slot_of_pos[pos[i]] = i;
        wins.emplace_back();
        make_window(T, a, b, slot_of_pos, pos, wins.back());
        for (uint32_t q : pos) slot_of_pos[q] = UINT32_MAX;
        b = a;
    }
    std::reverse(wins.begin(), wins.end());
}

// Segments of the factorization schedule: windows (one workgroup each) and,
// between them, runs of global-memory steps split into fused and grid
// segments.
void factor_segments(glu_plan &p, const Tables &T, const std::vector<Window> &wins)
{
    p.fsegs.clear();
    int l = 0;
    auto global_run = [&](int a, int b) {
        if (a >= b) return;
        std::vector<uint32_t> work(b - a);
        for (int q = a; q < b; ++q) work[q - a] = level_lanes(T, q);
        std::vector<Segment> segs;
        make_segments(work, (uint32_t)p.fuse_units, segs);
        for (Segment s : segs) {
            s.l0 += a;
            s.l1 += a;
            p.fsegs.push_back(s);
        }
    };
    for (size_t w = 0; w < wins.size(); ++w) {
        global_run(l, wins[w].l0);
        Segment s{wins[w].l0, wins[w].l1, true, 0};
        s.win = (int)w;
        p.fsegs.push_back(s);
        l = wins[w].l1;
    }
    global_run(l, p.num_lev + 1);
}

void tri_segments(glu_plan &p, TriHost &H)
{
    const int nlev = (int)H.lvl_ptr.size() - 1;
    std::vector<uint32_t> work(nlev);
    for (int l = 0; l < nlev; ++l) {
        const uint32_t rows = H.lvl_ptr[l + 1] - H.lvl_ptr[l];
        work[l] = (rows - H.lvl_long[l]) + H.lvl_long[l] * 64;   // lanes
    }
    make_segments(work, (uint32_t)p.fused_threads * 4, H.segs);
}

// Dataflow tasks of a sweep: in schedule order, each row with more than
// kDfLaneRow entries is a task of its own (a wave reads its entries at
// once), and consecutive shorter rows form tasks of up to 64 (one wave, a
// row per lane, every entry prefetched at the start), which may span
// levels.  No task crosses a segment boundary, so a fused segment runs
// tasks [k0, k1) by itself.
void tri_tasks(TriHost &H)
{
    H.tasks.clear();
    for (Segment &s : H.segs) {
        s.k0 = (uint32_t)H.tasks.size();
        uint32_t run = 0;   // rows in the open short task
        for (uint32_t q = H.lvl_ptr[s.l0]; q < H.lvl_ptr[s.l1]; ++q) {
            if (H.len[H.rows[q]] > kDfLaneRow) {
                H.tasks.push_back(q | kLongTask);
                run = 0;
                continue;
            }
            if (run == 0) H.tasks.push_back(q);
            run = run + 1 == 64 ? 0 : run + 1;   // a wave of rows
        }
        s.k1 = (uint32_t)H.tasks.size();
    }
    H.tasks.push_back(H.lvl_ptr.back());
}

// Task-contiguous layout of one sweep for the pipelined dataflow kernels
// (refinement sweeps of the batch).  entries(r, out) lists row r's
// {column, CSC position} pairs in the order the row consumes them; pivot
// (div sweeps) gives the CSC position of row r's divisor.
struct DfHost {
    std::vector<uint4> task;
    std::vector<uint2> rl;
    std::vector<uint32_t> idx, vpos;   // per slot
    std::vector<uint32_t> pvpos;       // per position: pivot CSC position (div)
};

// Lanes per row of a wave row: 16, 32 or 64 (longer rows take several
// passes of 64).
inline uint32_t df_group(uint32_t len) { return len <= 16 ? 16u : len <= 32 ? 32u : 64u; }

template <typename Entries, typename Pivot>
void build_df(const TriHost &H, Entries entries, bool div, Pivot pivot, bool long_lanes, DfHost &D)
{
    const uint32_t n = (uint32_t)H.rows.size();
    const int nlev = (int)H.lvl_ptr.size() - 1;
    std::vector<uint32_t> len(n);
    for (uint32_t r = 0; r < n; ++r) {
        std::vector<uint2> tmp;
        entries(r, tmp);
        len[r] = (uint32_t)tmp.size();
    }
    // Rows are taken a lane each (lane rows) or by groups of lanes (wave
    // rows).  Short rows (<= kDfK entries) are always lane rows.  With
    // long_lanes (orders whose sweep vector does not fit in LDS, where a
    // sweep is bound by the one workgroup's throughput), longer rows of up
    // to kDfLaneMax entries are lane rows too in a level with at least
    // kDfWideLevel of them: a lane walks its row kDfK entries at a time, 64
    // rows per task, where a wave would take 1 to 4 of them; in narrower
    // levels, which make up the sweep's critical path, they stay wave rows
    // (one memory round trip per row).  Measured: n=500k solve 55.7 ->
    // 34.6 ms, n=50k 2.57 -> 1.81 ms; the n=20k batch (vector in LDS) loses
    // 0.3 ms with them, so smaller orders keep wave rows.
    // Within each level (whose rows are independent) the short lane rows
    // are kept together, after the wave rows on even levels and before them
    // on odd ones, so that lane runs of adjacent levels meet and pack into
    // full tasks; long lane rows sit between the two, sorted by length so
    // that a task's rows have similar lengths (a task's slots are sized by
    // its longest row); wave rows are sorted by length so that rows sharing
    // a wave (16 or 32 lanes each) have the same lane count.
    std::vector<uint32_t> order(n);
    std::vector<uint8_t> lane_row(n);   // by position
    for (int l = 0; l < nlev; ++l) {
        std::vector<uint32_t> a, c, b;   // short lane rows, long lane rows, wave rows
        for (uint32_t q = H.lvl_ptr[l]; q < H.lvl_ptr[l + 1]; ++q) {
            const uint32_t r = H.rows[q];
            (len[r] <= kDfK ? a : len[r] <= kDfLaneMax ? c : b).push_back(r);
        }
        if (!long_lanes || c.size() < kDfWideLevel) {
            b.insert(b.end(), c.begin(), c.end());
            c.clear();
        }
        auto by_len = [&](uint32_t x, uint32_t y) { return len[x] < len[y]; };
        std::stable_sort(b.begin(), b.end(), by_len);
        std::stable_sort(c.begin(), c.end(), by_len);
        uint32_t q = H.lvl_ptr[l];
        auto put_rows = [&](auto first, auto last, uint8_t lane) {
            for (; first != last; ++first, ++q) {
                order[q] = *first;
                lane_row[q] = lane;
            }
        };
        if (l & 1) {
            put_rows(a.begin(), a.end(), 1);
            put_rows(c.begin(), c.end(), 1);
            put_rows(b.begin(), b.end(), 0);
        } else {
            put_rows(b.begin(), b.end(), 0);
            put_rows(c.rbegin(), c.rend(), 1);
            put_rows(a.begin(), a.end(), 1);
        }
    }
    std::vector<int> lev_of(n);
    for (int l = 0; l < nlev; ++l)
        for (uint32_t q = H.lvl_ptr[l]; q < H.lvl_ptr[l + 1]; ++q) lev_of[q] = l;
    std::vector<std::vector<uint2>> ent(n);
    D.rl.resize(n);
    for (uint32_t p = 0; p < n; ++p) {
        entries(order[p], ent[p]);
        D.rl[p] = make_uint2(order[p], (uint32_t)ent[p].size());
    }
    auto put = [&](const uint2 &e) {
        D.idx.push_back(e.x);
        D.vpos.push_back(e.y);
    };
    const uint2 pad = make_uint2(0u, UINT32_MAX);
    uint32_t p = 0;
    while (p < n) {
        const uint32_t base = (uint32_t)D.idx.size();
        const uint32_t p0 = p;
        if (!lane_row[p]) {
            // rows of one lane count and one level (no dependencies inside)
            const uint32_t gw = df_group((uint32_t)ent[p].size()), per = 64 / gw;
            uint32_t m = 0;
            while (p < n && p - p0 < per && lev_of[p] == lev_of[p0] && !lane_row[p] &&
                   df_group((uint32_t)ent[p].size()) == gw)
                m = std::max(m, (uint32_t)ent[p++].size());
            if (gw == 64) {   // padded to whole passes of 64 (every lane loads its slot)
                for (const uint2 &e : ent[p0]) put(e);
                while ((D.idx.size() - base) % 64) put(pad);
            } else {
                for (uint32_t g = 0; g < per; ++g)
                    for (uint32_t j = 0; j < gw; ++j)
                        put(p0 + g < p && j < ent[p0 + g].size() ? ent[p0 + g][j] : pad);
            }
            const uint32_t lg = gw == 16 ? 4u : gw == 32 ? 5u : 6u;
            D.task.push_back(make_uint4(base, p0, kLongTask | (lg << 24) | (p - p0), m));
            continue;
        }
        // lane rows, short and long ones in separate tasks
        const bool lng = ent[p].size() > kDfK;
        uint32_t m = 0;
        while (p < n && p - p0 < 64 && lane_row[p] && (ent[p].size() > kDfK) == lng)
            m = std::max(m, (uint32_t)ent[p++].size());
        for (uint32_t e = 0; e < m; ++e)
            for (uint32_t l = 0; l < 64; ++l) put(p0 + l < p && e < ent[p0 + l].size() ? ent[p0 + l][e] : pad);
        D.task.push_back(make_uint4(base, p0, p - p0, m));
    }
    // padding, so that the kernels' loads need no bounds tests: kDfK passes
    // of 64 slots after the last task, 64 positions after the last row
    for (uint32_t q = 0; q < kDfK * 64; ++q) put(pad);
    D.rl.resize(n + 64, make_uint2(0u, 0u));
    if (div) {
        D.pvpos.resize(n);
        for (uint32_t q = 0; q < n; ++q) D.pvpos[q] = pivot(order[q]);
    }
}

// Host check of a dataflow sweep table, before any launch.  Every load the
// sweep kernels issue stays inside the tables (a lane task's prefetch of
// kDfK slots per lane and its chunk refills, a wave task's slots, rows and
// pivots); every slot's value position is in the pattern or padding; every
// row runs exactly once; and every dependency of a row is finished by an
// earlier task or, in a lane task, by an earlier lane of the same task
// (the rows of a wave task never depend on each other).  Tasks are taken in
// order, so the lowest unfinished task always has its dependencies finished
// or inside itself: every ready-bit wait of the sweeps ends.
bool check_df(const DfHost &D, uint32_t n, uint32_t nnz, const char *name, std::string &err)
{
    const size_t nslots = D.idx.size();
    std::vector<uint8_t> done(n, 0);
    auto bad = [&](const std::string &what) {
        err = std::string(name) + " sweep table: " + what;
        return false;
    };
    if (D.rl.size() < (size_t)n + 64) return bad("row list not padded");
    for (size_t s = 0; s < nslots; ++s)
        if (D.vpos[s] != UINT32_MAX && D.vpos[s] >= nnz) return bad("value position out of range");
    for (size_t k = 0; k < D.task.size(); ++k) {
        const uint4 d = D.task[k];
        const uint32_t base = d.x, p0 = d.y, z = d.z, m = d.w;
        if ((size_t)p0 + 64 > D.rl.size()) return bad("task rows out of range");
        if (!(z & kLongTask)) {
            const uint32_t cnt = z;
            if (cnt > 64 || (size_t)base + (size_t)kDfK * 64 > nslots) return bad("lane task out of range");
            for (uint32_t lane = 0; lane < cnt; ++lane) {
                const uint32_t row = D.rl[p0 + lane].x, len = D.rl[p0 + lane].y;
                if (row >= n || len > m) return bad("lane row out of range");
                // a long lane row loads its entries kDfK at a time: up to the
                // end of its last chunk
                const size_t chunks = std::max<size_t>(1, (len + kDfK - 1) / kDfK);
                if (base + (chunks * kDfK - 1) * 64 + lane >= nslots) return bad("lane row slots out of range");
                for (uint32_t e = 0; e < len; ++e) {
                    const uint32_t dep = D.idx[base + (size_t)e * 64 + lane];
                    if (dep >= n || !done[dep]) return bad("row runs before its dependency");
                }
                if (done[row]) return bad("row runs twice");
                done[row] = 1;
            }
        } else {
            const uint32_t lg = (z >> 24) & 0x7fu, cnt = z & 0xffffffu, gw = 1u << lg;
            if (lg < 4 || lg > 6 || cnt == 0 || cnt * gw > 64) return bad("wave task shape");
            if ((size_t)base + 64 > nslots) return bad("wave task slots out of range");
            for (uint32_t g = 0; g < cnt; ++g) {
                const uint32_t row = D.rl[p0 + g].x, len = D.rl[p0 + g].y;
                if (row >= n || (gw < 64 && len > gw)) return bad("wave row out of range");
                if (gw == 64 && (cnt != 1 || (size_t)base + len > nslots)) return bad("wave row slots out of range");
                for (uint32_t j = 0; j < len; ++j) {
                    const uint32_t dep = D.idx[base + (size_t)gw * g + j];
                    if (dep >= n || !done[dep]) return bad("row runs before its dependency");
                }
            }
            for (uint32_t g = 0; g < cnt; ++g) {
                const uint32_t row = D.rl[p0 + g].x;
                if (done[row]) return bad("row runs twice");
                done[row] = 1;
            }
        }
    }
    for (uint32_t r = 0; r < n; ++r)
        if (!done[r]) return bad("a row never runs");
    return true;
}

struct Uploader {
    Arena arena;
    struct Item {
        size_t off;
        const void *src;
        size_t bytes;
    };
    std::vector<Item> items;
    template <typename T>
    size_t put(const T *src, size_t count)
    {
        size_t off = arena.add(count * sizeof(T));
        items.push_back(Item{off, src, count * sizeof(T)});
        return off;
    }
    template <typename T>
    size_t put(const std::vector<T> &v)
    {
        return put(v.data(), v.size());
    }
    size_t reserve(size_t bytes) { return arena.add(bytes); }
};

glu_status build_plan(const glu_pattern &P, const glu_analysis *a, const glu_plan_opts *opts, glu_plan **out)
{
    *out = nullptr;
    std::string err;
    if (P.n == 0 || P.nnz == 0) return fail(GLU_ERR_EMPTY, "matrix is empty");
    if (!validate_pattern(P, err)) return fail(GLU_ERR_INVALID, err);

    std::unique_ptr<glu_plan> p(new (std::nothrow) glu_plan());
    if (!p) return fail(GLU_ERR_MEMORY, "out of host memory");
    if (opts) {
        if (opts->block_threads) {
            if (opts->block_threads < 64 || opts->block_threads > 1024 || opts->block_threads % 64)
                return fail(GLU_ERR_ARGUMENT, "block_threads must be a multiple of 64 in [64, 1024]");
            p->fused_threads = opts->block_threads;
        }
        if (opts->fuse_units < 0) return fail(GLU_ERR_ARGUMENT, "fuse_units must be >= 0");
        if (opts->fuse_units) p->fuse_units = opts->fuse_units;
        p->use_graph = opts->use_graph != 0;
        if (opts->lds_windows < -1) return fail(GLU_ERR_ARGUMENT, "lds_windows must be >= -1");
        if (opts->lds_windows) p->max_windows = opts->lds_windows < 0 ? 0 : opts->lds_windows;
        if (opts->solve_dataflow < -1 || opts->solve_dataflow > 1)
            return fail(GLU_ERR_ARGUMENT, "solve_dataflow must be -1, 0 or 1");
        p->solve_df_opt = opts->solve_dataflow;
        if (opts->krylov < -1 || opts->krylov > kMaxKrylov)
            return fail(GLU_ERR_ARGUMENT, "krylov must be in [-1, 4]");
        if (opts->krylov) p->krylov = opts->krylov < 0 ? 0 : opts->krylov;
        if (opts->lds_window_values < 0) return fail(GLU_ERR_ARGUMENT, "lds_window_values must be >= 0");
        p->win_values = (uint32_t)opts->lds_window_values;
        if (opts->batch_grid_levels < -1) return fail(GLU_ERR_ARGUMENT, "batch_grid_levels must be >= -1");
        p->batch_grid_opt = opts->batch_grid_levels;
    }
    p->n = P.n;
    p->nnz = P.nnz;
    p->num_lev = P.num_lev;

    Tables T;
    try {
        if (!build_factor_tables(P, T, p->updates, err)) return fail(GLU_ERR_INVALID, err);
    } catch (const std::bad_alloc &) {
        return fail(GLU_ERR_MEMORY, "out of host memory building tables");
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(GLU_ERR_DEVICE, "no HIP device");
    HIP_TRY(hipGetDevice(&p->device));
    p->lds_cap = launch::enable_lds();
    std::vector<Window> wins;
    uint32_t budget = (uint32_t)(p->lds_cap / sizeof(float));
    if (p->win_values) budget = std::min(budget, p->win_values);
    choose_windows(T, P.num_lev, P.nnz, budget, (uint32_t)p->fuse_units, p->max_windows, wins);
    p->nwin = (int)wins.size();
    for (const Window &w : wins) {
        p->win_steps += w.l1 - w.l0;
        p->win_nslot = std::max<uint32_t>(p->win_nslot, (uint32_t)w.pos.size());
        p->win_slots.push_back((uint32_t)w.pos.size());
    }
    factor_segments(*p, T, wins);
    p->units = (int64_t)(T.cln.size() / 64);
    // The batched kernels' leading grid steps: those ahead of the first LDS
    // window with at least a workgroup's worth of lanes per matrix (a
    // narrower step takes one pass of the per-matrix workgroup anyway).
    {
        const int lim = std::min({wins.empty() ? P.num_lev : std::min(wins[0].l0, P.num_lev),
                                  p->batch_grid_opt < 0 ? 0 : p->batch_grid_opt > 0 ? p->batch_grid_opt : INT32_MAX});
        const uint32_t per_wave = launch::il_items_per_wave();
        for (int l = 0; l < lim && level_lanes(T, l) >= (uint32_t)p->fused_threads; ++l) {
            const uint32_t ml = T.ment_long[l], nm = T.ment_ptr[l + 1] - T.ment_ptr[l];
            const uint32_t items = (T.ctask_ptr[l + 1] - T.ctask_ptr[l]) * 64 + (nm - ml) +
                                   (l > 0 ? T.dvd_ptr[l] - T.dvd_ptr[l - 1] : 0u);   // + previous step's divisions
            p->grid_waves.push_back(ml + (items + per_wave - 1) / per_wave);
        }
        p->batch_grid_lev = (int)p->grid_waves.size();
    }
    // Division items: those of the columns of each of the batched kernels'
    // grid steps but the last run with the next grid step (kept step by
    // step); the rest run in one pass after the last step (the single path
    // runs them all there).  Each step's items, and the rest, are in
    // position order: a wave's items are then consecutive L entries of a
    // column (each column's first entry first) and read and write whole
    // lines.
    {
        const auto by_pos = [](const uint2 &a, const uint2 &b) { return a.x < b.x; };
        const int early = std::max(p->batch_grid_lev - 1, 0);
        for (int l = 0; l < early; ++l)
            std::sort(T.dvd.begin() + T.dvd_ptr[l], T.dvd.begin() + T.dvd_ptr[l + 1], by_pos);
        std::sort(T.dvd.begin() + T.dvd_ptr[early], T.dvd.end(), by_pos);
    }

    const uint32_t n = P.n;
    // Solve schedules.
    tri_levels(n, +1, [&](uint32_t r, auto f) {
        for (uint32_t q = P.csr_r_ptr[r]; q < P.csr_diag_ptr[r]; ++q) f(P.csr_c_idx[q]);
    }, [&](uint32_t r) { return P.csr_diag_ptr[r] - P.csr_r_ptr[r]; }, p->tri[0]);
    tri_levels(n, -1, [&](uint32_t r, auto f) {
        for (uint32_t q = P.csr_diag_ptr[r] + 1; q < P.csr_r_ptr[r + 1]; ++q) f(P.csr_c_idx[q]);
    }, [&](uint32_t r) { return P.csr_r_ptr[r + 1] - P.csr_diag_ptr[r] - 1; }, p->tri[1]);
    tri_levels(n, +1, [&](uint32_t r, auto f) {
        for (uint32_t q = P.sym_c_ptr[r]; q < P.l_col_ptr[r]; ++q) f(P.sym_r_idx[q]);
    }, [&](uint32_t r) { return P.l_col_ptr[r] - P.sym_c_ptr[r]; }, p->tri[2]);
    tri_levels(n, -1, [&](uint32_t r, auto f) {
        for (uint32_t q = P.l_col_ptr[r] + 1; q < P.sym_c_ptr[r + 1]; ++q) f(P.sym_r_idx[q]);
    }, [&](uint32_t r) { return P.sym_c_ptr[r + 1] - P.l_col_ptr[r] - 1; }, p->tri[3]);
    for (auto &h : p->tri) {
        tri_segments(*p, h);
        tri_tasks(h);
    }
    // sweep vector beyond LDS (the richest sweep mode does not fit): long
    // lane rows in wide levels
    const bool long_lanes = launch::batch_lds_bytes(2, n) > p->lds_cap;
    // the four sweeps' dataflow tables (L, U; U', L' for A' x = b: rows of
    // U' are columns of U with the pivot last, forward; rows of L' columns
    // of L, backward), built and checked on four host threads
    DfHost dlo, dup, dtu, dtl;
    std::string errs[4];
    bool ok[4] = {false, false, false, false};
    parallel_items(4, std::min(4u, table_threads()), [&](unsigned, size_t which) {
        switch (which) {
        case 0:
            build_df(p->tri[0], [&](uint32_t r, std::vector<uint2> &out) {
                for (uint32_t q = P.csr_r_ptr[r]; q < P.csr_diag_ptr[r]; ++q)
                    out.push_back(make_uint2(P.csr_c_idx[q], T.upos[q]));
            }, false, [&](uint32_t) { return 0u; }, long_lanes, dlo);
            ok[0] = check_df(dlo, n, P.nnz, "L", errs[0]);
            break;
        case 1:
            build_df(p->tri[1], [&](uint32_t r, std::vector<uint2> &out) {
                for (uint32_t q = P.csr_r_ptr[r + 1]; q > P.csr_diag_ptr[r] + 1; --q)
                    out.push_back(make_uint2(P.csr_c_idx[q - 1], T.upos[q - 1]));
            }, true, [&](uint32_t r) { return P.l_col_ptr[r]; }, long_lanes, dup);
            ok[1] = check_df(dup, n, P.nnz, "U", errs[1]);
            break;
        case 2:
            build_df(p->tri[2], [&](uint32_t r, std::vector<uint2> &out) {
                for (uint32_t q = P.sym_c_ptr[r]; q < P.l_col_ptr[r]; ++q) out.push_back(make_uint2(P.sym_r_idx[q], q));
            }, true, [&](uint32_t r) { return P.l_col_ptr[r]; }, long_lanes, dtu);
            ok[2] = check_df(dtu, n, P.nnz, "U'", errs[2]);
            break;
        default:
            build_df(p->tri[3], [&](uint32_t r, std::vector<uint2> &out) {
                for (uint32_t q = P.sym_c_ptr[r + 1]; q > P.l_col_ptr[r] + 1; --q)
                    out.push_back(make_uint2(P.sym_r_idx[q - 1], q - 1));
            }, false, [&](uint32_t) { return 0u; }, long_lanes, dtl);
            ok[3] = check_df(dtl, n, P.nnz, "L'", errs[3]);
        }
    });
    for (int q = 0; q < 4; ++q)
        if (!ok[q]) return fail(GLU_ERR_INVALID, errs[q]);
    // the value copies are gathered in slot order (whole-line stores): the
    // CSC position of each slot, ~0 for padding slots (which get 0)
    std::vector<uint32_t> tdf_vpos(dtu.vpos);   // U' slots, U' pivots (+64 pad), L' slots
    tdf_vpos.insert(tdf_vpos.end(), dtu.pvpos.begin(), dtu.pvpos.end());
    tdf_vpos.resize(tdf_vpos.size() + 64, UINT32_MAX);
    const uint32_t tlo_voff = (uint32_t)tdf_vpos.size();
    tdf_vpos.insert(tdf_vpos.end(), dtl.vpos.begin(), dtl.vpos.end());
    p->tdf_nval = (uint32_t)tdf_vpos.size();
    // per-matrix value copy: L slots, U slots, U pivots
    std::vector<uint32_t> df_vpos(dlo.vpos);
    df_vpos.insert(df_vpos.end(), dup.vpos.begin(), dup.vpos.end());
    df_vpos.insert(df_vpos.end(), dup.pvpos.begin(), dup.pvpos.end());
    df_vpos.resize(df_vpos.size() + 64, UINT32_MAX);   // pivot reads of the last positions' lanes
    p->df_nval = (uint32_t)df_vpos.size();
    // The batch kernels build the copy in LDS, copy_range slots at a time:
    // per range, {CSC position, slot - range start} of its non-padding
    // slots in increasing position order (so that the reads sweep the
    // factors in order), then the range is written out whole.
    p->copy_range = std::min<uint32_t>({p->df_nval, kCopyRange, (uint32_t)(p->lds_cap / sizeof(float)) & ~63u});
    if (p->copy_range == 0) return fail(GLU_ERR_DEVICE, "no LDS for the batch value copy");
    std::vector<uint2> df_cpy;
    std::vector<uint32_t> df_cpy_ptr(1, 0);
    df_cpy.reserve(p->df_nval);
    for (uint32_t b = 0; b < p->df_nval; b += p->copy_range) {
        const uint32_t e = std::min<uint32_t>(p->df_nval, b + p->copy_range);
        const size_t first = df_cpy.size();
        for (uint32_t q = b; q < e; ++q)
            if (df_vpos[q] != UINT32_MAX) df_cpy.push_back(make_uint2(df_vpos[q], q - b));
        std::sort(df_cpy.begin() + first, df_cpy.end(),
                  [](const uint2 &x, const uint2 &y) { return x.x < y.x || (x.x == y.x && x.y < y.y); });
        df_cpy_ptr.push_back((uint32_t)df_cpy.size());
    }

    // Refinement / refresh tables from the analysis.
    std::vector<uint32_t> a_rptr, a_cidx, a_src, lu_src, lu_col, a_rows, a_cols;
    std::vector<double> rs_col, a_rs, a_cs;
    std::vector<uint32_t> a_dst, lu_fill;
    const Analysis *an = a ? &a->an : nullptr;
    if (an) {
        const uint32_t nnz_a = an->nnz;
        p->nnz_a = nnz_a;
        // CSR of the caller's A (rows = caller rows), remembering each
        // entry's position in the caller's CSC value array.
        a_rptr.assign(n + 1, 0);
        for (uint32_t k = 0; k < nnz_a; ++k) ++a_rptr[an->in_ai[k] + 1];
        for (uint32_t i = 0; i < n; ++i) a_rptr[i + 1] += a_rptr[i];
        a_cidx.resize(nnz_a);
        a_src.resize(nnz_a);
        std::vector<uint32_t> cur(a_rptr.begin(), a_rptr.end() - 1);
        for (uint32_t c = 0; c < n; ++c)
            for (uint32_t k = an->in_ap[c]; k < an->in_ap[c + 1]; ++k) {
                const uint32_t q = cur[an->in_ai[k]]++;
                a_cidx[q] = c;
                a_src[q] = k;
            }
        lu_src.resize(P.nnz);
        for (uint32_t q = 0; q < P.nnz; ++q) {
            const uint32_t d = a->lu_from_dump[q];
            lu_src[q] = (d == UINT32_MAX) ? UINT32_MAX : an->src_pos[d];
        }
        rs_col.resize(n);
        for (uint32_t c = 0; c < n; ++c) rs_col[c] = an->row_scale[an->row_perm[c]];
        lu_col.resize(P.nnz);
        for (uint32_t c = 0; c < n; ++c)
            for (uint32_t q = P.sym_c_ptr[c]; q < P.sym_c_ptr[c + 1]; ++q) lu_col[q] = c;
        // the inverse map: every caller value lands on exactly one LU position
        a_dst.assign(nnz_a, UINT32_MAX);
        a_rs.assign(nnz_a, 0.0);
        a_cs.assign(nnz_a, 0.0);
        for (uint32_t q = 0; q < P.nnz; ++q) {
            const uint32_t k = lu_src[q];
            if (k == UINT32_MAX) continue;
            if (k >= nnz_a || a_dst[k] != UINT32_MAX) return fail(GLU_ERR_INVALID, "value map is not one-to-one");
            a_dst[k] = q;
            a_rs[k] = rs_col[lu_col[q]];
            a_cs[k] = an->col_scale_perm[P.sym_r_idx[q]];
        }
        for (uint32_t k = 0; k < nnz_a; ++k)
            if (a_dst[k] == UINT32_MAX) return fail(GLU_ERR_INVALID, "a value of A has no LU position");
        for (uint32_t q = 0; q < P.nnz; ++q)
            if (lu_src[q] == UINT32_MAX) lu_fill.push_back(q);
        // rows / columns of A with more than kLongRow entries first
        auto long_first = [&](const uint32_t *ptr, std::vector<uint32_t> &order, uint32_t &nlong) {
            order.clear();
            for (uint32_t i = 0; i < n; ++i)
                if (ptr[i + 1] - ptr[i] > kLongRow) order.push_back(i);
            nlong = (uint32_t)order.size();
            for (uint32_t i = 0; i < n; ++i)
                if (ptr[i + 1] - ptr[i] <= kLongRow) order.push_back(i);
        };
        long_first(a_rptr.data(), a_rows, p->a_nlong);
        long_first(an->in_ap.data(), a_cols, p->a_nlong_t);
        for (uint32_t i = 0; i < n; ++i)
            p->a_kmax = std::max({p->a_kmax, a_rptr[i + 1] - a_rptr[i], an->in_ap[i + 1] - an->in_ap[i]});
        p->has_scaling = p->has_refine = true;
    }

    // ---------------- upload everything into one arena ------------------
    Uploader U;
    const size_t o_colptr = U.put(P.sym_c_ptr, n + 1);
    const size_t o_rowidx = U.put(P.sym_r_idx, P.nnz);
    const size_t o_diag = U.put(P.l_col_ptr, n);
    const size_t o_crptr = U.put(P.csr_r_ptr, n + 1);
    const size_t o_ccidx = U.put(P.csr_c_idx, P.nnz);
    const size_t o_cdiag = U.put(P.csr_diag_ptr, n);
    const size_t o_upos = U.put(T.upos);
    const size_t o_cptr = U.put(T.ctask_ptr);
    std::vector<uint2> cln2(T.cln.size());   // {L position, pivot position} per lane
    for (size_t q = 0; q < T.cln.size(); ++q) cln2[q] = make_uint2(T.cln[q].x, T.cln[q].y);
    const size_t o_cln = U.put(cln2);
    const size_t o_ctgt = U.put(T.ctgt);
    const size_t o_cup = U.put(T.cup);
    const size_t o_ment = U.put(T.ment);
    const size_t o_mptr = U.put(T.ment_ptr);
    const size_t o_mlong = U.put(T.ment_long);
    const size_t o_msptr = U.put(T.msrc_ptr);
    const size_t o_msrc = U.put(T.msrc);
    std::vector<uint4> mrec;
    merge_records(T.ment, T.msrc_ptr, T.msrc, 0, (uint32_t)T.ment.size(), [](uint32_t q) { return q; }, mrec);
    const size_t o_mrec = U.put(mrec);
    const size_t o_dvd = U.put(T.dvd);
    const size_t o_dptr = U.put(T.dvd_ptr);
    p->ndiv = (uint32_t)T.dvd.size();
    std::vector<size_t> o_wpos(wins.size()), o_wcln(wins.size()), o_wctgt(wins.size()), o_wcup(wins.size()),
        o_wment(wins.size()), o_wmsrc(wins.size()), o_wmrec(wins.size()), o_wdvd(wins.size());
    for (size_t w = 0; w < wins.size(); ++w) {
        o_wpos[w] = U.put(wins[w].pos);
        o_wcln[w] = U.put(wins[w].cln);
        o_wctgt[w] = U.put(wins[w].ctgt);
        o_wcup[w] = U.put(wins[w].cup);
        o_wment[w] = U.put(wins[w].ment);
        o_wmsrc[w] = U.put(wins[w].msrc);
        o_wmrec[w] = U.put(wins[w].mrec);
        o_wdvd[w] = U.put(wins[w].dvd);
    }
    const size_t o_wdev = U.reserve(sizeof(WinDev) * std::max<size_t>(wins.size(), 1));
    const size_t o_dlo_task = U.put(dlo.task), o_dlo_rl = U.put(dlo.rl), o_dlo_idx = U.put(dlo.idx);
    const size_t o_dup_task = U.put(dup.task), o_dup_rl = U.put(dup.rl), o_dup_idx = U.put(dup.idx);
    const size_t o_dtu_task = U.put(dtu.task), o_dtu_rl = U.put(dtu.rl), o_dtu_idx = U.put(dtu.idx);
    const size_t o_dtl_task = U.put(dtl.task), o_dtl_rl = U.put(dtl.rl), o_dtl_idx = U.put(dtl.idx);
    const size_t o_tdfsrc = U.put(tdf_vpos);
    const size_t o_dfsrc = U.put(df_vpos);
    const size_t o_dfcpy = U.put(df_cpy), o_dfcptr = U.put(df_cpy_ptr);
    size_t o_trows[4], o_tptr[4], o_tlong[4], o_ttask[4];
    for (int s = 0; s < 4; ++s) {
        o_trows[s] = U.put(p->tri[s].rows);
        o_tptr[s] = U.put(p->tri[s].lvl_ptr);
        o_tlong[s] = U.put(p->tri[s].lvl_long);
        o_ttask[s] = U.put(p->tri[s].tasks);
    }
    size_t o_rp = 0, o_rpi = 0, o_cp = 0, o_cpi = 0, o_rs = 0, o_csp = 0;
    size_t o_arptr = 0, o_acidx = 0, o_asrc = 0, o_acptr = 0, o_aridx = 0, o_lusrc = 0, o_rscol = 0;
    size_t o_lucol = 0, o_arows = 0, o_acols = 0, o_adst = 0, o_ars = 0, o_acs = 0, o_lufill = 0;
    if (an) {
        o_rp = U.put(an->row_perm);
        o_rpi = U.put(an->row_perm_inv);
        o_cp = U.put(an->col_perm);
        o_cpi = U.put(an->col_perm_inv);
        o_rs = U.put(an->row_scale);
        o_csp = U.put(an->col_scale_perm);
        o_arptr = U.put(a_rptr);
        o_acidx = U.put(a_cidx);
        o_asrc = U.put(a_src);
        o_acptr = U.put(an->in_ap);
        o_aridx = U.put(an->in_ai);
        o_lusrc = U.put(lu_src);
        o_rscol = U.put(rs_col);
        o_lucol = U.put(lu_col);
        o_arows = U.put(a_rows);
        o_acols = U.put(a_cols);
        o_adst = U.put(a_dst);
        o_lufill = U.put(lu_fill);
        o_ars = U.put(a_rs);
        o_acs = U.put(a_cs);
    }
    const size_t o_desc = U.reserve(sizeof(MatDesc));
    const size_t o_normbits = U.reserve(sizeof(unsigned int));
    const size_t o_pert = U.reserve(sizeof(float));
    const size_t o_norms = U.reserve(4 * sizeof(unsigned long long));
    const size_t o_t32 = U.reserve(sizeof(float) * n);
    const size_t o_r64 = U.reserve(sizeof(double) * n);

    HIP_TRY(hipMalloc(&p->d_arena, U.arena.total));
    p->arena_bytes = U.arena.total;
    for (auto &it : U.items)
        if (it.bytes) HIP_TRY(hipMemcpy(p->d_arena + it.off, it.src, it.bytes, hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(p->d_arena + o_normbits, 0, sizeof(unsigned int)));

    char *A = p->d_arena;
    auto u32p = [&](size_t o) { return reinterpret_cast<const uint32_t *>(A + o); };
    FactorDev &F = p->F;
    F.col_ptr = u32p(o_colptr);
    F.ctask_ptr = u32p(o_cptr);
    F.cln = reinterpret_cast<const uint2 *>(A + o_cln);
    F.ctgt = u32p(o_ctgt);
    F.cup = u32p(o_cup);
    F.ment = u32p(o_ment);
    F.ment_ptr = u32p(o_mptr);
    F.ment_long = u32p(o_mlong);
    F.msrc_ptr = u32p(o_msptr);
    F.msrc = reinterpret_cast<const MergeSrc *>(A + o_msrc);
    F.mrec = reinterpret_cast<const uint4 *>(A + o_mrec);
    F.dvd = reinterpret_cast<const uint2 *>(A + o_dvd);
    F.dvd_ptr = u32p(o_dptr);
    const uint32_t *upos_d = u32p(o_upos);
    (void)upos_d;
    F.num_lev = P.num_lev;
    F.n = n;
    F.nnz = P.nnz;
    // window items are indexed like the full tables: shift each base back
    // by the index of its first item (only indices >= that are used)
    auto shifted = [&](size_t off, size_t first, size_t item) { return (uintptr_t)(A + off) - first * item; };
    {
        std::vector<WinDev> wd(wins.size());
        for (size_t w = 0; w < wins.size(); ++w) {
            const Window &x = wins[w];
            WinDev &d = wd[w];
            d.l0 = x.l0;
            d.l1 = x.l1;
            d.nslot = (uint32_t)x.pos.size();
            d.pos = u32p(o_wpos[w]);
            d.cln = reinterpret_cast<const uint2 *>(shifted(o_wcln[w], (size_t)x.c0 * 64, sizeof(uint2)));
            d.ctgt = reinterpret_cast<const uint32_t *>(shifted(o_wctgt[w], x.g0, sizeof(uint32_t)));
            d.cup = reinterpret_cast<const uint32_t *>(shifted(o_wcup[w], x.g0, sizeof(uint32_t)));
            d.ment = reinterpret_cast<const uint32_t *>(shifted(o_wment[w], x.m0, sizeof(uint32_t)));
            d.msrc = reinterpret_cast<const MergeSrc *>(shifted(o_wmsrc[w], x.s0, sizeof(MergeSrc)));
            d.mrec = reinterpret_cast<const uint4 *>(shifted(o_wmrec[w], 2 * (size_t)x.m0, sizeof(uint4)));
            d.dvd = reinterpret_cast<const uint2 *>(shifted(o_wdvd[w], x.d0, sizeof(uint2)));
        }
        if (!wd.empty())
            HIP_TRY(hipMemcpy(A + o_wdev, wd.data(), sizeof(WinDev) * wd.size(), hipMemcpyHostToDevice));
        F.win = reinterpret_cast<const WinDev *>(A + o_wdev);
        F.nwin = (int32_t)wins.size();
        F.win_nslot = p->win_nslot;
    }

    auto tri = [&](int s, const uint32_t *beg, uint32_t begoff, const uint32_t *end, const uint32_t *idx,
                   const uint32_t *vpos, const uint32_t *dpos, int desc) {
        TriDev t{};
        t.beg = beg;
        t.begoff = begoff;
        t.end = end;
        t.idx = idx;
        t.vpos = vpos;
        t.dpos = dpos;
        t.desc = desc;
        t.rows = u32p(o_trows[s]);
        t.lvl_ptr = u32p(o_tptr[s]);
        t.lvl_long = u32p(o_tlong[s]);
        t.nlev = (int)p->tri[s].lvl_ptr.size() - 1;
        t.dtask = u32p(o_ttask[s]);
        t.ntask = (uint32_t)p->tri[s].tasks.size() - 1;
        return t;
    };
    const uint32_t *colptr = u32p(o_colptr), *rowidx = u32p(o_rowidx), *diag = u32p(o_diag);
    p->d_rowidx = rowidx;
    p->d_diag = diag;
    const uint32_t *crptr = u32p(o_crptr), *ccidx = u32p(o_ccidx), *cdiag = u32p(o_cdiag);
    SolveDev &S = p->S;
    S.lo = tri(0, crptr, 0, cdiag, ccidx, upos_d, nullptr, 0);
    S.up = tri(1, cdiag, 1, crptr + 1, ccidx, upos_d, diag, 1);
    S.tup = tri(2, colptr, 0, diag, rowidx, nullptr, diag, 0);
    S.tlo = tri(3, diag, 1, colptr + 1, rowidx, nullptr, nullptr, 0);
    // the same L and U sweeps, pipelined, over the per-matrix value copy
    auto df = [&](const DfHost &h, size_t ot, size_t orl, size_t oi, int div, uint32_t voff) {
        DfDev d{};
        d.task = reinterpret_cast<const uint4 *>(A + ot);
        d.rl = reinterpret_cast<const uint2 *>(A + orl);
        d.idx = u32p(oi);
        d.ntask = (uint32_t)h.task.size();
        d.nslots = (uint32_t)h.idx.size();
        d.div = div;
        d.voff = voff;
        return d;
    };
    S.lo_df = df(dlo, o_dlo_task, o_dlo_rl, o_dlo_idx, 0, 0);
    S.up_df = df(dup, o_dup_task, o_dup_rl, o_dup_idx, 1, (uint32_t)dlo.idx.size());
    S.tup_df = df(dtu, o_dtu_task, o_dtu_rl, o_dtu_idx, 1, 0);
    S.tlo_df = df(dtl, o_dtl_task, o_dtl_rl, o_dtl_idx, 0, tlo_voff);
    S.tdf_src = u32p(o_tdfsrc);
    S.tdf_nval = p->tdf_nval;
    S.df_src = u32p(o_dfsrc);
    S.df_nval = p->df_nval;
    S.df_cpy = reinterpret_cast<const uint2 *>(A + o_dfcpy);
    S.df_cpy_ptr = u32p(o_dfcptr);
    S.df_cpy_range = p->copy_range;
    p->tri_dataflow = p->solve_df_opt >= 0 && p->lds_cap >= 4 * (size_t)((n + 31) / 32);
    if (an) {
        S.row_perm = u32p(o_rp);
        S.row_perm_inv = u32p(o_rpi);
        S.col_perm = u32p(o_cp);
        S.col_perm_inv = u32p(o_cpi);
        S.row_scale = reinterpret_cast<const double *>(A + o_rs);
        S.col_scale_perm = reinterpret_cast<const double *>(A + o_csp);
        RefineDev &R = p->R;
        R.a_rptr = u32p(o_arptr);
        R.a_cidx = u32p(o_acidx);
        R.a_src = u32p(o_asrc);
        R.a_cptr = u32p(o_acptr);
        R.a_ridx = u32p(o_aridx);
        R.lu_src = u32p(o_lusrc);
        R.row_idx = rowidx;
        R.lu_col = u32p(o_lucol);
        R.a_rows = u32p(o_arows);
        R.a_cols = u32p(o_acols);
        R.a_nlong = p->a_nlong;
        R.a_nlong_t = p->a_nlong_t;
        R.rs_col = reinterpret_cast<const double *>(A + o_rscol);
        R.a_dst = u32p(o_adst);
        R.lu_fill = u32p(o_lufill);
        R.nfill = (uint32_t)lu_fill.size();
        R.a_rs = reinterpret_cast<const double *>(A + o_ars);
        R.a_cs = reinterpret_cast<const double *>(A + o_acs);
        R.csp = S.col_scale_perm;
        R.nnz_a = p->nnz_a;
    }
    MatDesc md{};
    md.f = F;
    md.s = S;
    md.r = p->R;
    md.n = n;
    md.nnz = P.nnz;
    md.krylov = (uint32_t)p->krylov;
    md.a_kmax = p->a_kmax;
    p->d_desc = reinterpret_cast<MatDesc *>(A + o_desc);
    HIP_TRY(hipMemcpy(p->d_desc, &md, sizeof(MatDesc), hipMemcpyHostToDevice));
    p->d_norm_bits = reinterpret_cast<unsigned int *>(A + o_normbits);
    p->d_pert = reinterpret_cast<float *>(A + o_pert);
    p->d_norms = reinterpret_cast<unsigned long long *>(A + o_norms);
    p->d_t32 = reinterpret_cast<float *>(A + o_t32);
    p->d_r64 = reinterpret_cast<double *>(A + o_r64);
    *out = p.release();
    return GLU_OK;
}

glu_pattern pattern_of(const Symbolic &s)
{
    glu_pattern P{};
    P.n = s.n;
    P.nnz = s.nnz;
    P.num_lev = s.num_lev;
    P.sym_c_ptr = s.sym_c_ptr.data();
    P.sym_r_idx = s.sym_r_idx.data();
    P.l_col_ptr = s.l_col_ptr.data();
    P.csr_r_ptr = s.csr_r_ptr.data();
    P.csr_c_idx = s.csr_c_idx.data();
    P.csr_diag_ptr = s.csr_diag_ptr.data();
    P.level_ptr = s.level_ptr.data();
    P.level_idx = s.level_idx.data();
    return P;
}

glu_status run_factor(glu_plan *p, float *d_val, int perturb, hipStream_t st)
{
    if (perturb) launch::pert(p->F, d_val, p->d_norm_bits, p->d_pert, st);
    for (const Segment &s : p->fsegs) {
        if (s.win >= 0)
            launch::factor_window(p->F, s.win, p->win_slots[s.win], d_val, p->d_pert, perturb, p->fused_threads, st);
        else if (s.fused)
            launch::factor_fused(p->F, d_val, p->d_pert, s.l0, s.l1, perturb, p->fused_threads, st);
        else
            launch::factor_grid(p->F, d_val, p->d_pert, s.l0, s.items, perturb, kGridThreads, st);
    }
    launch::factor_divide(p->F, d_val, p->d_pert, p->ndiv, perturb, st);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(GLU_ERR_DEVICE, std::string("factor launch: ") + hipGetErrorString(e));
    return GLU_OK;
}

void run_tri(glu_plan *p, const TriDev &T, const TriHost &H, const float *lu, float *x, hipStream_t st)
{
    for (const Segment &s : H.segs) {
        if (s.fused && p->tri_dataflow)
            launch::tri_df(T, lu, x, s.l0, s.l1, s.k0, s.k1, p->n, p->fused_threads, st);
        else if (s.fused)
            launch::tri_fused(T, lu, x, s.l0, s.l1, p->fused_threads, st);
        else
            launch::tri_grid(T, lu, x, s.l0, s.items, st);
    }
}

// t (scaled, permuted) -> LU sweeps in place
void run_sweeps(glu_plan *p, const float *lu, float *t, int transpose, hipStream_t st)
{
    if (!transpose) {
        run_tri(p, p->S.lo, p->tri[0], lu, t, st);
        run_tri(p, p->S.up, p->tri[1], lu, t, st);
    } else {
        run_tri(p, p->S.tup, p->tri[2], lu, t, st);
        run_tri(p, p->S.tlo, p->tri[3], lu, t, st);
    }
}

// Rows whose values a one-workgroup sweep with ready bits of `bits` bytes
// keeps in the rest of `cap` bytes of LDS: the last ones, when they are at
// least a quarter of the n rows, else none.
uint32_t split_rows(uint32_t n, size_t bits, size_t cap)
{
    const size_t keep = cap > bits ? std::min<size_t>(n, (cap - bits) / sizeof(float)) : 0;
    return 4 * keep >= n ? (uint32_t)keep : 0u;
}

// The richest sweep mode of the batch kernels whose LDS fits order n:
// 2 = ready bits and correction vector in LDS, 1 = ready bits only,
// 0 = level-synchronous sweeps.
int sweep_mode(uint32_t n, size_t cap)
{
    for (int mode = 2; mode > 0; --mode)
        if (launch::batch_lds_bytes(mode, n) <= cap) return mode;
    return 0;
}

glu_status ensure_batch(glu_plan *p, int batch)
{
    if (batch <= p->batch_cap) return GLU_OK;
    if (p->d_bt32) (void)hipFree(p->d_bt32);
    if (p->d_br64) (void)hipFree(p->d_br64);
    if (p->d_blucsr) (void)hipFree(p->d_blucsr);
    if (p->d_bnorms) (void)hipFree(p->d_bnorms);
    if (p->d_bkry) (void)hipFree(p->d_bkry);
    p->d_bkry = nullptr;
    p->d_bt32 = nullptr;
    p->d_br64 = nullptr;
    p->d_blucsr = nullptr;
    p->d_bnorms = nullptr;
    p->batch_cap = 0;
    HIP_TRY(hipMalloc(&p->d_bt32, sizeof(float) * (size_t)p->n * batch));
    HIP_TRY(hipMalloc(&p->d_br64, sizeof(double) * (size_t)p->n * batch));
    HIP_TRY(hipMalloc(&p->d_blucsr, sizeof(float) * (size_t)p->df_nval * batch));
    HIP_TRY(hipMalloc(&p->d_bnorms, 2 * sizeof(double) * (size_t)batch));
    if (p->krylov)
        HIP_TRY(hipMalloc(&p->d_bkry, sizeof(double) * (3 * (size_t)p->krylov + 1) * p->n * (size_t)batch));
    p->batch_cap = batch;
    return GLU_OK;
}

}  // namespace

extern "C" glu_status glu_plan_create(const glu_pattern *pat, const glu_plan_opts *opts, glu_plan **out)
{
    if (!pat || !out) return fail(GLU_ERR_ARGUMENT, "null argument");
    try {
        return build_plan(*pat, nullptr, opts, out);
    } catch (const std::bad_alloc &) {
        return fail(GLU_ERR_MEMORY, "out of host memory");
    }
}

extern "C" glu_status glu_plan_create_from_analysis(const glu_analysis *a, const glu_plan_opts *opts,
                                                    glu_plan **out)
{
    if (!a || !out) return fail(GLU_ERR_ARGUMENT, "null argument");
    glu_pattern P = pattern_of(a->sym);
    try {
        return build_plan(P, a, opts, out);
    } catch (const std::bad_alloc &) {
        return fail(GLU_ERR_MEMORY, "out of host memory");
    }
}

extern "C" void glu_plan_destroy(glu_plan *p) { delete p; }

extern "C" glu_status glu_plan_info(const glu_plan *p, int64_t info[12])
{
    if (!p || !info) return fail(GLU_ERR_ARGUMENT, "null argument");
    info[0] = p->n;
    info[1] = p->nnz;
    info[2] = p->nnz_a;
    info[3] = p->num_lev;
    info[4] = p->updates;
    info[5] = (int64_t)p->fsegs.size();
    info[6] = (int64_t)p->arena_bytes;
    info[7] = p->units;
    info[8] = p->nwin;
    info[9] = p->win_steps;
    info[10] = p->win_nslot;
    info[11] = p->batch_grid_lev;
    return GLU_OK;
}

namespace {

// Replays slot's graph on st, capturing `issue` (which enqueues the launch
// sequence on the stream it is given) first when the slot is empty or was
// captured for other buffers / another mode.
template <typename Issue>
glu_status replay(glu_plan *p, glu_plan::GraphSlot &slot, const void *k0, const void *k1, const void *k2, int mode,
                  hipStream_t st, Issue issue)
{
    if (!slot.exec || slot.k[0] != k0 || slot.k[1] != k1 || slot.k[2] != k2 || slot.mode != mode) {
        if (slot.exec) {
            HIP_TRY(hipStreamSynchronize(st));   // the old graph may still run
            HIP_TRY(hipGraphExecDestroy(slot.exec));
            slot.exec = nullptr;
        }
        if (!p->cap_stream) HIP_TRY(hipStreamCreateWithFlags(&p->cap_stream, hipStreamNonBlocking));
        HIP_TRY(hipStreamBeginCapture(p->cap_stream, hipStreamCaptureModeThreadLocal));
        glu_status s = issue(p->cap_stream);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(p->cap_stream, &g);
        if (s != GLU_OK) {
            if (g) (void)hipGraphDestroy(g);
            return s;
        }
        if (e != hipSuccess) return fail(GLU_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(e));
        e = hipGraphInstantiate(&slot.exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            slot.exec = nullptr;
            return fail(GLU_ERR_DEVICE, std::string("graph instantiate: ") + hipGetErrorString(e));
        }
        slot.k[0] = k0;
        slot.k[1] = k1;
        slot.k[2] = k2;
        slot.mode = mode;
    }
    HIP_TRY(hipGraphLaunch(slot.exec, st));
    return GLU_OK;
}

}  // namespace

extern "C" glu_status glu_factor(glu_plan *p, float *d_val, int perturb, void *stream)
{
    if (!p || !d_val) return fail(GLU_ERR_ARGUMENT, "null argument");
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (!p->use_graph) return run_factor(p, d_val, perturb, st);
    return replay(p, p->g_factor, d_val, nullptr, nullptr, perturb ? 1 : 0, st,
                  [&](hipStream_t cs) { return run_factor(p, d_val, perturb, cs); });
}

extern "C" glu_status glu_factor_host(glu_plan *p, float *h_val, int perturb)
{
    if (!p || !h_val) return fail(GLU_ERR_ARGUMENT, "null argument");
    if (!p->d_val) HIP_TRY(hipMalloc(&p->d_val, sizeof(float) * p->nnz));
    HIP_TRY(hipMemcpy(p->d_val, h_val, sizeof(float) * p->nnz, hipMemcpyHostToDevice));
    glu_status s = run_factor(p, p->d_val, perturb, nullptr);
    if (s != GLU_OK) return s;
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(h_val, p->d_val, sizeof(float) * p->nnz, hipMemcpyDeviceToHost));
    return GLU_OK;
}

extern "C" glu_status glu_solve(glu_plan *p, const float *d_lu, const float *d_b, float *d_x, int transpose,
                                void *stream)
{
    if (!p || !d_lu || !d_b || !d_x) return fail(GLU_ERR_ARGUMENT, "null argument");
    if (!p->has_scaling) return fail(GLU_ERR_ARGUMENT, "plan has no permutation/scaling (create it from an analysis)");
    hipStream_t st = static_cast<hipStream_t>(stream);
    // A x = b in one workgroup (exact dataflow sweeps) when the ready bits
    // fit in LDS and the dataflow solve is on
    const int mode = !transpose && p->tri_dataflow ? sweep_mode(p->n, launch::enable_lds()) : 0;
    if (mode != 0) {
        glu_status s = ensure_batch(p, 1);   // the sweep-layout copy (before any graph capture)
        if (s != GLU_OK) return s;
    }
    auto issue = [&](hipStream_t s) {
        const SolveDev &S = p->S;
        if (mode != 0) {
            // mode 1: the LDS the ready bits leave free holds the last rows
            // (r0..n-1) when they are a quarter of the rows or more
            const size_t bits = launch::batch_lds_bytes(mode, p->n), cap = launch::enable_lds();
            const uint32_t keep = mode == 1 ? split_rows(p->n, bits, cap) : 0u;
            launch::solve_one(p->d_desc, p->df_nval, d_lu, p->d_blucsr, d_b, d_x, p->d_bt32,
                              keep ? p->n - keep : UINT32_MAX, p->fused_threads, mode, bits + sizeof(float) * keep, s);
        } else if (!transpose) {
            launch::perm_scale_f32(d_b, p->d_t32, S.col_perm, S.col_scale_perm, 0, p->n, s);
            run_sweeps(p, d_lu, p->d_t32, 0, s);
            launch::perm_scale_f32(p->d_t32, d_x, S.row_perm_inv, S.row_scale, 0, p->n, s);
        } else {
            launch::perm_scale_f32(d_b, p->d_t32, S.row_perm, S.row_scale, 1, p->n, s);
            run_sweeps(p, d_lu, p->d_t32, 1, s);
            launch::perm_scale_f32(p->d_t32, d_x, S.col_perm_inv, S.col_scale_perm, 1, p->n, s);
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return fail(GLU_ERR_DEVICE, std::string("solve launch: ") + hipGetErrorString(e));
        return GLU_OK;
    };
    if (!p->use_graph) return issue(st);
    return replay(p, p->g_solve[transpose ? 1 : 0], d_lu, d_b, d_x, 0, st, issue);
}

extern "C" glu_status glu_refresh(glu_plan *p, const double *d_ax, float *d_lu, int batch, void *stream)
{
    if (!p || !d_ax || !d_lu || batch <= 0) return fail(GLU_ERR_ARGUMENT, "bad argument");
    if (!p->has_refine) return fail(GLU_ERR_ARGUMENT, "plan has no analysis");
    launch::refresh(p->R, d_ax, d_lu, p->nnz, batch, static_cast<hipStream_t>(stream));
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(GLU_ERR_DEVICE, std::string("refresh launch: ") + hipGetErrorString(e));
    return GLU_OK;
}

static double bits_to_double(unsigned long long b)
{
    double d;
    std::memcpy(&d, &b, sizeof d);
    return d;
}

namespace {

// Plain iterative refinement driven from the host, one preconditioner
// solve per iteration: x = 0, r = b; then t = float(r scaled into the
// factors' frame), the two sweeps on t, x += t scaled back, r = b - A x with
// |r|, |x| (and |A|, |b| on the first pass), eta on the host.  The FP64 work
// runs in grid kernels; the sweeps run in one workgroup over the
// sweep-layout copy `lucsr` (dataflow, ready bits in LDS) or, when lucsr is
// null, level by level over the factors.
glu_status refine_host_loop(glu_plan *p, const double *d_ax, const float *d_lu, const double *d_b, double *d_x,
                            double tol, int maxit, int transpose, const float *lucsr, float *t32, double *r64,
                            unsigned long long *norms, int *h_iters, double *h_eta, hipStream_t st)
{
    const SolveDev &S = p->S;
    const uint32_t n = p->n;
    // the sweeps keep the results of the last rows in the LDS their ready
    // bits leave free (rows r0..n-1), when that is a quarter of the rows or
    // more (below that the test per access costs more than it saves)
    const size_t bits = launch::batch_lds_bytes(1, n), cap = launch::enable_lds();
    const uint32_t keep = split_rows(n, bits, cap), r0 = keep ? n - keep : UINT32_MAX;
    HIP_TRY(hipMemsetAsync(d_x, 0, sizeof(double) * n, st));
    HIP_TRY(hipMemcpyAsync(r64, d_b, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    double anorm = 0.0, bnorm = 0.0, eta = 0.0;
    int it = 0;
    for (;;) {
        if (!transpose)
            launch::perm_scale_f64(r64, t32, S.col_perm, S.col_scale_perm, 0, n, st);
        else
            launch::perm_scale_f64(r64, t32, S.row_perm, S.row_scale, 1, n, st);
        if (lucsr)
            launch::sweeps_df(p->d_desc, lucsr, t32, r0, p->fused_threads, bits + sizeof(float) * keep, transpose, st);
        else
            run_sweeps(p, d_lu, t32, transpose, st);
        if (!transpose)
            launch::update_x(t32, d_x, S.row_perm_inv, S.row_scale, 0, n, st);
        else
            launch::update_x(t32, d_x, S.col_perm_inv, S.col_scale_perm, 1, n, st);
        launch::zero_u64(norms, 4, st);
        const uint32_t nl = transpose ? p->a_nlong_t : p->a_nlong;
        launch::residual(p->R, d_ax, d_b, d_x, r64, n, transpose, it == 0, norms, (n - nl) + nl * 64, st);
        HIP_TRY(hipGetLastError());
        unsigned long long nb[4];
        HIP_TRY(hipMemcpyAsync(nb, norms, sizeof nb, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        if (it == 0) {
            anorm = bits_to_double(nb[2]);
            bnorm = bits_to_double(nb[3]);
        }
        const double rn = bits_to_double(nb[0]), xn = bits_to_double(nb[1]);
        eta = rn / (anorm * xn + bnorm);
        ++it;
        if (!(eta > tol) || it >= maxit) break;
    }
    if (h_iters) *h_iters = it;
    if (h_eta) *h_eta = eta;
    if (!(eta == eta)) return fail(GLU_ERR_NUMERIC, "refinement produced NaN");
    return GLU_OK;
}

// The one-workgroup refinement (sweeps, FGMRES and the products with A in
// one workgroup; sweep vector and ready bits in LDS), then its iteration
// count and eta read back.
glu_status refine_one_wg(glu_plan *p, const double *d_ax, const float *d_lu, const double *d_b, double *d_x,
                         double tol, int maxit, int transpose, float *lucsr, float *t32, double *r64,
                         double *bnorms, double *kry, char *out16, int *h_iters, double *h_eta,
                         hipStream_t st)
{
    const uint32_t n = p->n;
    int32_t *d_it = reinterpret_cast<int32_t *>(out16);
    double *d_eta = reinterpret_cast<double *>(out16 + 8);
    launch::refine_one(p->d_desc, transpose ? p->tdf_nval : p->df_nval, d_ax, d_b, d_x, d_lu, lucsr, t32, r64,
                       bnorms, kry, d_it, d_eta, tol, maxit, p->fused_threads, launch::batch_lds_bytes(2, n),
                       transpose, st);
    HIP_TRY(hipGetLastError());
    int32_t it = 0;
    double eta = 0.0;
    HIP_TRY(hipMemcpyAsync(&it, d_it, sizeof it, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(&eta, d_eta, sizeof eta, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (h_iters) *h_iters = it;
    if (h_eta) *h_eta = eta;
    if (!(eta == eta)) return fail(GLU_ERR_NUMERIC, "refinement produced NaN");
    return GLU_OK;
}

}  // namespace

extern "C" glu_status glu_refine(glu_plan *p, const double *d_ax, const float *d_lu, const double *d_b,
                                 double *d_x, double tol, int maxit, int transpose, int *h_iters, double *h_eta,
                                 void *stream)
{
    if (!p || !d_ax || !d_lu || !d_b || !d_x || maxit <= 0) return fail(GLU_ERR_ARGUMENT, "bad argument");
    if (!p->has_refine) return fail(GLU_ERR_ARGUMENT, "plan has no analysis");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const uint32_t n = p->n;
    // 2: the sweep vector fits in LDS -> everything in one workgroup (FGMRES,
    // plan.krylov); 1: only the ready bits fit -> dataflow sweeps in one
    // workgroup, FP64 work across the grid (plain refinement); 0: level by
    // level.
    const int mode = sweep_mode(n, launch::enable_lds());
    float *lucsr = nullptr;
    if (mode != 0) {
        glu_status s = ensure_batch(p, 1);
        if (s != GLU_OK) return s;
        lucsr = p->d_blucsr;
        if (transpose) {
            if (!p->d_tlucsr) HIP_TRY(hipMalloc(&p->d_tlucsr, sizeof(float) * p->tdf_nval));
            lucsr = p->d_tlucsr;
        }
    }
    if (mode == 2) {
        if (!p->d_rout) HIP_TRY(hipMalloc(&p->d_rout, 16));
        return refine_one_wg(p, d_ax, d_lu, d_b, d_x, tol, maxit, transpose, lucsr, p->d_bt32, p->d_br64,
                             p->d_bnorms, p->d_bkry, p->d_rout, h_iters, h_eta, st);
    }
    if (mode == 1) launch::copy_df(p->d_desc, transpose ? p->tdf_nval : p->df_nval, d_lu, lucsr, transpose, st);
    return refine_host_loop(p, d_ax, d_lu, d_b, d_x, tol, maxit, transpose, lucsr, p->d_t32, p->d_r64, p->d_norms,
                            h_iters, h_eta, st);
}

// ---------------------------------------------------------------------------
// glu_refine_noalloc: glu_refine's one-workgroup path with its scratch in a
// file-scope device array instead of plan buffers grown on first use.
// ---------------------------------------------------------------------------

namespace {

constexpr size_t kScratchBytes = GLU_NOALLOC_SCRATCH_BYTES;
constexpr int kScratchDevices = 64;

// One copy per device, held until the process ends.  Host code never uses
// this name as an address: each device's copy is reached through the
// pointer hipGetSymbolAddress gives for it (scratch_base).
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[kScratchDevices] = {};

// The calling thread's current device's copy of `scratch`, looked up once
// per device (with that device current).
glu_status scratch_base(unsigned char **out)
{
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kScratchDevices) return fail(GLU_ERR_DEVICE, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(GLU_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<unsigned char *>(ptr);
    }
    *out = g_scratch_ptr[dev];
    return GLU_OK;
}

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

extern "C" glu_status glu_refine_noalloc(glu_plan *p, const double *d_ax, const float *d_lu, const double *d_b,
                                         double *d_x, double tol, int maxit, int transpose, int *h_iters,
                                         double *h_eta, void *stream)
{
    if (!p || !d_ax || !d_lu || !d_b || !d_x || maxit <= 0) return fail(GLU_ERR_ARGUMENT, "bad argument");
    if (!p->has_refine) return fail(GLU_ERR_ARGUMENT, "plan has no analysis");
    const uint32_t n = p->n;
    const int mode = sweep_mode(n, launch::enable_lds());
    if (mode == 0)
        return fail(GLU_ERR_ARGUMENT, "order " + std::to_string(n) +
                                          " is beyond the dataflow sweeps (their ready bits exceed LDS): "
                                          "use glu_refine");
    // scratch layout: sweep value copy, correction vector, residual, norms
    // (one-workgroup path: FP64 |b| slots, FGMRES vectors, iteration count +
    // eta; grid path: the four norm words)
    const uint32_t nval = transpose ? p->tdf_nval : p->df_nval;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off = align256(off + bytes);
        return o;
    };
    const size_t o_val = take(sizeof(float) * (size_t)nval);
    const size_t o_t32 = take(sizeof(float) * (size_t)n);
    const size_t o_r64 = take(sizeof(double) * (size_t)n);
    const size_t o_norms = take(4 * sizeof(unsigned long long));
    const size_t o_kry = take(mode == 2 && p->krylov ? sizeof(double) * (3 * (size_t)p->krylov + 1) * n : 0);
    const size_t o_out = take(16);
    if (off > kScratchBytes)
        return fail(GLU_ERR_ARGUMENT, "refinement needs " + std::to_string(off) + " bytes of scratch; the static " +
                                          "scratch holds " + std::to_string(kScratchBytes) + ": use glu_refine");
    unsigned char *base = nullptr;
    glu_status s = scratch_base(&base);
    if (s != GLU_OK) return s;
    hipStream_t st = static_cast<hipStream_t>(stream);
    float *lucsr = reinterpret_cast<float *>(base + o_val);
    float *t32 = reinterpret_cast<float *>(base + o_t32);
    double *r64 = reinterpret_cast<double *>(base + o_r64);
    if (mode == 2)   // as glu_refine: everything in one workgroup
        return refine_one_wg(p, d_ax, d_lu, d_b, d_x, tol, maxit, transpose, lucsr, t32, r64,
                             reinterpret_cast<double *>(base + o_norms),
                             p->krylov ? reinterpret_cast<double *>(base + o_kry) : nullptr,
                             reinterpret_cast<char *>(base + o_out), h_iters, h_eta, st);
    launch::copy_df(p->d_desc, nval, d_lu, lucsr, transpose, st);   // as glu_refine: sweeps in one workgroup
    return refine_host_loop(p, d_ax, d_lu, d_b, d_x, tol, maxit, transpose, lucsr, t32, r64,
                            reinterpret_cast<unsigned long long *>(base + o_norms), h_iters, h_eta, st);
}

extern "C" glu_status glu_adjoint_values(glu_plan *p, const double *d_lam, const double *d_x, double *d_grad_ax,
                                         void *stream)
{
    if (!p || !d_lam || !d_x || !d_grad_ax) return fail(GLU_ERR_ARGUMENT, "null argument");
    if (!p->has_refine) return fail(GLU_ERR_ARGUMENT, "plan has no analysis");
    launch::adjoint_values(p->R, d_lam, d_x, d_grad_ax, p->n, static_cast<hipStream_t>(stream));
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(GLU_ERR_DEVICE, std::string("adjoint launch: ") + hipGetErrorString(e));
    return GLU_OK;
}

namespace {

// Scratch of the interleaved steps for `batch` matrices (grown on demand;
// the GESP maxima start at zero and every call leaves them so).
glu_status ensure_il(glu_plan *p, int batch)
{
    if (p->batch_grid_lev == 0 || batch <= p->il_cap) return GLU_OK;
    for (void *q : {(void *)p->d_il, (void *)p->d_il_bits, (void *)p->d_fnorms})
        if (q) (void)hipFree(q);
    p->d_il = nullptr;
    p->d_il_bits = nullptr;
    p->d_fnorms = nullptr;
    p->il_cap = 0;
    const size_t padded = ((size_t)batch + launch::il_group() - 1) / launch::il_group() * launch::il_group();
    HIP_TRY(hipMalloc(&p->d_il, sizeof(float) * padded * p->nnz));
    HIP_TRY(hipMalloc(&p->d_il_bits, sizeof(unsigned int) * (size_t)batch));
    HIP_TRY(hipMalloc(&p->d_fnorms, 2 * sizeof(double) * (size_t)batch));
    HIP_TRY(hipMemset(p->d_il_bits, 0, sizeof(unsigned int) * (size_t)batch));
    p->il_cap = batch;
    return GLU_OK;
}

// The batched kernels' first steps: the values of every matrix into the
// interleaved copy (with each GESP value in norms[2 m]); steps
// [0, batch_grid_lev) as one grid launch each; the factors so far into
// d_lu.  The per-matrix workgroups go on from there (with a right-hand
// side, from |b|_inf into norms[2 m + 1], x = 0 and r = b).
void batch_grid_steps(glu_plan *p, int batch, const double *d_ax, float *d_lu, double *norms, int perturb,
                      hipStream_t st)
{
    if (p->batch_grid_lev == 0) return;
    launch::il_values(p->R, p->F, d_ax, p->d_il, p->d_il_bits, norms, p->nnz, p->n, batch, perturb, st);
    for (int l = 0; l < p->batch_grid_lev; ++l)
        launch::factor_grid_il(p->F, p->d_il, norms, l, p->grid_waves[l], batch, perturb, st);
    launch::il_to_rows(p->d_il, d_lu, p->nnz, batch, st);
}

}  // namespace

extern "C" glu_status glu_batch_solve(glu_plan *p, int batch, const double *d_ax, const double *d_b, double *d_x,
                                      float *d_lu, double tol, int maxit, int32_t *d_iters, double *d_eta,
                                      unsigned flags, void *stream)
{
    if (!p || batch <= 0 || !d_ax || !d_b || !d_x || !d_lu || !d_iters || !d_eta || maxit <= 0)
        return fail(GLU_ERR_ARGUMENT, "bad argument");
    if (!p->has_refine) return fail(GLU_ERR_ARGUMENT, "plan has no analysis");
    glu_status s = ensure_batch(p, batch);
    if (s == GLU_OK) s = ensure_il(p, batch);
    if (s != GLU_OK) return s;
    const size_t cap = launch::enable_lds();
    const int mode = sweep_mode(p->n, cap);
    const int perturb = (int)(flags & 1u);
    hipStream_t st = static_cast<hipStream_t>(stream);
    batch_grid_steps(p, batch, d_ax, d_lu, p->d_bnorms, perturb, st);
    launch::batch_same(p->d_desc, batch, d_ax, d_b, d_x, d_lu, p->d_blucsr, p->d_bt32, p->d_br64, p->d_bnorms,
                       p->d_bkry, d_iters, d_eta, tol, maxit, perturb, p->fused_threads, mode,
                       sizeof(float) * std::max(p->win_nslot, p->copy_range), launch::batch_lds_bytes(mode, p->n),
                       p->batch_grid_lev, st);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(GLU_ERR_DEVICE, std::string("batch launch: ") + hipGetErrorString(e));
    return GLU_OK;
}

extern "C" glu_status glu_batch_factor(glu_plan *p, int batch, const double *d_ax, float *d_lu, unsigned flags,
                                       void *stream)
{
    if (!p || batch <= 0 || !d_ax || !d_lu) return fail(GLU_ERR_ARGUMENT, "bad argument");
    if (!p->has_refine) return fail(GLU_ERR_ARGUMENT, "plan has no analysis");
    (void)launch::enable_lds();
    const int perturb = (int)(flags & 1u);
    hipStream_t st = static_cast<hipStream_t>(stream);
    glu_status s = ensure_il(p, batch);
    if (s != GLU_OK) return s;
    batch_grid_steps(p, batch, d_ax, d_lu, p->d_fnorms, perturb, st);
    launch::batch_factor(p->d_desc, batch, d_ax, d_lu, p->d_fnorms, perturb, p->fused_threads,
                         sizeof(float) * p->win_nslot, p->batch_grid_lev, st);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(GLU_ERR_DEVICE, std::string("batch factor launch: ") + hipGetErrorString(e));
    return GLU_OK;
}

extern "C" glu_status glu_abft_sums(glu_plan *p, int batch, const float *d_val, float *d_cca, void *stream)
{
    if (!p || batch <= 0 || !d_val || !d_cca) return fail(GLU_ERR_ARGUMENT, "bad argument");
    if (batch > 65535) return fail(GLU_ERR_ARGUMENT, "batch above 65535");
    launch::abft_cca(p->F, d_val, d_cca, batch, static_cast<hipStream_t>(stream));
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(GLU_ERR_DEVICE, std::string("abft launch: ") + hipGetErrorString(e));
    return GLU_OK;
}

extern "C" glu_status glu_abft_check(glu_plan *p, int batch, const float *d_lu, const float *d_cca, float tol,
                                     int32_t *d_first_bad, float *d_max_diff, void *stream)
{
    if (!p || batch <= 0 || !d_lu || !d_cca || !d_first_bad || !d_max_diff || !(tol >= 0.f))
        return fail(GLU_ERR_ARGUMENT, "bad argument");
    if (batch > 65535) return fail(GLU_ERR_ARGUMENT, "batch above 65535");
    if (batch > p->abft_cap) {
        if (p->d_abft) (void)hipFree(p->d_abft);
        p->d_abft = nullptr;
        p->abft_cap = 0;
        HIP_TRY(hipMalloc(&p->d_abft, sizeof(float) * (size_t)p->n * batch));
        p->abft_cap = batch;
    }
    launch::abft_check(p->F, p->d_rowidx, p->d_diag, d_lu, d_cca, p->d_abft, tol, d_first_bad, d_max_diff, batch,
                       static_cast<hipStream_t>(stream));
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(GLU_ERR_DEVICE, std::string("abft launch: ") + hipGetErrorString(e));
    return GLU_OK;
}

// ---------------------------------------------------------------------------
// uneven groups
// ---------------------------------------------------------------------------

struct MatIOHost {            // mirrors the device-side MatIO layout
    const double *ax;
    const double *b;
    double *x;
    float *lu;
    float *t32;
    double *r64;
    float *lucsr;
    double *norms;
    double *kry;
    int32_t *iters;
    double *eta;
};

struct glu_group {
    std::vector<glu_plan *> plans;
    MatDesc *d_descs = nullptr;
    MatIOHost *d_ios = nullptr;
    uint32_t *d_order = nullptr;
    float *d_t32 = nullptr;
    double *d_r64 = nullptr;
    float *d_lucsr = nullptr;
    double *d_norms = nullptr;
    double *d_kry = nullptr;
    std::vector<size_t> voff;   // per-matrix offset into the scratch vectors
    std::vector<size_t> koff;   // per-matrix offset into d_kry (~0: plain refinement)
    std::vector<size_t> foff;   // per-matrix offset into d_lucsr
    std::vector<uint32_t> order;   // members, heaviest first (d_order)
    int threads = kDefaultFusedThreads;
    int split_req = -1;            // heaviest members on the single-matrix path (-1: automatic)
    hipStream_t aux[kMaxGroupSplit] = {};
    hipEvent_t ev_in = nullptr, ev_aux[kMaxGroupSplit] = {};
    ~glu_group()
    {
        for (int j = 0; j < kMaxGroupSplit; ++j) {
            if (aux[j]) (void)hipStreamDestroy(aux[j]);
            if (ev_aux[j]) (void)hipEventDestroy(ev_aux[j]);
        }
        if (ev_in) (void)hipEventDestroy(ev_in);
        if (d_lucsr) (void)hipFree(d_lucsr);
        if (d_norms) (void)hipFree(d_norms);
        if (d_kry) (void)hipFree(d_kry);
        if (d_descs) (void)hipFree(d_descs);
        if (d_ios) (void)hipFree(d_ios);
        if (d_order) (void)hipFree(d_order);
        if (d_t32) (void)hipFree(d_t32);
        if (d_r64) (void)hipFree(d_r64);
    }
};

extern "C" glu_status glu_group_create(glu_plan *const *plans, int count, glu_group **out)
{
    if (!plans || count <= 0 || !out) return fail(GLU_ERR_ARGUMENT, "bad argument");
    *out = nullptr;
    if (launch::matio_size() != sizeof(MatIOHost)) return fail(GLU_ERR_GENERAL, "MatIO layout mismatch");
    std::unique_ptr<glu_group> g(new glu_group());
    std::vector<MatDesc> descs(count);
    std::vector<std::pair<int64_t, uint32_t>> work(count);
    size_t vtot = 0, ftot = 0, ktot = 0;
    for (int i = 0; i < count; ++i) {
        glu_plan *p = plans[i];
        if (!p || !p->has_refine) return fail(GLU_ERR_ARGUMENT, "every plan must come from an analysis");
        g->plans.push_back(p);
        descs[i].f = p->F;
        descs[i].s = p->S;
        descs[i].r = p->R;
        descs[i].n = p->n;
        descs[i].nnz = p->nnz;
        descs[i].krylov = (uint32_t)p->krylov;
        descs[i].a_kmax = p->a_kmax;
        g->koff.push_back(p->krylov ? ktot : ~(size_t)0);
        ktot += (3 * (size_t)p->krylov + 1) * p->n * (p->krylov != 0);
        g->voff.push_back(vtot);
        vtot += p->n;
        g->foff.push_back(ftot);
        ftot += p->df_nval;
        g->threads = std::max(g->threads, p->fused_threads);
        // heaviest first: levels dominate the latency, updates the work
        work[i] = {-(int64_t)(p->num_lev + p->tri[0].lvl_ptr.size() + p->tri[1].lvl_ptr.size()) * 64 - p->updates,
                   (uint32_t)i};
    }
    std::sort(work.begin(), work.end());
    std::vector<uint32_t> &order = g->order;
    order.resize(count);
    for (int i = 0; i < count; ++i) order[i] = work[i].second;
    HIP_TRY(hipMalloc(&g->d_descs, sizeof(MatDesc) * count));
    HIP_TRY(hipMalloc(&g->d_ios, sizeof(MatIOHost) * count));
    HIP_TRY(hipMalloc(&g->d_order, sizeof(uint32_t) * count));
    HIP_TRY(hipMalloc(&g->d_t32, sizeof(float) * vtot));
    HIP_TRY(hipMalloc(&g->d_r64, sizeof(double) * vtot));
    HIP_TRY(hipMalloc(&g->d_lucsr, sizeof(float) * ftot));
    HIP_TRY(hipMalloc(&g->d_norms, 2 * sizeof(double) * count));
    if (ktot) HIP_TRY(hipMalloc(&g->d_kry, sizeof(double) * ktot));
    HIP_TRY(hipMemcpy(g->d_descs, descs.data(), sizeof(MatDesc) * count, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(g->d_order, order.data(), sizeof(uint32_t) * count, hipMemcpyHostToDevice));
    *out = g.release();
    return GLU_OK;
}

extern "C" void glu_group_destroy(glu_group *g) { delete g; }

extern "C" glu_status glu_group_set_split(glu_group *g, int count)
{
    if (!g || count < -1 || count > kMaxGroupSplit) return fail(GLU_ERR_ARGUMENT, "split count must be -1..2");
    g->split_req = count;
    return GLU_OK;
}

extern "C" glu_status glu_group_solve(glu_group *g, const double *const *d_ax, const double *const *d_b,
                                      double *const *d_x, float *const *d_lu, double tol, int maxit,
                                      int32_t *d_iters, double *d_eta, unsigned flags, void *stream)
{
    if (!g || !d_ax || !d_b || !d_x || !d_lu || !d_iters || !d_eta || maxit <= 0)
        return fail(GLU_ERR_ARGUMENT, "bad argument");
    const int count = (int)g->plans.size();
    std::vector<MatIOHost> ios(count);
    for (int i = 0; i < count; ++i) {
        if (!d_ax[i] || !d_b[i] || !d_x[i] || !d_lu[i]) return fail(GLU_ERR_ARGUMENT, "null matrix pointer");
        ios[i] = MatIOHost{d_ax[i], d_b[i], d_x[i], d_lu[i], g->d_t32 + g->voff[i], g->d_r64 + g->voff[i],
                           g->d_lucsr + g->foff[i], g->d_norms + 2 * i,
                           g->koff[i] == ~(size_t)0 ? nullptr : g->d_kry + g->koff[i], d_iters + i, d_eta + i};
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    HIP_TRY(hipMemcpyAsync(g->d_ios, ios.data(), sizeof(MatIOHost) * count, hipMemcpyHostToDevice, st));
    const size_t cap = launch::enable_lds();
    const int perturb = (int)(flags & 1u);
    // The heaviest members (as many as asked; automatic: one in a group of
    // four or more) run on the single-matrix path on streams of their own,
    // alongside the fused launch of the rest: the grid factorization on the
    // CUs the fused launch leaves idle, then the one-workgroup refinement,
    // the same kernel the fused launch runs (same workgroup size and restart
    // length), so the results are the same bits.  Only members whose
    // refinement keeps its vector in LDS qualify.
    int want = g->split_req >= 0 ? g->split_req : (count >= 4 ? 1 : 0);
    want = std::min(want, count - 1);
    int nsplit = 0;
    while (nsplit < want && sweep_mode(g->plans[g->order[nsplit]]->n, cap) == 2) ++nsplit;
    if (nsplit > 0) {
        if (!g->ev_in) HIP_TRY(hipEventCreateWithFlags(&g->ev_in, hipEventDisableTiming));
        HIP_TRY(hipEventRecord(g->ev_in, st));
    }
    for (int j = 0; j < nsplit; ++j) {
        const uint32_t m = g->order[j];
        glu_plan *p = g->plans[m];
        glu_status s = ensure_batch(p, 1);
        if (s != GLU_OK) return s;
        if (!g->aux[j]) HIP_TRY(hipStreamCreateWithFlags(&g->aux[j], hipStreamNonBlocking));
        if (!g->ev_aux[j]) HIP_TRY(hipEventCreateWithFlags(&g->ev_aux[j], hipEventDisableTiming));
        hipStream_t as = g->aux[j];
        HIP_TRY(hipStreamWaitEvent(as, g->ev_in, 0));
        launch::refresh(p->R, d_ax[m], d_lu[m], p->nnz, 1, as);
        s = run_factor(p, d_lu[m], perturb, as);
        if (s != GLU_OK) return s;
        launch::refine_one(p->d_desc, p->df_nval, d_ax[m], d_b[m], d_x[m], d_lu[m], p->d_blucsr, p->d_bt32,
                           p->d_br64, p->d_bnorms, p->d_bkry, d_iters + m, d_eta + m, tol, maxit, g->threads,
                           launch::batch_lds_bytes(2, p->n), 0, as);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(g->ev_aux[j], as));
    }
    size_t nmax = 0, tmax = 0;
    for (int i = nsplit; i < count; ++i) {
        const glu_plan *p = g->plans[g->order[i]];
        nmax = std::max<size_t>(nmax, p->n);
        tmax = std::max<size_t>(tmax, sizeof(float) * std::max(p->win_nslot, p->copy_range));
    }
    const int mode = sweep_mode((uint32_t)nmax, cap);
    // ready bits only for the largest order: each matrix whose vector fits
    // keeps it in LDS too (the launch's LDS: the largest such need)
    size_t ref_lds = launch::batch_lds_bytes(mode, (uint32_t)nmax);
    if (mode == 1)
        for (int i = nsplit; i < count; ++i) {
            const glu_plan *p = g->plans[g->order[i]];
            const size_t full = launch::batch_lds_bytes(2, p->n);
            ref_lds = std::max(ref_lds, full <= cap ? full : launch::batch_lds_bytes(1, p->n));
        }
    launch::batch_group(g->d_descs, g->d_ios, g->d_order + nsplit, count - nsplit, tol, maxit, perturb, g->threads,
                        mode, tmax, ref_lds, st);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(GLU_ERR_DEVICE, std::string("group launch: ") + hipGetErrorString(e));
    for (int j = 0; j < nsplit; ++j) HIP_TRY(hipStreamWaitEvent(st, g->ev_aux[j], 0));
    // ios is read by the kernel from device memory; the host copy may go.
    HIP_TRY(hipStreamSynchronize(st));
    return GLU_OK;
}
