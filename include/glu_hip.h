/*
 * glu_hip.h — C-ABI of the MI355X (gfx950) numeric phase for GLU 3.0.
 *
 * Plain C types only (no HIP, torch or C++ types in the signatures); device
 * buffers are passed as plain pointers obtained from hipMalloc (or any
 * allocator of device memory), streams as an opaque `void *` holding a
 * hipStream_t (NULL = the null stream).  Every entry point returns a
 * glu_status; glu_last_error() gives the thread's last error text.
 *
 * Reference interfaces replaced (paths relative to the reference tree):
 *   preprocess_from_arrays + Symbolic_Matrix::{fill_in,csr,predictLU,leveling}
 *       src/preprocess/preprocess_arrays.c:18, src/symbolic.cc:11-151,
 *       as called in sequence by src/glu_binding.cpp:87-109  -> glu_analyze
 *   LUonDevice(Symbolic_Matrix&, ostream&, ostream&, bool PERTURB)
 *       include/numeric.h:6, src/numeric.cu:347-639
 *       -> glu_plan_create + glu_factor_host (+ glu_plan_destroy)
 *   Symbolic_Matrix::solve(SNicsLU*, const vector<REAL>&)
 *       src/symbolic.cc:200-251                             -> glu_solve
 *   Symbolic_Matrix::ABFTCalculateCCA / ABFTCheckResult
 *       src/symbolic.cc:163-197 (GLU_DEBUG)  -> glu_abft_sums / glu_abft_check
 *   GLUFactorization(data, indices, indptr, n, perturb) / .solve(b)
 *       src/glu_binding.cpp:44-155 (pybind11 module _pyglu) -> glu_analyze +
 *       glu_plan_create_from_analysis + glu_factor + glu_solve
 *   No counterpart in the reference (the refactor semantics follow NICSLU's
 *   _I_NicsLU_MC64ScaleForRefact, src/nicslu/source/nicslu_mc64scale.c:58-85):
 *       glu_refresh, glu_refine, glu_adjoint_values, glu_batch_solve, glu_batch_factor,
 *       glu_group_*.
 */
#ifndef GLU_HIP_H
#define GLU_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum glu_status {
    GLU_OK = 0,
    GLU_ERR_GENERAL = -1,
    GLU_ERR_ARGUMENT = -2,      /* bad pointer / size / option */
    GLU_ERR_MEMORY = -3,        /* host or device allocation failed */
    GLU_ERR_DEVICE = -4,        /* HIP runtime error or no device */
    GLU_ERR_STRUCT_SINGULAR = -5, /* MC64 found no perfect matching (nicslu.h:16) */
    GLU_ERR_NUMERIC = -6,       /* non-finite factors / no convergence */
    GLU_ERR_INVALID = -7,       /* malformed CSC (nicslu.h:18) */
    GLU_ERR_DUPLICATE = -8,     /* duplicated entry (nicslu.h:19) */
    GLU_ERR_EMPTY = -9          /* n == 0 or nnz == 0 (numeric.cu:349-352) */
} glu_status;

/* Thread-local text of the last error on this thread ("" if none). */
const char *glu_last_error(void);

/* ------------------------------------------------------------------------
 * Host analysis: MC64 + AMD + permute/scale (NICSLU NicsLU_Analyze and
 * my_DumpA), then fill-in, CSR, predictLU and leveling.  Bit-exact with the
 * reference on the same input.  `ax`/`ai`/`ap` are the caller's CSC arrays
 * (ap[n] == nnz); they are copied.
 * ---------------------------------------------------------------------- */
typedef struct glu_analysis glu_analysis;

glu_status glu_analyze(uint32_t n, uint32_t nnz, const double *ax, const uint32_t *ai,
                       const uint32_t *ap, glu_analysis **out);
void glu_analysis_destroy(glu_analysis *a);

/* Independent matrices analysed on `threads` host threads (<= 0: one per
 * core), each exactly as glu_analyze (same results); out[i] / status[i] per
 * matrix.  Returns GLU_OK, or the first failing matrix's status with its
 * index in glu_last_error().  No counterpart in the reference: the
 * patterns of an uneven batch are independent. */
glu_status glu_analyze_many(int count, const uint32_t *n, const uint32_t *nnz, const double *const *ax,
                            const uint32_t *const *ai, const uint32_t *const *ap, int threads,
                            glu_analysis **out, glu_status *status);
/* info[0]=n info[1]=nnz(A) info[2]=nnz(L+U) info[3]=num_lev
 * info[4]=scalar updates of one factorization  info[5]=L-solve levels
 * info[6]=U-solve levels  info[7]=reserved */
glu_status glu_analysis_info(const glu_analysis *a, int64_t info[8]);

/* Copy one array of the analysis into `dst` (`bytes` must equal its size). */
typedef enum glu_field {
    GLU_F_SYM_C_PTR = 0,   /* u32[n+1] */
    GLU_F_SYM_R_IDX,       /* u32[nnz_lu] */
    GLU_F_L_COL_PTR,       /* u32[n]  (diagonal positions) */
    GLU_F_CSR_R_PTR,       /* u32[n+1] */
    GLU_F_CSR_C_IDX,       /* u32[nnz_lu] */
    GLU_F_CSR_DIAG_PTR,    /* u32[n] */
    GLU_F_VAL,             /* f32[nnz_lu]  predictLU values */
    GLU_F_LEVEL_PTR,       /* i32[num_lev+1] */
    GLU_F_LEVEL_IDX,       /* i32[n] */
    GLU_F_ROW_PERM,        /* u32[n] */
    GLU_F_ROW_PERM_INV,    /* u32[n] */
    GLU_F_COL_PERM,        /* u32[n] */
    GLU_F_COL_PERM_INV,    /* u32[n] */
    GLU_F_ROW_SCALE,       /* f64[n] */
    GLU_F_COL_SCALE_PERM,  /* f64[n] */
    GLU_F_DUMP_AX,         /* f64[nnz]  permuted + scaled A (my_DumpA) */
    GLU_F_DUMP_AI,         /* u32[nnz] */
    GLU_F_DUMP_AP          /* u32[n+1] */
} glu_field;

glu_status glu_analysis_get(const glu_analysis *a, glu_field f, void *dst, size_t bytes);

/* ------------------------------------------------------------------------
 * Device plan: the pattern, levels, update maps and solve schedules,
 * uploaded once and reused by every factorization/solve.
 * ---------------------------------------------------------------------- */
typedef struct glu_plan glu_plan;

/* The Symbolic_Matrix fields LUonDevice reads (include/symbolic.h:12-23). */
typedef struct glu_pattern {
    uint32_t n, nnz;
    int32_t num_lev;
    const uint32_t *sym_c_ptr, *sym_r_idx, *l_col_ptr;
    const uint32_t *csr_r_ptr, *csr_c_idx, *csr_diag_ptr;
    const int32_t *level_ptr, *level_idx;
} glu_pattern;

typedef struct glu_plan_opts {
    int32_t fuse_units;     /* level fusion: levels with <= this many update
                               units run inside one workgroup; 0 = default */
    int32_t block_threads;  /* threads per workgroup of the fused and batched
                               kernels (64..1024, multiple of 64); 0 = default */
    int32_t use_graph;      /* 1 = replay the level schedule from a hipGraph */
    int32_t lds_windows;    /* runs of narrow levels factored with their values
                               staged in LDS, counted from the last level:
                               0 = default (64), -1 = none, k > 0 = at most k */
    int32_t solve_dataflow; /* narrow levels of the single-matrix solves run
                               barrier-free in one workgroup: 0 = default (on
                               when LDS holds the ready bits), 1 = on, -1 = off */
    int32_t krylov;         /* refinement of glu_batch_solve/glu_group_solve after
                               the first solve: 0 = default (2), -1 = plain
                               (x += LU^-1 r per step), m in 1..4 = restarted
                               flexible GMRES(m) in FP64 preconditioned by the
                               FP32 factors (m solves per restart cycle) */
    int32_t lds_window_values; /* values one LDS window may stage: 0 = as many
                               as the workgroup's LDS holds (gfx950: ~39.8k),
                               k > 0 = at most k (smaller windows, more of
                               them: a schedule knob and a test hook) */
    int32_t batch_grid_levels; /* glu_batch_solve/glu_batch_factor: the leading
                               factorization steps run as one grid launch per
                               step over the whole batch (a lane's table data
                               serves up to 4 matrices) before the
                               workgroup-per-matrix kernel: 0 = default (the
                               steps ahead of the first LDS window with at
                               least block_threads lanes), -1 = none, k > 0 =
                               at most k of those.  Same bits either way. */
} glu_plan_opts;

glu_status glu_plan_create(const glu_pattern *pat, const glu_plan_opts *opts, glu_plan **out);
/* As glu_plan_create, plus the permutation/scaling needed by glu_solve and
 * the original-A structure used by glu_refresh/glu_refine/glu_batch_solve. */
glu_status glu_plan_create_from_analysis(const glu_analysis *a, const glu_plan_opts *opts,
                                         glu_plan **out);
void glu_plan_destroy(glu_plan *p);

/* info[0]=n info[1]=nnz_lu info[2]=nnz_a info[3]=num_lev info[4]=updates
 * info[5]=segments in the factor schedule info[6]=device bytes held
 * info[7]=64-lane column tasks info[8]=LDS windows of the factorization
 * info[9]=steps inside them info[10]=values the largest window stages in
 * LDS info[11]=leading steps the batched kernels run as grid launches */
glu_status glu_plan_info(const glu_plan *p, int64_t info[12]);

/* ------------------------------------------------------------------------
 * Numeric factorization (FP32, in place, L unit-lower strictly below the
 * diagonal, U including the diagonal — the layout of Symbolic_Matrix::val).
 * perturb != 0: GESP as numeric.cu:490-506 and :139-140 — pert =
 * 3.45e-4 * max column sum of |val| (FP32, before factorization); a pivot
 * with |d| < pert becomes +pert, only in columns whose L part is non-empty.
 * ---------------------------------------------------------------------- */
glu_status glu_factor(glu_plan *p, float *d_val, int perturb, void *stream);
/* Host buffer in/out, synchronous: the LUonDevice contract. */
glu_status glu_factor_host(glu_plan *p, float *h_val, int perturb);

/* FP32 solve with the factors `d_lu`, Symbolic_Matrix::solve semantics
 * (permute+scale rhs, unit-L forward, U backward, unpermute+scale), bit-exact
 * with the reference's host solve.  transpose != 0 solves A^T x = b.
 * Needs a plan created from an analysis.  d_b may equal d_x. */
glu_status glu_solve(glu_plan *p, const float *d_lu, const float *d_b, float *d_x,
                     int transpose, void *stream);

/* Scale + place `batch` FP64 value sets of A (caller CSC order, nnz(A) each)
 * into the LU pattern with the frozen MC64 scales and permutation
 * ((ax * rs) * cs, nicslu_mc64scale.c:82), cast to FP32, fill entries 0.
 * d_lu receives batch * nnz_lu floats. */
glu_status glu_refresh(glu_plan *p, const double *d_ax, float *d_lu, int batch, void *stream);

/* Iterative refinement of A x = b (or A^T x = b): FP32 solves with d_lu,
 * FP64 residual r = b - A x on the original A values d_ax and FP64 update,
 * until eta = |r|_inf / (|A|_inf |x|_inf + |b|_inf) <= tol or maxit.
 * x starts from 0.  h_iters/h_eta (host, may be NULL) receive the result
 * (h_iters counts preconditioner solves).  Orders whose FP32 sweep vector
 * fits in LDS run FGMRES(plan krylov) entirely in one workgroup; larger
 * orders run plain refinement with the two sweeps in one workgroup and the
 * FP64 work in grid kernels (or, beyond the sweeps' LDS ready bits, level
 * by level).  Synchronous. */
glu_status glu_refine(glu_plan *p, const double *d_ax, const float *d_lu, const double *d_b,
                      double *d_x, double tol, int maxit, int transpose, int *h_iters,
                      double *h_eta, void *stream);

/* glu_refine without any allocation: the same refinement and results (the
 * same kernels), with its scratch (the sweep-layout copy of the factors,
 * the correction and residual vectors, the FGMRES vectors, the iteration
 * count and eta) in a static device array of GLU_NOALLOC_SCRATCH_BYTES per
 * device, which each device that runs it holds until the process ends.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  It calls no hipMalloc/hipFree.  Limits: the
 * order must fit the one-workgroup refinement (n up to ~1.2M) and its
 * scratch must fit the array, else GLU_ERR_ARGUMENT (use glu_refine).
 * Every call on one device shares that device's array: calls that overlap
 * on one device (other streams or threads) race and must not be issued. */
#define GLU_NOALLOC_SCRATCH_BYTES (64u << 20)
glu_status glu_refine_noalloc(glu_plan *p, const double *d_ax, const float *d_lu, const double *d_b,
                              double *d_x, double tol, int maxit, int transpose, int *h_iters,
                              double *h_eta, void *stream);

/* Adjoint of x = A^-1 b with respect to the stored values of A:
 * grad_ax[k] = -lam[row(k)] * x[col(k)] for every entry k of the caller's
 * CSC (lam = A^-T g).  All pointers are device buffers. */
glu_status glu_adjoint_values(glu_plan *p, const double *d_lam, const double *d_x, double *d_grad_ax,
                              void *stream);

/* Same-pattern batch: for every b < batch, refresh the values
 * d_ax + b*nnz_a into d_lu + b*nnz_lu, factor, solve and refine d_b + b*n
 * into d_x + b*n (the wide leading steps as grid launches over an
 * interleaved copy of the batch, the rest one workgroup per matrix).
 * d_iters[b], d_eta[b] (device) receive the per-matrix iteration count and
 * backward error.  flags bit0: perturb. */
glu_status glu_batch_solve(glu_plan *p, int batch, const double *d_ax, const double *d_b,
                           double *d_x, float *d_lu, double tol, int maxit, int32_t *d_iters,
                           double *d_eta, unsigned flags, void *stream);

/* Same-pattern batched refactor: for every b < batch, refresh the values
 * d_ax + b*nnz_a into d_lu + b*nnz_lu (frozen scaling, as glu_refresh) and
 * factor them (as glu_factor: bit-identical factors), scheduled as
 * glu_batch_solve's factorization.  This is LUonDevice (include/numeric.h:6,
 * src/numeric.cu:347-639) over many value sets of one pattern, with the
 * values staged from A instead of from Symbolic_Matrix::predictLU
 * (src/symbolic.cc:82-103).  flags bit0: perturb. */
glu_status glu_batch_factor(glu_plan *p, int batch, const double *d_ax, float *d_lu, unsigned flags,
                            void *stream);

/* ABFT column checksums, Symbolic_Matrix::ABFTCalculateCCA and
 * ABFTCheckResult (src/symbolic.cc:163-197; called around LUonDevice by
 * src/lu_cmd.cpp:110-120 under GLU_DEBUG), for `batch` value sets of the
 * plan's pattern (stride nnz_lu; sums stride n), FP32 with sequential sums
 * in the reference's order.
 * glu_abft_sums: d_cca[i] = sum of column i of d_val (the values before the
 * factorization).
 * glu_abft_check: with CCL[r] = 1 + sum of L(:,r) and CCA_ABFT[i] =
 * sum_{U(:,i), diagonal included} CCL[row] * U, d_first_bad[b] = the first
 * column with |CCA_ABFT - CCA| > tol (the reference uses tol = 1e-5; as
 * there, a NaN difference does not fail), or -1; d_max_diff[b] = the
 * largest |CCA_ABFT - CCA| (NaN if any is NaN).  batch <= 65535. */
glu_status glu_abft_sums(glu_plan *p, int batch, const float *d_val, float *d_cca, void *stream);
glu_status glu_abft_check(glu_plan *p, int batch, const float *d_lu, const float *d_cca, float tol,
                          int32_t *d_first_bad, float *d_max_diff, void *stream);

/* ------------------------------------------------------------------------
 * Uneven batch: distinct patterns refactored + refined in one fused launch.
 * ---------------------------------------------------------------------- */
typedef struct glu_group glu_group;

glu_status glu_group_create(glu_plan *const *plans, int count, glu_group **out);
void glu_group_destroy(glu_group *g);
/* Members run through the single-matrix path alongside the fused launch:
 * the `count` heaviest (0..2; -1, the default: one in a group of four or
 * more), on streams of their own, while the fused launch takes the rest.
 * Results are the same bits either way.  Only members whose refinement
 * vector fits in LDS are moved.  (No counterpart in the reference: part of
 * the uneven batch this library adds.) */
glu_status glu_group_set_split(glu_group *g, int count);
/* Host arrays of `count` device pointers each, in plan order. */
glu_status glu_group_solve(glu_group *g, const double *const *d_ax, const double *const *d_b,
                           double *const *d_x, float *const *d_lu, double tol, int maxit,
                           int32_t *d_iters, double *d_eta, unsigned flags, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* GLU_HIP_H */
