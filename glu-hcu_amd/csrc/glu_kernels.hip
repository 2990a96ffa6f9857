// gfx950 kernels of the GLU numeric phase: level-scheduled right-looking
// FP32 column update, FP32 triangular solves, FP64 residual/refinement, and
// the persistent per-matrix pipeline used by the batched launches.
//
// Semantics (what every kernel computes) follow the reference:
//   column update   src/numeric.cu:10-103 (RL), :105-202 (RL_perturb),
//                   :204-345 (the one-column stream-mode trio)
//   GESP value      src/numeric.cu:490-506
//   FP32 solve      src/symbolic.cc:200-251
// The structure does not: per level, one wave per work unit (pivot column k
// times a slice of U_k); targets come from a precomputed scatter map instead
// of a dense n-length scratch column; the L(:,k) /= U(k,k) write-back of a
// level is deferred to the next level's launch so a level costs one barrier;
// updates are fire-and-forget FP32 atomics (the reference's own update
// operator, val += -(l*u)) where an entry gets one update per level, and
// lane-owned sequential merges where several columns of a level update the
// same entry, so the factors are deterministic and equal a sequential pass
// over each level's columns in increasing order; solves pull along CSR/CSC
// rows so each unknown is owned by one lane and accumulates in the
// reference's order.
//
// Build flags (see the package Makefile): -ffp-contract=off so products are
// rounded before they are subtracted, as the reference's FP32 kernels and
// host solve do; packed FP32 disabled; -munsafe-fp-atomics so FP32 atomic
// adds lower to global_atomic_add_f32.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "glu_dev.h"

namespace glu {

constexpr int kWave = 64;
constexpr float kPertFactor = 3.45e-4f;   // numeric.cu:503

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------

template <int Scope>
__device__ __forceinline__ void atomic_sub(float *p, float v)
{
    __hip_atomic_fetch_add(p, -v, __ATOMIC_RELAXED, Scope);
}

__device__ __forceinline__ uint32_t pow2_ceil(uint32_t x)
{
    return x <= 1u ? 1u : (1u << (32 - __clz((int)(x - 1u))));
}

// max that propagates NaN (a NaN residual must not look converged)
__device__ __forceinline__ double nmax(double a, double b)
{
    return (a > b || a != a) ? a : b;
}

__device__ __forceinline__ double wave_max(double v)
{
    for (int o = kWave / 2; o > 0; o >>= 1) v = nmax(v, __shfl_xor(v, o, kWave));
    return v;
}

// Workgroup max; `sh` is LDS with one slot per wave.  Ends with a barrier.
__device__ double wg_max(double v, double *sh)
{
    v = wave_max(v);
    const int nw = blockDim.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) sh[threadIdx.x / kWave] = v;
    __syncthreads();
    double r = sh[0];
    for (int i = 1; i < nw; ++i) r = nmax(r, sh[i]);
    __syncthreads();
    return r;
}

// ---------------------------------------------------------------------------
// factorization
// ---------------------------------------------------------------------------

// One work unit: pivot column k, U_k entries [jb, jb + jc).
//   l_i = A(i,k) / d          (d = A(k,k), GESP-perturbed if requested)
//   A(i,j) += -(l_i * A(k,j)) for i in L_k, j in the slice of U_k
// except targets that another column of the same level also updates; those
// belong to a merge lane (factor_merge).  Every remaining target receives one
// update per level, so the atomic is a plain fire-and-forget add and the
// result does not depend on the order waves run in.
// Lanes are split into groups of G = pow2 >= |L_k| (|L_k| <= 64), each group
// owning one j at a time; longer columns loop over L_k in strides of 64.
template <int Scope>
__device__ __forceinline__ void factor_unit(const FactorDev &F, float *__restrict__ val,
                                            const Unit u, int lane, float pert, int perturb)
{
    const uint32_t k = u.col;
    const uint32_t dpos = F.diag[k];
    const uint32_t L = F.lcnt[k];
    const uint32_t *__restrict__ up = F.upos + F.ufirst[k] + u.jb;
    const uint32_t *__restrict__ map = F.umap + (F.moff[k] + (uint64_t)u.jb * L);
    float d = val[dpos];
    if (perturb && fabsf(d) < pert) d = pert;
    if (L <= (uint32_t)kWave) {
        const uint32_t G = pow2_ceil(L);
        const uint32_t sh = (uint32_t)(__ffs((int)G) - 1);
        const uint32_t ii = (uint32_t)lane & (G - 1u);
        const uint32_t g = (uint32_t)lane >> sh;
        const uint32_t step = (uint32_t)kWave >> sh;
        const bool act = ii < L;
        float l = 0.f;
        if (act) l = val[dpos + 1 + ii] / d;
        for (uint32_t jj = g; jj < u.jc; jj += step) {
            const float uv = val[up[jj]];
            if (act) {
                const uint32_t t = map[(uint64_t)jj * L + ii];
                if (!(t & kMergedBit)) atomic_sub<Scope>(val + t, l * uv);
            }
        }
    } else {
        for (uint32_t jj = 0; jj < u.jc; ++jj) {
            const float uv = val[up[jj]];
            const uint32_t *__restrict__ mrow = map + (uint64_t)jj * L;
            for (uint32_t ii = (uint32_t)lane; ii < L; ii += kWave) {
                const uint32_t t = mrow[ii];
                if (t & kMergedBit) continue;
                const float l = val[dpos + 1 + ii] / d;
                atomic_sub<Scope>(val + t, l * uv);
            }
        }
    }
}

// Merge entries e0 + lane: apply every source of this level in increasing
// column order, t += -(l * u) with l = L(i,k) / U(k,k) — the order in which
// a sequential pass over the level's columns would apply them.
__device__ __forceinline__ void factor_merge(const FactorDev &F, float *__restrict__ val, uint32_t e,
                                             uint32_t e_end, float pert, int perturb)
{
    if (e >= e_end) return;
    const uint32_t t = F.ment[e];
    float v = val[t];
    const uint32_t s1 = F.msrc_ptr[e + 1];
    for (uint32_t s = F.msrc_ptr[e]; s < s1; ++s) {
        const MergeSrc m = F.msrc[s];
        float d = val[m.dpos];
        if (perturb && fabsf(d) < pert) d = pert;
        const float l = val[m.lpos] / d;
        v = v + (-(l * val[m.upos]));
    }
    val[t] = v;
}

// Deferred write-back of column k: L(:,k) /= d, and d := pert when GESP fired.
__device__ __forceinline__ void factor_div(const FactorDev &F, float *__restrict__ val,
                                           uint32_t k, int lane, float pert, int perturb)
{
    const uint32_t dpos = F.diag[k];
    const uint32_t L = F.lcnt[k];
    float d = val[dpos];
    const bool pt = perturb && fabsf(d) < pert;
    if (pt) d = pert;
    for (uint32_t ii = (uint32_t)lane; ii < L; ii += kWave) val[dpos + 1 + ii] = val[dpos + 1 + ii] / d;
    if (pt && lane == 0) val[dpos] = pert;
}

__device__ __forceinline__ uint32_t merge_waves(const FactorDev &F, int lvl)
{
    return (F.ment_ptr[lvl + 1] - F.ment_ptr[lvl] + kWave - 1) / kWave;
}

__device__ __forceinline__ uint32_t level_items(const FactorDev &F, int lvl)
{
    uint32_t items = F.unit_ptr[lvl + 1] - F.unit_ptr[lvl] + merge_waves(F, lvl);
    if (lvl > 0) items += F.dcol_ptr[lvl] - F.dcol_ptr[lvl - 1];
    return items;
}

// Item t of level lvl: update units of lvl, then merge waves (64 entries
// each), then the write-backs of the columns of level lvl-1.  lvl == num_lev
// has write-backs only.
template <int Scope>
__device__ __forceinline__ void factor_item(const FactorDev &F, float *__restrict__ val, int lvl,
                                            uint32_t item, int lane, float pert, int perturb)
{
    const uint32_t u0 = F.unit_ptr[lvl];
    const uint32_t nu = F.unit_ptr[lvl + 1] - u0;
    if (item < nu) {
        factor_unit<Scope>(F, val, F.units[u0 + item], lane, pert, perturb);
        return;
    }
    item -= nu;
    const uint32_t nm = merge_waves(F, lvl);
    if (item < nm) {
        factor_merge(F, val, F.ment_ptr[lvl] + item * kWave + (uint32_t)lane, F.ment_ptr[lvl + 1], pert, perturb);
        return;
    }
    item -= nm;
    if (lvl > 0) {
        const uint32_t d0 = F.dcol_ptr[lvl - 1];
        if (item < F.dcol_ptr[lvl] - d0) factor_div(F, val, F.dcols[d0 + item], lane, pert, perturb);
    }
}

// GESP: pert = 3.45e-4 * max_c sum_p |val[p]| (FP32 sums in column order).
__global__ void k_pert(const uint32_t *__restrict__ col_ptr, const float *__restrict__ val,
                       uint32_t n, unsigned int *__restrict__ norm_bits)
{
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    float s = 0.f;
    for (uint32_t p = col_ptr[c]; p < col_ptr[c + 1]; ++p) s += fabsf(val[p]);
    if (s == s) atomicMax(norm_bits, __float_as_uint(s));
}

// Also re-arms the accumulator for the next call (no memset in the stream).
__global__ void k_pert_finish(unsigned int *__restrict__ norm_bits, float *__restrict__ pert)
{
    *pert = kPertFactor * __uint_as_float(*norm_bits);
    *norm_bits = 0u;
}

__global__ void k_zero_u64(unsigned long long *__restrict__ p, int count)
{
    if ((int)threadIdx.x < count) p[threadIdx.x] = 0ull;
}

// One level across the grid: one wave per item.
__global__ void k_factor_grid(FactorDev F, float *__restrict__ val, const float *__restrict__ d_pert,
                              int lvl, int perturb)
{
    const uint32_t item = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    const float pert = perturb ? *d_pert : 0.f;
    factor_item<__HIP_MEMORY_SCOPE_AGENT>(F, val, lvl, item, lane, pert, perturb);
}

// A run of narrow levels inside one workgroup, a barrier between levels.
__global__ void k_factor_fused(FactorDev F, float *__restrict__ val, const float *__restrict__ d_pert,
                               int l0, int l1, int perturb)
{
    const int lane = threadIdx.x & (kWave - 1);
    const uint32_t w = threadIdx.x / kWave, nw = blockDim.x / kWave;
    const float pert = perturb ? *d_pert : 0.f;
    for (int lvl = l0; lvl < l1; ++lvl) {
        const uint32_t items = level_items(F, lvl);
        for (uint32_t t = w; t < items; t += nw)
            factor_item<__HIP_MEMORY_SCOPE_AGENT>(F, val, lvl, t, lane, pert, perturb);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// triangular solves
// ---------------------------------------------------------------------------

__device__ __forceinline__ void tri_row(const TriDev &T, const float *__restrict__ lu,
                                        float *__restrict__ x, uint32_t r)
{
    const uint32_t b = T.beg[r] + T.begoff, e = T.end[r];
    float acc = x[r];
    if (!T.desc) {
        for (uint32_t q = b; q < e; ++q) {
            const uint32_t p = T.vpos ? T.vpos[q] : q;
            acc = acc - lu[p] * x[T.idx[q]];
        }
    } else {
        for (uint32_t q = e; q > b; --q) {
            const uint32_t p = T.vpos ? T.vpos[q - 1] : (q - 1);
            acc = acc - lu[p] * x[T.idx[q - 1]];
        }
    }
    if (T.dpos) acc = acc / lu[T.dpos[r]];
    x[r] = acc;
}

__global__ void k_tri_grid(TriDev T, const float *__restrict__ lu, float *__restrict__ x, int lvl)
{
    const uint32_t t = T.lvl_ptr[lvl] + blockIdx.x * blockDim.x + threadIdx.x;
    if (t < T.lvl_ptr[lvl + 1]) tri_row(T, lu, x, T.rows[t]);
}

__global__ void k_tri_fused(TriDev T, const float *__restrict__ lu, float *__restrict__ x, int l0, int l1)
{
    for (int lvl = l0; lvl < l1; ++lvl) {
        for (uint32_t t = T.lvl_ptr[lvl] + threadIdx.x; t < T.lvl_ptr[lvl + 1]; t += blockDim.x)
            tri_row(T, lu, x, T.rows[t]);
        __syncthreads();
    }
}

// out[j] = float(double(in[P[j]]) * S[by_perm ? P[j] : j])
template <typename TIn>
__global__ void k_perm_scale(const TIn *__restrict__ in, float *__restrict__ out,
                             const uint32_t *__restrict__ P, const double *__restrict__ S,
                             int by_perm, uint32_t n)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint32_t pj = P[j];
    out[j] = (float)((double)in[pj] * S[by_perm ? pj : j]);
}

// x[j] += double(t[P[j]]) * S[by_perm ? P[j] : j]
__global__ void k_update_x(const float *__restrict__ t, double *__restrict__ x,
                           const uint32_t *__restrict__ P, const double *__restrict__ S,
                           int by_perm, uint32_t n)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint32_t pj = P[j];
    x[j] = x[j] + (double)t[pj] * S[by_perm ? pj : j];
}

// ---------------------------------------------------------------------------
// refresh, residual, norms
// ---------------------------------------------------------------------------

// lu[p] = float((ax[src[p]] * rs_col[c]) * csp[row[p]]), 0 for fill.
__device__ __forceinline__ void refresh_col(const RefineDev &R, const uint32_t *__restrict__ col_ptr,
                                            const double *__restrict__ ax, float *__restrict__ lu,
                                            uint32_t c)
{
    const double rsc = R.rs_col[c];
    for (uint32_t p = col_ptr[c]; p < col_ptr[c + 1]; ++p) {
        const uint32_t s = R.lu_src[p];
        lu[p] = (s == 0xffffffffu) ? 0.f : (float)((ax[s] * rsc) * R.csp[R.row_idx[p]]);
    }
}

__global__ void k_refresh(RefineDev R, const uint32_t *__restrict__ col_ptr, const double *__restrict__ ax,
                          float *__restrict__ lu, uint32_t n, uint32_t nnz_lu, int batch)
{
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (uint64_t)n * batch) return;
    const uint32_t b = (uint32_t)(g / n), c = (uint32_t)(g % n);
    refresh_col(R, col_ptr, ax + (uint64_t)b * R.nnz_a, lu + (uint64_t)b * nnz_lu, c);
}

// Row i of A (transpose=0, CSR) or of A' (transpose=1, the caller's CSC).
__device__ __forceinline__ double a_row_dot(const RefineDev &R, const double *__restrict__ ax,
                                            const double *__restrict__ x, uint32_t i, int transpose,
                                            double acc, double *abs_sum)
{
    double s = 0.0;
    if (!transpose) {
        for (uint32_t q = R.a_rptr[i]; q < R.a_rptr[i + 1]; ++q) {
            const double a = ax[R.a_src[q]];
            acc = fma(-a, x[R.a_cidx[q]], acc);
            s += fabs(a);
        }
    } else {
        for (uint32_t q = R.a_cptr[i]; q < R.a_cptr[i + 1]; ++q) {
            const double a = ax[q];
            acc = fma(-a, x[R.a_ridx[q]], acc);
            s += fabs(a);
        }
    }
    if (abs_sum) *abs_sum = s;
    return acc;
}

// norms[0] = |r|_inf, [1] = |x|_inf, [2] = |A|_inf, [3] = |b|_inf (as u64
// bits of non-negative doubles; NaN compares largest).
__device__ __forceinline__ void umax64(unsigned long long *p, double v)
{
    unsigned long long bits = (v == v) ? (unsigned long long)__double_as_longlong(v) : 0x7ff8000000000000ull;
    atomicMax(p, bits);
}

__global__ void k_residual(RefineDev R, const double *__restrict__ ax, const double *__restrict__ b,
                           const double *__restrict__ x, double *__restrict__ r, uint32_t n,
                           int transpose, int with_anorm, unsigned long long *__restrict__ norms)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double rmax = 0.0, xmax = 0.0, amax = 0.0, bmax = 0.0;
    if (i < n) {
        double asum;
        const double ri = a_row_dot(R, ax, x, i, transpose, b[i], with_anorm ? &asum : nullptr);
        r[i] = ri;
        rmax = fabs(ri);
        xmax = fabs(x[i]);
        bmax = fabs(b[i]);
        if (with_anorm) amax = asum;
    }
    rmax = wave_max(rmax);
    xmax = wave_max(xmax);
    amax = wave_max(amax);
    bmax = wave_max(bmax);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        umax64(norms + 0, rmax);
        umax64(norms + 1, xmax);
        if (with_anorm) {
            umax64(norms + 2, amax);
            umax64(norms + 3, bmax);
        }
    }
}

// grad[k] = -lam[row(k)] * x[col(k)] over the caller's CSC entries.
__global__ void k_adjoint_values(RefineDev R, const double *__restrict__ lam, const double *__restrict__ x,
                                 double *__restrict__ grad, uint32_t n)
{
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const double xc = x[c];
    for (uint32_t k = R.a_cptr[c]; k < R.a_cptr[c + 1]; ++k) grad[k] = -(lam[R.a_ridx[k]] * xc);
}

// ---------------------------------------------------------------------------
// persistent per-matrix pipeline (one workgroup owns one matrix)
// ---------------------------------------------------------------------------

struct MatIO {
    const double *ax;   // nnz_a values, caller CSC order
    const double *b;    // n
    double *x;          // n
    float *lu;          // nnz_lu
    float *t32;         // n scratch
    double *r64;        // n scratch
    int32_t *iters;
    double *eta;
};

__device__ __forceinline__ void wg_tri(const TriDev &T, const float *__restrict__ lu, float *__restrict__ x)
{
    for (int lvl = 0; lvl < T.nlev; ++lvl) {
        for (uint32_t t = T.lvl_ptr[lvl] + threadIdx.x; t < T.lvl_ptr[lvl + 1]; t += blockDim.x)
            tri_row(T, lu, x, T.rows[t]);
        __syncthreads();
    }
}

__device__ void wg_pipeline(const MatDesc &M, const MatIO &io, double tol, int maxit, int perturb,
                            double *sh)
{
    const uint32_t n = M.n;
    const uint32_t tid = threadIdx.x, nt = blockDim.x;
    const int lane = tid & (kWave - 1);
    const uint32_t w = tid / kWave, nw = nt / kWave;
    const FactorDev &F = M.f;
    const SolveDev &S = M.s;
    const RefineDev &R = M.r;

    // 1. values: scale + place A into the LU pattern; |A|_inf and |b|_inf
    double amax = 0.0, bmax = 0.0;
    for (uint32_t c = tid; c < n; c += nt) {
        refresh_col(R, F.col_ptr, io.ax, io.lu, c);
        double s = 0.0;
        for (uint32_t q = R.a_rptr[c]; q < R.a_rptr[c + 1]; ++q) s += fabs(io.ax[R.a_src[q]]);
        amax = nmax(amax, s);
        bmax = nmax(bmax, fabs(io.b[c]));
        io.x[c] = 0.0;
        io.r64[c] = io.b[c];
    }
    const double anorm = wg_max(amax, sh);
    const double bnorm = wg_max(bmax, sh);

    // 2. GESP value from the freshly placed values
    float pert = 0.f;
    if (perturb) {
        double cm = 0.0;
        for (uint32_t c = tid; c < n; c += nt) {
            float s = 0.f;
            for (uint32_t p = F.col_ptr[c]; p < F.col_ptr[c + 1]; ++p) s += fabsf(io.lu[p]);
            if (s == s) cm = fmax(cm, (double)s);
        }
        pert = kPertFactor * (float)wg_max(cm, sh);
    }

    // 3. factorization, level by level
    for (int lvl = 0; lvl <= F.num_lev; ++lvl) {
        const uint32_t items = level_items(F, lvl);
        for (uint32_t t = w; t < items; t += nw)
            factor_item<__HIP_MEMORY_SCOPE_WORKGROUP>(F, io.lu, lvl, t, lane, pert, perturb);
        __syncthreads();
    }

    // 4. refinement: x += LU^-1 r (FP32), r = b - A x (FP64)
    int it = 0;
    double eta = 0.0;
    for (;;) {
        for (uint32_t j = tid; j < n; j += nt) {
            const uint32_t pj = S.col_perm[j];
            io.t32[j] = (float)(io.r64[pj] * S.col_scale_perm[j]);
        }
        __syncthreads();
        wg_tri(S.lo, io.lu, io.t32);
        wg_tri(S.up, io.lu, io.t32);
        for (uint32_t j = tid; j < n; j += nt)
            io.x[j] = io.x[j] + (double)io.t32[S.row_perm_inv[j]] * S.row_scale[j];
        __syncthreads();
        double rmax = 0.0, xmax = 0.0;
        for (uint32_t i = tid; i < n; i += nt) {
            const double ri = a_row_dot(R, io.ax, io.x, i, 0, io.b[i], nullptr);
            io.r64[i] = ri;
            rmax = nmax(rmax, fabs(ri));
            xmax = nmax(xmax, fabs(io.x[i]));
        }
        const double rn = wg_max(rmax, sh);
        const double xn = wg_max(xmax, sh);
        eta = rn / (anorm * xn + bnorm);
        ++it;
        if (!(eta > tol) || it >= maxit) break;   // also stops on NaN
    }
    if (tid == 0) {
        *io.iters = it;
        *io.eta = eta;
    }
}

// Same pattern: matrix b = blockIdx.x uses desc[0] and contiguous buffers.
__global__ void k_batch_same(const MatDesc *__restrict__ desc, const double *__restrict__ ax,
                             const double *__restrict__ b, double *__restrict__ x, float *__restrict__ lu,
                             float *__restrict__ t32, double *__restrict__ r64, int32_t *__restrict__ iters,
                             double *__restrict__ eta, double tol, int maxit, int perturb)
{
    __shared__ double sh[1024 / kWave];
    const MatDesc &M = desc[0];
    const uint64_t m = blockIdx.x;
    MatIO io;
    io.ax = ax + m * M.r.nnz_a;
    io.b = b + m * M.n;
    io.x = x + m * M.n;
    io.lu = lu + m * M.nnz;
    io.t32 = t32 + m * M.n;
    io.r64 = r64 + m * M.n;
    io.iters = iters + m;
    io.eta = eta + m;
    wg_pipeline(M, io, tol, maxit, perturb, sh);
}

// Distinct patterns: workgroup g runs matrix order[g] with its own tables.
__global__ void k_batch_group(const MatDesc *__restrict__ descs, const MatIO *__restrict__ ios,
                              const uint32_t *__restrict__ order, double tol, int maxit, int perturb)
{
    __shared__ double sh[1024 / kWave];
    const uint32_t m = order[blockIdx.x];
    wg_pipeline(descs[m], ios[m], tol, maxit, perturb, sh);
}

// ---------------------------------------------------------------------------
// host-side launch wrappers (C++ linkage, used by glu_plan.hip)
// ---------------------------------------------------------------------------

namespace launch {

static inline unsigned blocks_for(uint64_t work, unsigned per) { return (unsigned)((work + per - 1) / per); }

// norm_bits must be zero on entry; k_pert_finish leaves it zero again.
void pert(const FactorDev &F, const float *val, unsigned int *norm_bits, float *d_pert, hipStream_t st)
{
    k_pert<<<blocks_for(F.n, 256), 256, 0, st>>>(F.col_ptr, val, F.n, norm_bits);
    k_pert_finish<<<1, 1, 0, st>>>(norm_bits, d_pert);
}

void factor_grid(const FactorDev &F, float *val, const float *d_pert, int lvl, uint32_t items,
                 int perturb, int threads, hipStream_t st)
{
    const unsigned wpb = (unsigned)threads / kWave;
    k_factor_grid<<<blocks_for(items, wpb), threads, 0, st>>>(F, val, d_pert, lvl, perturb);
}

void factor_fused(const FactorDev &F, float *val, const float *d_pert, int l0, int l1, int perturb,
                  int threads, hipStream_t st)
{
    k_factor_fused<<<1, threads, 0, st>>>(F, val, d_pert, l0, l1, perturb);
}

void tri_grid(const TriDev &T, const float *lu, float *x, int lvl, uint32_t rows, hipStream_t st)
{
    k_tri_grid<<<blocks_for(rows, 256), 256, 0, st>>>(T, lu, x, lvl);
}

void tri_fused(const TriDev &T, const float *lu, float *x, int l0, int l1, int threads, hipStream_t st)
{
    k_tri_fused<<<1, threads, 0, st>>>(T, lu, x, l0, l1);
}

void perm_scale_f32(const float *in, float *out, const uint32_t *P, const double *S, int by_perm,
                    uint32_t n, hipStream_t st)
{
    k_perm_scale<float><<<blocks_for(n, 256), 256, 0, st>>>(in, out, P, S, by_perm, n);
}

void perm_scale_f64(const double *in, float *out, const uint32_t *P, const double *S, int by_perm,
                    uint32_t n, hipStream_t st)
{
    k_perm_scale<double><<<blocks_for(n, 256), 256, 0, st>>>(in, out, P, S, by_perm, n);
}

void update_x(const float *t, double *x, const uint32_t *P, const double *S, int by_perm, uint32_t n,
              hipStream_t st)
{
    k_update_x<<<blocks_for(n, 256), 256, 0, st>>>(t, x, P, S, by_perm, n);
}

void refresh(const RefineDev &R, const uint32_t *col_ptr, const double *ax, float *lu, uint32_t n,
             uint32_t nnz_lu, int batch, hipStream_t st)
{
    k_refresh<<<blocks_for((uint64_t)n * batch, 256), 256, 0, st>>>(R, col_ptr, ax, lu, n, nnz_lu, batch);
}

void residual(const RefineDev &R, const double *ax, const double *b, const double *x, double *r, uint32_t n,
              int transpose, int with_anorm, unsigned long long *norms, hipStream_t st)
{
    k_residual<<<blocks_for(n, 256), 256, 0, st>>>(R, ax, b, x, r, n, transpose, with_anorm, norms);
}

void batch_same(const MatDesc *desc, int batch, const double *ax, const double *b, double *x, float *lu,
                float *t32, double *r64, int32_t *iters, double *eta, double tol, int maxit, int perturb,
                int threads, hipStream_t st)
{
    k_batch_same<<<batch, threads, 0, st>>>(desc, ax, b, x, lu, t32, r64, iters, eta, tol, maxit, perturb);
}

void batch_group(const MatDesc *descs, const void *ios, const uint32_t *order, int count, double tol,
                 int maxit, int perturb, int threads, hipStream_t st)
{
    k_batch_group<<<count, threads, 0, st>>>(descs, static_cast<const MatIO *>(ios), order, tol, maxit,
                                            perturb);
}

void adjoint_values(const RefineDev &R, const double *lam, const double *x, double *grad, uint32_t n,
                    hipStream_t st)
{
    k_adjoint_values<<<blocks_for(n, 256), 256, 0, st>>>(R, lam, x, grad, n);
}

void zero_u64(unsigned long long *p, int count, hipStream_t st)
{
    k_zero_u64<<<1, 64, 0, st>>>(p, count);
}

size_t matio_size() { return sizeof(MatIO); }

}  // namespace launch
}  // namespace glu
