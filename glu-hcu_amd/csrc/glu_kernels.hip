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
// an entry that gets one update per level takes it as a plain load + store
// of the reference's own update operator, val += -(l*u) (no other item of
// the level reads or writes it, so no atomic is needed), and lane-owned
// sequential merges take entries that several columns of a level update,
// so the factors are deterministic and equal a sequential pass
// over each level's columns in increasing order; solves pull along CSR/CSC
// rows so each unknown is owned by one lane and accumulates in the
// reference's order.
//
// Build flags (see the package Makefile): -ffp-contract=off so products are
// rounded before they are subtracted, as the reference's FP32 kernels and
// host solve do; packed FP32 disabled.
#include <hip/hip_runtime.h>

#include <mutex>

#include <algorithm>
#include <cstdint>

#include "glu_dev.h"

namespace glu {

// Pointers read from device memory (the matrix descriptors) are generic to
// the compiler, which then emits flat loads: these count against the LDS
// wait counter too, so every LDS wait would also wait for global loads in
// flight.  The tables and buffers are all in global memory; rebuilding each
// pointer as a global-address-space pointer (through its integer value, a
// cast the optimizer keeps, unlike a generic->global->generic round trip)
// lets address-space inference turn its accesses into global loads.
template <typename T>
__device__ __forceinline__ T *glb(T *p)
{
    return (T *)(__attribute__((address_space(1))) T *)(uintptr_t)p;
}

constexpr int kWave = 64;
constexpr float kPertFactor = 3.45e-4f;   // numeric.cu:503

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------

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

__device__ __forceinline__ float gesp(float d, float pert, int perturb)
{
    return (perturb && fabsf(d) < pert) ? pert : d;
}

// Column tasks.  A wave takes a task; each lane owns one L entry A(i,k) of a
// pivot column k of the level and up to four entries u = A(k,j) of its U
// list, forms l = A(i,k) / A(k,k) (A(k,k) GESP-perturbed) and applies
//   A(i,j) += -(l * u)
// Each target here receives exactly this one update in this level and no
// other item of the level reads it (a target A(i,j) of pivot column k lies
// in a column of a later level), so a plain load + store is race-free and
// equals the sequential pass.  Plain stores beat L2 atomics here: the
// target's line is usually already in flight with the task's other values.
//
// CTab holds a task's table data for this lane (independent of the values,
// so it can be loaded ahead): the L and pivot positions, the U positions and
// the targets (~0u: no update in that slot).
struct CTab {
    uint32_t lpos, dpos;
    uint32_t tt[4], up[4];
};

__device__ __forceinline__ void ctab_load(const FactorDev &F, uint32_t t, uint32_t lane, CTab &c)
{
    const uint2 ln = F.cln[(size_t)t * kWave + lane];
    c.lpos = ln.x;
    c.dpos = ln.y;
    const size_t q = (size_t)t * (4 * kWave) + lane;
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u) {
        c.tt[u] = F.ctgt[q + u * kWave];
        c.up[u] = F.cup[q + u * kWave];
    }
}

__device__ __forceinline__ void ctab_run(const FactorDev &F, float *__restrict__ val, const CTab &c, float pert,
                                         int perturb)
{
    if ((c.tt[0] & c.tt[1] & c.tt[2] & c.tt[3]) == ~0u) return;   // nothing to update
    // every load without a branch (an empty slot's U position is 0 and its
    // target is read at 0, valid positions), so that they issue together
    const float d = val[c.dpos], a = val[c.lpos];
    float uv[4], tv[4];
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u) {
        uv[u] = val[c.up[u]];
        tv[u] = val[c.tt[u] != ~0u ? c.tt[u] : 0u];
    }
    const float l = a / gesp(d, pert, perturb);
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u)
        if (c.tt[u] != ~0u) val[c.tt[u]] = tv[u] + (-(l * uv[u]));
}

// Column tasks [c0, c0 + nc) of one level for a group of cooperating
// waves (w of nw), one at a time.
__device__ __forceinline__ void factor_ctasks(const FactorDev &F, float *__restrict__ val, uint32_t c0, uint32_t nc,
                                              uint32_t tid, uint32_t nthreads, float pert, int perturb)
{
    const uint32_t lane = tid & (kWave - 1), w = tid / kWave, nw = nthreads / kWave;
    for (uint32_t t = c0 + w; t < c0 + nc; t += nw) {
        CTab c;
        ctab_load(F, t, lane, c);
        ctab_run(F, val, c, pert, perturb);
    }
}

// Sources 2.. of merge entry e subtracted from v in order (four loaded at a
// time so that the loads overlap; the additions stay sequential).
__device__ __forceinline__ void factor_merge_rest(const FactorDev &F, float *__restrict__ val, uint32_t e, float &v,
                                                  float pert, int perturb)
{
    const uint32_t s1 = F.msrc_ptr[e + 1];
    uint32_t s = F.msrc_ptr[e] + 2;
    for (; s + 4 <= s1; s += 4) {
        MergeSrc m[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) m[u] = F.msrc[s + u];
        __builtin_amdgcn_sched_barrier(0);   // the four table loads issue before any value load
        float p[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) p[u] = (val[m[u].lpos] / gesp(val[m[u].dpos], pert, perturb)) * val[m[u].upos];
#pragma unroll
        for (int u = 0; u < 4; ++u) v = v + (-p[u]);
    }
    for (; s < s1; ++s) {
        const MergeSrc m = F.msrc[s];
        v = v + (-((val[m.lpos] / gesp(val[m.dpos], pert, perturb)) * val[m.upos]));
    }
}

// Merge entry e: every update of this level in increasing column order,
// t += -(l * u) — the order a sequential pass over the level applies them.
// Sources are fetched four at a time so the loads overlap; the additions
// stay sequential.
__device__ __forceinline__ void factor_merge(const FactorDev &F, float *__restrict__ val, uint32_t e,
                                             float pert, int perturb)
{
    // the target and the first two sources from the entry's record (one
    // dependent table load instead of three)
    const uint4 r0 = F.mrec[2 * e], r1 = F.mrec[2 * e + 1];
    const uint32_t t = r0.x;
    float v = val[t];
    const float p0 = (val[r0.y] / gesp(val[r0.w], pert, perturb)) * val[r0.z];
    const float p1 = (val[r1.x] / gesp(val[r1.z], pert, perturb)) * val[r1.y];
    v = v + (-p0);
    v = v + (-p1);
    if (r1.w > 2) factor_merge_rest(F, val, e, v, pert, perturb);
    val[t] = v;
}

// acc - p[0] - p[1] - ... - p[cnt-1] (desc: from p[cnt-1] down to p[0]), each
// subtraction rounded, for a wave's 64-float LDS slice p (256-byte
// aligned).  A full slice is read as two halves of eight 16-byte loads
// issued together, so only the subtraction chain is sequential.
__device__ __forceinline__ float serial_sub(float acc, const float *p, uint32_t cnt, bool desc)
{
    if (cnt == kWave) {
        const float4 *p4 = reinterpret_cast<const float4 *>(p);
        constexpr uint32_t kHalf = kWave / 8;
#pragma unroll
        for (uint32_t h = 0; h < 2; ++h) {
            float4 v[kHalf];
            const uint32_t base = desc ? (1 - h) * kHalf : h * kHalf;
#pragma unroll
            for (uint32_t k = 0; k < kHalf; ++k) v[k] = p4[base + k];
            if (!desc) {
#pragma unroll
                for (uint32_t k = 0; k < kHalf; ++k) {
                    acc = acc - v[k].x;
                    acc = acc - v[k].y;
                    acc = acc - v[k].z;
                    acc = acc - v[k].w;
                }
            } else {
#pragma unroll
                for (uint32_t k = kHalf; k > 0; --k) {
                    acc = acc - v[k - 1].w;
                    acc = acc - v[k - 1].z;
                    acc = acc - v[k - 1].y;
                    acc = acc - v[k - 1].x;
                }
            }
        }
        return acc;
    }
    if (!desc)
        for (uint32_t i = 0; i < cnt; ++i) acc = acc - p[i];
    else
        for (uint32_t i = cnt; i > 0; --i) acc = acc - p[i - 1];
    return acc;
}

// The same for an entry with many sources, by one wave: the products of 64
// sources at a time in parallel, then lane 0 adds them in source order from
// the wave's LDS slice (64 floats) — the same arithmetic, sequence included.
// kFull: every pass subtracts all 64 slots (lanes past the entry's last
// source hold +0, and v - 0 is v exactly), so partial passes take the
// full-slice path too; it holds more registers, which the batch kernels
// (one matrix per workgroup, near their register limit) cannot spare.
template <bool kFull = true>
__device__ __forceinline__ void factor_merge_wave(const FactorDev &F, float *__restrict__ val, uint32_t e,
                                                  float pert, int perturb, uint32_t lane, float *lds)
{
    const uint32_t t = F.ment[e];
    const uint32_t s0 = F.msrc_ptr[e], s1 = F.msrc_ptr[e + 1];
    float v = val[t];
    for (uint32_t base = s0; base < s1; base += kWave) {
        const uint32_t s = base + lane;
        float p = 0.f;
        if (s < s1) {
            const MergeSrc m = F.msrc[s];
            p = (val[m.lpos] / gesp(val[m.dpos], pert, perturb)) * val[m.upos];
        }
        lds[lane] = p;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        if (lane == 0) {
            const uint32_t cnt = kFull ? kWave : min((uint32_t)kWave, s1 - base);
            v = serial_sub(v, lds, cnt, false);   // v + (-p) is v - p exactly
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
    if (lane == 0) val[t] = v;
}

// Division item of a column of the previous level: L(i,k) /= U(k,k); the
// lane of the column's first L entry stores the GESP pivot.  Other lanes of
// the same column may read the pivot before or after that store: both
// values perturb to the same pert.
__device__ __forceinline__ void factor_divide(float *__restrict__ val, uint2 it, float pert, int perturb)
{
    const uint32_t dpos = it.y & ~kFirstBit;
    const float d0 = val[dpos];
    const bool pt = perturb && fabsf(d0) < pert;
    val[it.x] = val[it.x] / (pt ? pert : d0);
    if (pt && (it.y & kFirstBit)) val[dpos] = pert;
}

struct LevelSpan {
    uint32_t c0, nc, m0, nm, ml;
};

__device__ __forceinline__ LevelSpan level_span(const FactorDev &F, int lvl)
{
    LevelSpan s;
    s.c0 = F.ctask_ptr[lvl];
    s.nc = F.ctask_ptr[lvl + 1] - s.c0;
    s.m0 = F.ment_ptr[lvl];
    s.nm = F.ment_ptr[lvl + 1] - s.m0;
    s.ml = F.ment_long[lvl];
    return s;
}

// All items of one level for a group of `nthreads` cooperating lanes (a
// workgroup or the whole grid), `tid` = this lane's index in the group:
// long merges and column tasks one wave each, short merges and divisions
// one lane each.  lds: this wave's 64-float slice.
// (Division items: none here; the factorization ends with one pass over all
// of them, k_factor_divide.)
__device__ __forceinline__ void factor_level(const FactorDev &F, float *__restrict__ val, const LevelSpan &s,
                                             uint32_t tid, uint32_t nthreads, float pert, int perturb, float *lds)
{
    const uint32_t lane = tid & (kWave - 1), w = tid / kWave, nw = nthreads / kWave;
    for (uint32_t e = w; e < s.ml; e += nw) factor_merge_wave(F, val, s.m0 + e, pert, perturb, lane, lds);
    factor_ctasks(F, val, s.c0, s.nc, tid, nthreads, pert, perturb);
    for (uint32_t t = s.ml + tid; t < s.nm; t += nthreads) factor_merge(F, val, s.m0 + t, pert, perturb);
}

// Steps [l0, l1) inside one workgroup, a barrier after each, with the column
// tasks software-pipelined: a wave loads the table data of its next task
// (the next one of the level, or its first one of the next level, issued
// before the barrier) while it runs the current one, so a task waits only
// for its values.  Same items and order as factor_level (no division items:
// the caller divides after the last step; no update reads a divided L
// value, each divides on the fly).
// kFullMerge: see factor_merge_wave.
template <bool kFullMerge = true>
__device__ __forceinline__ void factor_levels_wg(const FactorDev &F, float *__restrict__ val, int l0, int l1,
                                                 float pert, int perturb, float *lds)
{
    const uint32_t tid = threadIdx.x, nt = blockDim.x;
    const uint32_t lane = tid & (kWave - 1), w = tid / kWave, nw = nt / kWave;
    CTab cur;
    if (l0 < l1 && F.ctask_ptr[l0] + w < F.ctask_ptr[l0 + 1]) ctab_load(F, F.ctask_ptr[l0] + w, lane, cur);
    for (int lvl = l0; lvl < l1; ++lvl) {
        const LevelSpan s = level_span(F, lvl);
        for (uint32_t e = w; e < s.ml; e += nw)
            factor_merge_wave<kFullMerge>(F, val, s.m0 + e, pert, perturb, lane, lds);
        for (uint32_t t = s.c0 + w; t < s.c0 + s.nc; t += nw) {
            const bool more = t + nw < s.c0 + s.nc;
            CTab nxt;
            if (more) ctab_load(F, t + nw, lane, nxt);
            ctab_run(F, val, cur, pert, perturb);
            if (more) cur = nxt;
        }
        for (uint32_t t = s.ml + tid; t < s.nm; t += nt) factor_merge(F, val, s.m0 + t, pert, perturb);
        if (lvl + 1 < l1) {   // the next step's table data, loaded across the barrier
            const LevelSpan q = level_span(F, lvl + 1);
            if (q.c0 + w < q.c0 + q.nc) ctab_load(F, q.c0 + w, lane, cur);
        }
        __syncthreads();
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

// One level across the grid; the host sizes the grid with the lanes the
// level needs (column tasks and long merges a wave each).
__global__ void __launch_bounds__(256) k_factor_grid(FactorDev F, float *__restrict__ val,
                                                     const float *__restrict__ d_pert, int lvl, int perturb)
{
    __shared__ __attribute__((aligned(16))) float lds[256 / kWave][kWave];
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    const float pert = perturb ? *d_pert : 0.f;
    factor_level(F, val, level_span(F, lvl), g, gridDim.x * blockDim.x, pert, perturb,
                                           lds[threadIdx.x / kWave]);
}

// ---- interleaved batch steps ----------------------------------------------
// The wide leading steps of a same-pattern batch run over a scratch copy of
// the values with kIl matrices interleaved position by position: matrix
// m = kIl g + j keeps position p at S[(g nnz + p) kIl + j], so the kIl
// lanes that run one item for the kIl matrices of a group touch one 64-byte
// line per value (and share its table data), where matrix-contiguous
// values cost a line per lane.  A wave runs 64 / kIl items.
constexpr uint32_t kIl = 16;
constexpr uint32_t kIlItems = kWave / kIl;   // items per wave

// Column-task item `q` (lane q % 64 of task c0 + q / 64) of one matrix,
// values V[p kIl]: factor_ctasks' arithmetic.
__device__ __forceinline__ void il_ctask(const FactorDev &F, float *__restrict__ V, uint32_t c0, uint32_t q,
                                         float pert, int perturb)
{
    CTab c;
    ctab_load(F, c0 + q / kWave, q % kWave, c);
    if ((c.tt[0] & c.tt[1] & c.tt[2] & c.tt[3]) == ~0u) return;
    const float d = V[(size_t)c.dpos * kIl], a = V[(size_t)c.lpos * kIl];
    float uv[4], tv[4];
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u) {   // branch-free loads, as in ctab_run
        uv[u] = V[(size_t)c.up[u] * kIl];
        tv[u] = V[(size_t)(c.tt[u] != ~0u ? c.tt[u] : 0u) * kIl];
    }
    const float l = a / gesp(d, pert, perturb);
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u)
        if (c.tt[u] != ~0u) V[(size_t)c.tt[u] * kIl] = tv[u] + (-(l * uv[u]));
}

// Column-task items q0 + i kIlItems (i < kIlPer) of one matrix, their loads
// issued together (items of one step are independent: distinct targets,
// none read by another item).  An unused lane's positions are 0, a valid
// address.
constexpr uint32_t kIlPer = 2;   // column-task items per lane

__device__ __forceinline__ void il_ctasks(const FactorDev &F, float *__restrict__ V, uint32_t c0, uint32_t q0,
                                          float pert, int perturb)
{
    CTab c[kIlPer];
#pragma unroll
    for (uint32_t i = 0; i < kIlPer; ++i) {
        const uint32_t q = q0 + i * kIlItems;
        ctab_load(F, c0 + q / kWave, q % kWave, c[i]);
    }
    // every load without a branch (an empty slot's U position is 0 and its
    // target is read at 0, valid addresses), so all are issued together
    float d[kIlPer], a[kIlPer], uv[kIlPer][4], tv[kIlPer][4];
#pragma unroll
    for (uint32_t i = 0; i < kIlPer; ++i) {
        d[i] = V[(size_t)c[i].dpos * kIl];
        a[i] = V[(size_t)c[i].lpos * kIl];
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u) {
            uv[i][u] = V[(size_t)c[i].up[u] * kIl];
            tv[i][u] = V[(size_t)(c[i].tt[u] != ~0u ? c[i].tt[u] : 0u) * kIl];
        }
    }
#pragma unroll
    for (uint32_t i = 0; i < kIlPer; ++i) {
        if ((c[i].tt[0] & c[i].tt[1] & c[i].tt[2] & c[i].tt[3]) == ~0u) continue;
        const float l = a[i] / gesp(d[i], pert, perturb);
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u)
            if (c[i].tt[u] != ~0u) V[(size_t)c[i].tt[u] * kIl] = tv[i][u] + (-(l * uv[i][u]));
    }
}

// Merge entry e of one matrix, values V[p kIl]: factor_merge's arithmetic.
__device__ __forceinline__ void il_merge(const FactorDev &F, float *__restrict__ V, uint32_t e, float pert,
                                         int perturb)
{
    const uint4 r0 = F.mrec[2 * e], r1 = F.mrec[2 * e + 1];
    float v = V[(size_t)r0.x * kIl];
    const float p0 = (V[(size_t)r0.y * kIl] / gesp(V[(size_t)r0.w * kIl], pert, perturb)) * V[(size_t)r0.z * kIl];
    const float p1 = (V[(size_t)r1.x * kIl] / gesp(V[(size_t)r1.z * kIl], pert, perturb)) * V[(size_t)r1.y * kIl];
    v = v + (-p0);
    v = v + (-p1);
    if (r1.w > 2) {
        const uint32_t s1 = F.msrc_ptr[e + 1];
        uint32_t s = F.msrc_ptr[e] + 2;
        for (; s + 4 <= s1; s += 4) {
            MergeSrc m[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) m[u] = F.msrc[s + u];
            __builtin_amdgcn_sched_barrier(0);   // the four table loads issue before any value load
            float p[4];
#pragma unroll
            for (int u = 0; u < 4; ++u)
                p[u] = (V[(size_t)m[u].lpos * kIl] / gesp(V[(size_t)m[u].dpos * kIl], pert, perturb)) *
                       V[(size_t)m[u].upos * kIl];
#pragma unroll
            for (int u = 0; u < 4; ++u) v = v + (-p[u]);
        }
        for (; s < s1; ++s) {
            const MergeSrc m = F.msrc[s];
            v = v + (-((V[(size_t)m.lpos * kIl] / gesp(V[(size_t)m.dpos * kIl], pert, perturb)) *
                       V[(size_t)m.upos * kIl]));
        }
    }
    V[(size_t)r0.x * kIl] = v;
}

// Long merge entry e by one wave for the kIl matrices of a group: lane
// (slot, j) = (lane / kIl, lane % kIl) computes matrix j's products of
// sources base + 4 u + slot (kIlItems sources a pass, kU passes), and every
// lane subtracts matrix j's products in source order, taking them from the
// slot lanes with a permute: factor_merge_wave's arithmetic (sources past
// the end give +0, and v - 0 is v).  The loads take no branch (past the
// end: the last source again; a padding matrix's lanes read its padding
// values), so the kU passes' loads are issued together.
__device__ __forceinline__ void il_merge_long(const FactorDev &F, float *__restrict__ V, uint32_t e, float pert,
                                              int perturb, bool valid, uint32_t lane)
{
    constexpr uint32_t kU = 4;
    const uint32_t t = F.ment[e];
    const uint32_t s0 = F.msrc_ptr[e], s1 = F.msrc_ptr[e + 1];   // s1 > s0 + kLongMerge
    const uint32_t slot = lane / kIl, j = lane % kIl;
    float v = V[(size_t)t * kIl];
    for (uint32_t base = s0; base < s1; base += kIlItems * kU) {
        MergeSrc m[kU];
#pragma unroll
        for (uint32_t u = 0; u < kU; ++u) m[u] = F.msrc[min(base + u * kIlItems + slot, s1 - 1)];
        float lv[kU], dv[kU], uv[kU];
#pragma unroll
        for (uint32_t u = 0; u < kU; ++u) {
            lv[u] = V[(size_t)m[u].lpos * kIl];
            dv[u] = V[(size_t)m[u].dpos * kIl];
            uv[u] = V[(size_t)m[u].upos * kIl];
        }
        float p[kU];
#pragma unroll
        for (uint32_t u = 0; u < kU; ++u) {
            const float q = (lv[u] / gesp(dv[u], pert, perturb)) * uv[u];
            p[u] = base + u * kIlItems + slot < s1 ? q : 0.f;
        }
#pragma unroll
        for (uint32_t u = 0; u < kU; ++u)
#pragma unroll
            for (uint32_t k = 0; k < kIlItems; ++k) {
                const int src = (int)((k * kIl + j) * 4);
                v = v - __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(p[u])));
            }
    }
    if (valid && slot == 0) V[(size_t)t * kIl] = v;
}

// Division item `it` of one matrix, values V[p kIl]: factor_divide's
// arithmetic.
__device__ __forceinline__ void il_divide(float *__restrict__ V, uint2 it, float pert, int perturb)
{
    const uint32_t dpos = it.y & ~kFirstBit;
    const float d0 = V[(size_t)dpos * kIl];
    const bool pt = perturb && fabsf(d0) < pert;
    V[(size_t)it.x * kIl] = V[(size_t)it.x * kIl] / (pt ? pert : d0);
    if (pt && (it.y & kFirstBit)) V[(size_t)dpos * kIl] = pert;
}

// One step for every matrix of a same-pattern batch over the interleaved
// copy S (GESP value of matrix m in pertd[2 m] as a double).  Block b takes
// wave chunk b / groups of the step for group g = b % groups, so the blocks
// that share a chunk's table data are dispatched side by side (the tables
// come from L2, not once per group).  The step's first ml waves take its
// long merges, the rest its column-task items, its short merges, then the
// division items of the previous step's columns, kIlItems a wave: those
// columns are final once that step ran, and this step neither reads nor
// writes them.  A wave takes kIlPer kIlItems items, kIlItems apart (column-
// task items: their loads issued together).
__global__ void __launch_bounds__(256) k_factor_grid_il(FactorDev F, float *__restrict__ S,
                                                        const double *__restrict__ pertd, int lvl, int perturb,
                                                        uint32_t groups, uint32_t batch)
{
    const uint32_t g = blockIdx.x % groups, chunk = blockIdx.x / groups;
    const uint32_t lane = threadIdx.x % kWave, j = lane % kIl;
    const uint32_t gw = chunk * (blockDim.x / kWave) + threadIdx.x / kWave;
    const uint32_t m = g * kIl + j;
    const bool valid = m < batch;
    float *V = S + (size_t)g * F.nnz * kIl + j;
    const float pert = perturb && valid ? (float)pertd[2 * m] : 0.f;
    const LevelSpan s = level_span(F, lvl);
    if (gw < s.ml) {
        il_merge_long(F, V, s.m0 + gw, pert, perturb, valid, lane);
        return;
    }
    if (!valid) return;
    const uint32_t nct = s.nc * kWave, nsm = s.nm - s.ml;
    const uint32_t qa = (gw - s.ml) * kIlPer * kIlItems + lane / kIl, qz = qa + (kIlPer - 1) * kIlItems;
    if (qz < nct) {
        il_ctasks(F, V, s.c0, qa, pert, perturb);
        return;
    }
#pragma unroll 1
    for (uint32_t q = qa; q <= qz; q += kIlItems) {
        if (q < nct) {
            il_ctask(F, V, s.c0, q, pert, perturb);
        } else if (q - nct < nsm) {
            il_merge(F, V, s.m0 + s.ml + (q - nct), pert, perturb);
        } else if (lvl > 0) {
            const uint32_t d = F.dvd_ptr[lvl - 1] + (q - nct - nsm);
            if (d < F.dvd_ptr[lvl]) il_divide(V, F.dvd[d], pert, perturb);
        }
    }
}

// Zeros on the fill-in lines of group blockIdx.y (a line: one position's 16
// matrices, four float4 a line).
__global__ void __launch_bounds__(256) k_il_zero_fill(const uint32_t *__restrict__ fill, uint32_t nfill,
                                                      float4 *__restrict__ S, uint32_t nnz)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= 4 * nfill) return;
    S[((size_t)blockIdx.y * nnz + fill[i / 4]) * 4 + i % 4] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// S = 0 (float4 stores; count4 = floats / 4).
__global__ void k_il_zero(float4 *__restrict__ S, size_t count4)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count4; i += (size_t)gridDim.x * blockDim.x)
        S[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// The scaled entries of A into S (wg_values' step 1, interleaved): block
// (x, g) takes entries [64 x, 64 x + 64) of the kIl matrices of group g,
// read matrix by matrix (whole lines of ax), scaled, and turned in LDS so
// that each entry's kIl values go out as one line.
__global__ void __launch_bounds__(256) k_il_scatter(RefineDev R, const double *__restrict__ ax,
                                                    float *__restrict__ S, uint32_t nnz, uint32_t batch)
{
    constexpr uint32_t kK = 64;
    __shared__ float tile[kK][kIl + 1];
    const uint32_t g = blockIdx.y, k0 = blockIdx.x * kK, tid = threadIdx.x;
    // loads without branches (past the last entry or matrix: the last one
    // again, never stored), so that a thread's four issue together
    float v[kK * kIl / 256];
#pragma unroll
    for (uint32_t u = 0; u < kK * kIl / 256; ++u) {
        const uint32_t i = tid + 256 * u, j = i / kK, k = min(k0 + i % kK, R.nnz_a - 1);
        const uint32_t m = min(g * kIl + j, batch - 1);
        v[u] = (float)((ax[(size_t)m * R.nnz_a + k] * R.a_rs[k]) * R.a_cs[k]);
    }
#pragma unroll
    for (uint32_t u = 0; u < kK * kIl / 256; ++u) {
        const uint32_t i = tid + 256 * u;
        tile[i % kK][i / kK] = v[u];
    }
    __syncthreads();
    for (uint32_t i = tid; i < kK * kIl; i += blockDim.x) {
        const uint32_t kk = i / kIl, j = i % kIl, k = k0 + kk;
        if (k < R.nnz_a && g * kIl + j < batch) S[((size_t)g * nnz + R.a_dst[k]) * kIl + j] = tile[kk][j];
    }
}

// GESP: bits[m] = max over columns of sum_p |val[p]| (FP32, column order)
// for the matrices of group blockIdx.y, then k_il_pert_finish.
__global__ void k_il_pert(const uint32_t *__restrict__ col_ptr, const float *__restrict__ S, uint32_t n,
                          uint32_t nnz, uint32_t batch, unsigned int *__restrict__ bits)
{
    const uint32_t g = blockIdx.y, i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t c = i / kIl, j = i % kIl, m = g * kIl + j;
    if (c >= n || m >= batch) return;
    const float *V = S + (size_t)g * nnz * kIl + j;
    float s = 0.f;
    for (uint32_t p = col_ptr[c]; p < col_ptr[c + 1]; ++p) s += fabsf(V[(size_t)p * kIl]);
    if (s == s) atomicMax(bits + m, __float_as_uint(s));
}

// pertd[2 m] = the GESP value; re-arms bits for the next call.
__global__ void k_il_pert_finish(unsigned int *__restrict__ bits, double *__restrict__ pertd, uint32_t batch)
{
    const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= batch) return;
    pertd[2 * m] = (double)(kPertFactor * __uint_as_float(bits[m]));
    bits[m] = 0u;
}

// S (interleaved) -> lu (matrix-contiguous), 256 positions of one group a
// block through LDS: whole-line reads and writes.
__global__ void __launch_bounds__(256) k_il_to_rows(const float *__restrict__ S, float *__restrict__ lu,
                                                    uint32_t nnz, uint32_t batch)
{
    __shared__ float tile[kIl][256 + 1];
    const uint32_t g = blockIdx.y, p0 = blockIdx.x * 256, tid = threadIdx.x;
    const float4 *src = reinterpret_cast<const float4 *>(S + (size_t)g * nnz * kIl);
    // a position's kIl values are kIl / 4 float4; the loads take no branch
    // (past the last position: the last one again), so a thread's four
    // issue together
    constexpr uint32_t kF = 256 * kIl / 4 / 256;
    float4 v[kF];
#pragma unroll
    for (uint32_t u = 0; u < kF; ++u) {
        const uint32_t f = tid + 256 * u, p = min(p0 + f / (kIl / 4), nnz - 1);
        v[u] = src[(size_t)p * (kIl / 4) + f % (kIl / 4)];
    }
#pragma unroll
    for (uint32_t u = 0; u < kF; ++u) {
        const uint32_t f = tid + 256 * u, p = f / (kIl / 4), j = (f % (kIl / 4)) * 4;
        tile[j][p] = v[u].x;
        tile[j + 1][p] = v[u].y;
        tile[j + 2][p] = v[u].z;
        tile[j + 3][p] = v[u].w;
    }
    __syncthreads();
    if (p0 + tid >= nnz) return;
    for (uint32_t j = 0; j < kIl && g * kIl + j < batch; ++j)
        lu[(size_t)(g * kIl + j) * nnz + p0 + tid] = tile[j][tid];
}

// A run of narrow levels inside one workgroup, a barrier between levels.
__global__ void __launch_bounds__(1024) k_factor_fused(FactorDev F, float *__restrict__ val,
                                                       const float *__restrict__ d_pert, int l0, int l1, int perturb)
{
    __shared__ __attribute__((aligned(16))) float lds[1024 / kWave][kWave];
    const float pert = perturb ? *d_pert : 0.f;
    factor_levels_wg(F, val, l0, l1, pert, perturb, lds[threadIdx.x / kWave]);
}

// The divisions L(i,k) /= U(k,k) of every column, after the last step (no
// step reads a divided L value: updates divide on the fly; no step writes
// L(:,k) or U(k,k) after column k's step).
__global__ void __launch_bounds__(256) k_factor_divide(FactorDev F, float *__restrict__ val,
                                                       const float *__restrict__ d_pert, int perturb)
{
    const float pert = perturb ? *d_pert : 0.f;
    const uint32_t nd = F.dvd_ptr[F.num_lev];
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < nd; t += gridDim.x * blockDim.x)
        factor_divide(val, F.dvd[t], pert, perturb);
}

// A window's view of the tables: the same level spans, items that address
// LDS slots.
__device__ __forceinline__ FactorDev window_view(const FactorDev &F, const WinDev &W)
{
    FactorDev T = F;
    T.cln = W.cln;
    T.ctgt = W.ctgt;
    T.cup = W.cup;
    T.ment = W.ment;
    T.msrc = W.msrc;
    T.mrec = W.mrec;
    T.dvd = W.dvd;
    return T;
}

__device__ __forceinline__ WinDev load_window(const FactorDev &F, int w)
{
    WinDev W = F.win[w];
    W.pos = glb(W.pos);
    W.cln = glb(W.cln);
    W.ctgt = glb(W.ctgt);
    W.cup = glb(W.cup);
    W.ment = glb(W.ment);
    W.msrc = glb(W.msrc);
    W.mrec = glb(W.mrec);
    W.dvd = glb(W.dvd);
    return W;
}

// Steps [W.l0, W.l1) inside one workgroup with the values they touch staged
// in LDS (slots): load, one barrier per step, store back.
template <bool kFullMerge = true>
__device__ __forceinline__ void factor_window(const FactorDev &F, const WinDev &W, float *__restrict__ val,
                                              float *slots, float pert, int perturb, float *wlds)
{
    const uint32_t tid = threadIdx.x, nt = blockDim.x;
    for (uint32_t s = tid; s < W.nslot; s += nt) slots[s] = val[W.pos[s]];
    __syncthreads();
    const FactorDev T = window_view(F, W);
    factor_levels_wg<kFullMerge>(T, slots, W.l0, W.l1, pert, perturb, wlds);
    for (uint32_t s = tid; s < W.nslot; s += nt) val[W.pos[s]] = slots[s];
    __syncthreads();
}

// Steps l0.. of one factorization by one workgroup: windows in LDS, the
// steps between them in global memory; l0 is 0 or lies at or before the
// first window's first step (the steps before it ran as
// k_factor_grid_il launches).  The divisions L(i,k) /= U(k,k) of every
// column run after the last step, all at once: no step reads a divided L
// value (updates divide on the fly) and no step writes L(:,k) or U(k,k)
// after column k's step, so the values are those of the per-step order,
// one barrier-bound pass per step less.
__device__ __forceinline__ void factor_all_wg(const FactorDev &F, float *__restrict__ val, float *slots, float pert,
                                              int perturb, float *wlds, int l0)
{
    int l = l0;
    for (int w = 0; w < F.nwin; ++w) {
        const WinDev W = load_window(F, w);
        factor_levels_wg<false>(F, val, l, W.l0, pert, perturb, wlds);
        factor_window<false>(F, W, val, slots, pert, perturb, wlds);
        l = W.l1;
    }
    factor_levels_wg<false>(F, val, l, F.num_lev + 1, pert, perturb, wlds);
    // the divisions, eight items a lane at a time (their loads issued
    // together); items are independent: an item writes its L entry and, for
    // a perturbed pivot, the pivot, which the column's other items divide
    // by either way (see factor_divide).  Those of the columns of steps
    // before l0 - 1 ran with the interleaved steps (k_factor_grid_il).
    const uint32_t nd = F.dvd_ptr[F.num_lev], tid = threadIdx.x, nt = blockDim.x;
    constexpr int U = 8;
    for (uint32_t t0 = (l0 > 0 ? F.dvd_ptr[l0 - 1] : 0u) + tid; t0 < nd; t0 += U * nt) {
        // loads without branches (past the end: the last item again), so
        // that the eight items' table loads, then value loads, issue together
        uint2 it[U];
        float d[U], x[U];
#pragma unroll
        for (int u = 0; u < U; ++u) it[u] = F.dvd[min(t0 + u * nt, nd - 1)];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            d[u] = val[it[u].y & ~kFirstBit];
            x[u] = val[it[u].x];
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (t0 + u * nt < nd) {
                const uint32_t dpos = it[u].y & ~kFirstBit;
                const bool pt = perturb && fabsf(d[u]) < pert;
                val[it[u].x] = x[u] / (pt ? pert : d[u]);
                if (pt && (it[u].y & kFirstBit)) val[dpos] = pert;
            }
    }
    __syncthreads();
}

extern __shared__ float g_lds_vec[];   // dynamic LDS of the kernels below

// Window w of one factorization in one workgroup (dynamic LDS: its values).
__global__ void __launch_bounds__(1024) k_factor_window(FactorDev F, float *__restrict__ val,
                                                        const float *__restrict__ d_pert, int w, int perturb)
{
    __shared__ __attribute__((aligned(16))) float lds[1024 / kWave][kWave];
    const float pert = perturb ? *d_pert : 0.f;
    const WinDev W = load_window(F, w);
    factor_window(F, W, val, g_lds_vec, pert, perturb, lds[threadIdx.x / kWave]);
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

// A long row by one wave, with the reference's arithmetic: products of 64
// entries at a time in parallel (each rounded, as in the sequential loop),
// then lane 0 subtracts them in the row's order (ascending, or descending
// for the U sweep) from the wave's LDS slice.
__device__ __forceinline__ void tri_row_wave(const TriDev &T, const float *__restrict__ lu,
                                             float *__restrict__ x, uint32_t r, uint32_t lane, float *lds)
{
    const uint32_t b = T.beg[r] + T.begoff, e = T.end[r];
    float acc = x[r];
    const uint32_t nchunk = (e - b + kWave - 1) / kWave;
    for (uint32_t c = 0; c < nchunk; ++c) {
        // ascending: chunk c covers [b + 64c, ...); descending: from the top
        const uint32_t lo = T.desc ? (e - b > (c + 1) * kWave ? e - (c + 1) * kWave : b) : b + c * kWave;
        const uint32_t hi = T.desc ? e - c * kWave : min(e, lo + kWave);
        const uint32_t q = lo + lane;
        float p = 0.f;
        if (q < hi) p = lu[T.vpos ? T.vpos[q] : q] * x[T.idx[q]];
        lds[lane] = p;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        if (lane == 0) acc = serial_sub(acc, lds, kWave, T.desc != 0);   // lanes past hi - lo hold +0
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
    if (lane == 0) {
        if (T.dpos) acc = acc / lu[T.dpos[r]];
        x[r] = acc;
    }
}

// One level across the grid: waves [0, nlong) take the level's long rows,
// the remaining lanes one short row each.
__global__ void __launch_bounds__(256) k_tri_grid(TriDev T, const float *__restrict__ lu, float *__restrict__ x,
                                                  int lvl)
{
    __shared__ __attribute__((aligned(16))) float lds[256 / kWave][kWave];
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t r0 = T.lvl_ptr[lvl], nlong = T.lvl_long[lvl];
    const uint32_t w = g / kWave;
    if (w < nlong) {
        tri_row_wave(T, lu, x, T.rows[r0 + w], g & (kWave - 1), lds[threadIdx.x / kWave]);
        return;
    }
    const uint32_t t = r0 + nlong + (g - nlong * kWave);
    if (t < T.lvl_ptr[lvl + 1]) tri_row(T, lu, x, T.rows[t]);
}

__global__ void __launch_bounds__(1024) k_tri_fused(TriDev T, const float *__restrict__ lu, float *__restrict__ x,
                                                    int l0, int l1)
{
    __shared__ __attribute__((aligned(16))) float lds[1024 / kWave][kWave];
    const uint32_t lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave, nw = blockDim.x / kWave;
    for (int lvl = l0; lvl < l1; ++lvl) {
        const uint32_t r0 = T.lvl_ptr[lvl], nlong = T.lvl_long[lvl];
        for (uint32_t i = w; i < nlong; i += nw) tri_row_wave(T, lu, x, T.rows[r0 + i], lane, lds[w]);
        for (uint32_t t = r0 + nlong + threadIdx.x; t < T.lvl_ptr[lvl + 1]; t += blockDim.x)
            tri_row(T, lu, x, T.rows[t]);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// dataflow (barrier-free) triangular sweeps inside one workgroup
// ---------------------------------------------------------------------------
//
// Rows are taken in the topological (level) order of T.rows, grouped into
// tasks of up to 64 short rows or one long row; wave w runs tasks w, w + nw,
// ... in order.  A row waits for each dependency's ready bit (LDS) before
// reading its value, so no level barrier is needed.  Deadlock freedom: the
// first unfinished task always has all its dependencies in earlier
// (finished) tasks or inside itself, and its wave is at it; lanes of one
// task poll cooperatively (a wave-uniform loop), so a row blocked on a row
// of the same task never stops that row from completing.

__device__ __forceinline__ float wave_sum(float v)
{
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// Ready bit of row c.  df_peek is a relaxed read: a caller that reads x
// after it issues an acquire fence first.
__device__ __forceinline__ bool df_peek(const uint32_t *flags, uint32_t c)
{
    return (__hip_atomic_load(flags + (c >> 5), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >> (c & 31u)) & 1u;
}

__device__ __forceinline__ bool df_ready(const uint32_t *flags, uint32_t c)
{
    return (__hip_atomic_load(flags + (c >> 5), __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) >> (c & 31u)) & 1u;
}

// Bound on the polls of one wait for a ready bit (far above any real wait:
// a whole sweep takes at most tens of milliseconds, a poll tens of
// nanoseconds).
constexpr uint32_t kSpinCap = 1u << 24;

__device__ __forceinline__ void df_set(uint32_t *flags, uint32_t r)
{
    __hip_atomic_fetch_or(flags + (r >> 5), 1u << (r & 31u), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Waits for ready bits.  Every dependency of a row is finished in an
// earlier task or earlier in the row's own task, and a task's lanes poll
// together, so every wait ends; the plan builder checks exactly this on the
// host for every sweep table before any launch (an inconsistent table is
// refused with GLU_ERR_INVALID).  On top of that, the waits of the
// level-synchronous path's dataflow runs and of long lane rows are bounded
// (kSpinCap polls): a row whose wait gives up completes with NaN, so its
// dependents, and every wave, still finish, and the NaN result is reported
// (glu_refine: GLU_ERR_NUMERIC; batched solves: eta = NaN).  The waits of
// the short-row and wave-row paths, the hot loops of the batched
// refinement, stay unbounded: a bound there costs ~4% of the batch step.
__device__ __forceinline__ float df_nan() { return __builtin_nanf(""); }

// x[c] once row c's ready bit is set (acquire), polling at most kSpinCap
// times (with s_sleep between polls when kSleep); NaN if the wait gives up.
// A ready bit that is already set costs one load and a branch.
template <bool kSleep>
__device__ __forceinline__ float df_get(const uint32_t *flags, const float *x, uint32_t c)
{
    if (!df_ready(flags, c)) {
        uint32_t s = 0;
        do {
            if (kSleep) __builtin_amdgcn_s_sleep(1);
            if (++s >= kSpinCap) return df_nan();
        } while (!df_ready(flags, c));
    }
    return x[c];
}

// Up to 64 rows of at most kDfLaneRow entries, one per lane: each lane
// loads all its row's entries at once and consumes them in the row's order
// (ascending, or descending for the U sweep) as they become ready — the
// sequential arithmetic of tri_row.  (Rows longer than K would be refilled
// K at a time; the task builder gives them a wave instead.)
__device__ __forceinline__ void df_short_task(const TriDev &T, const float *__restrict__ vals, float *x,
                                              uint32_t *flags, uint32_t p0, uint32_t p1, uint32_t lane)
{
    constexpr uint32_t K = kDfLaneRow;
    const uint32_t pos = p0 + lane;
    bool pending = pos < p1;
    uint32_t r = 0, q = 0, qend = 0;
    float acc = 0.f;
    if (pending) {
        r = T.rows[pos];
        const uint32_t b = T.beg[r] + T.begoff, e = T.end[r];
        acc = x[r];
        q = T.desc ? e : b;
        qend = T.desc ? b : e;
    }
    uint32_t cidx[K];
    float cval[K];
    uint32_t cn = 0, ci = 0, stalls = 0;
    while (__ballot(pending)) {
        bool progressed = false;
        if (pending) {
            if (ci == cn && q != qend) {
                const uint32_t left = T.desc ? q - qend : qend - q;
                cn = left < K ? left : K;
                ci = 0;
#pragma unroll
                for (uint32_t j = 0; j < K; ++j)
                    if (j < cn) {
                        const uint32_t qq = T.desc ? q - 1 - j : q + j;
                        cidx[j] = T.idx[qq];
                        cval[j] = vals[T.vpos ? T.vpos[qq] : qq];
                    }
                q = T.desc ? q - cn : q + cn;
            }
            // ready bits of the chunk's unconsumed entries (independent
            // relaxed reads into one mask), one acquire fence, then
            // in-order consumption of the ready prefix
            uint32_t mask = 0;
#pragma unroll
            for (uint32_t j = 0; j < K; ++j)
                if (j >= ci && j < cn && df_peek(flags, cidx[j])) mask |= 1u << j;
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
#pragma unroll
            for (uint32_t j = 0; j < K; ++j)
                if (j == ci && ((mask >> j) & 1u)) {
                    acc = acc - cval[j] * x[cidx[j]];
                    ++ci;
                    progressed = true;
                }
            if (ci == cn && q == qend) {
                if (T.dpos) acc = acc / vals[T.dpos[r]];
                x[r] = acc;
                df_set(flags, r);
                pending = false;
                progressed = true;
            }
        }
        if (!__ballot(progressed)) {
            if (++stalls >= kSpinCap && pending) {   // give up: the row completes with NaN
                x[r] = df_nan();
                df_set(flags, r);
                pending = false;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
}

// One long row by a wave.  kExact: products of 64 entries in parallel, then
// lane 0 subtracts them in the row's order (tri_row_wave's arithmetic);
// otherwise a wave reduction (refinement sweeps, where order is free).
template <bool kExact>
__device__ __forceinline__ void df_long_task(const TriDev &T, const float *__restrict__ vals, float *x,
                                             uint32_t *flags, uint32_t pos, uint32_t lane, float *lds)
{
    const uint32_t r = T.rows[pos];
    const uint32_t b = T.beg[r] + T.begoff, e = T.end[r];
    if (!kExact) {
        float sum = 0.f;
        for (uint32_t q = b + lane; q < e; q += kWave) {
            const uint32_t c = T.idx[q];
            const float v = vals[T.vpos ? T.vpos[q] : q];
            sum += v * df_get<true>(flags, x, c);
        }
        sum = wave_sum(sum);
        if (lane == 0) {
            float acc = x[r] - sum;
            if (T.dpos) acc = acc / vals[T.dpos[r]];
            x[r] = acc;
            df_set(flags, r);
        }
        return;
    }
    float acc = x[r];
    const uint32_t nchunk = (e - b + kWave - 1) / kWave;
    for (uint32_t c = 0; c < nchunk; ++c) {
        const uint32_t lo = T.desc ? (e - b > (c + 1) * kWave ? e - (c + 1) * kWave : b) : b + c * kWave;
        const uint32_t hi = T.desc ? e - c * kWave : min(e, lo + kWave);
        const uint32_t q = lo + lane;
        float p = 0.f;
        if (q < hi) {
            const uint32_t cc = T.idx[q];
            const float v = vals[T.vpos ? T.vpos[q] : q];
            p = v * df_get<true>(flags, x, cc);
        }
        lds[lane] = p;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        if (lane == 0) acc = serial_sub(acc, lds, kWave, T.desc != 0);   // lanes past hi - lo hold +0
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
    if (lane == 0) {
        if (T.dpos) acc = acc / vals[T.dpos[r]];
        x[r] = acc;
        df_set(flags, r);
    }
}

// Tasks [k0, k1) of a sweep for the calling workgroup.  The caller has
// cleared the ready bits of the tasks' rows and set all others.
template <bool kExact>
__device__ __forceinline__ void df_sweep(const TriDev &T, const float *__restrict__ vals, float *x, uint32_t *flags,
                                         uint32_t k0, uint32_t k1, float *lds)
{
    const uint32_t lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave, nw = blockDim.x / kWave;
    for (uint32_t k = k0 + w; k < k1; k += nw) {
        const uint32_t a = T.dtask[k];
        if (a & kLongTask)
            df_long_task<kExact>(T, vals, x, flags, a & ~kLongTask, lane, lds);
        else
            df_short_task(T, vals, x, flags, a, T.dtask[k + 1] & ~kLongTask, lane);
    }
}

// A run of narrow levels [l0, l1) of one sweep (tasks [k0, k1)) inside one
// workgroup without level barriers, x in global memory; bit-exact with the
// reference's sequential row arithmetic.  Dynamic LDS: ceil(n/32) words.
__global__ void __launch_bounds__(1024) k_tri_df(TriDev T, const float *__restrict__ lu, float *__restrict__ x,
                                                 int l0, int l1, uint32_t k0, uint32_t k1, uint32_t n)
{
    __shared__ __attribute__((aligned(16))) float lds[1024 / kWave][kWave];
    uint32_t *flags = reinterpret_cast<uint32_t *>(g_lds_vec);
    const uint32_t nwords = (n + 31) / 32;
    for (uint32_t i = threadIdx.x; i < nwords; i += blockDim.x) flags[i] = 0xffffffffu;
    __syncthreads();
    for (uint32_t t = T.lvl_ptr[l0] + threadIdx.x; t < T.lvl_ptr[l1]; t += blockDim.x) {
        const uint32_t r = T.rows[t];
        atomicAnd(flags + (r >> 5), ~(1u << (r & 31u)));
    }
    __syncthreads();
    df_sweep<true>(T, lu, x, flags, k0, k1, lds[threadIdx.x / kWave]);
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

// lu[p] = float((ax[src[p]] * rs_col[col[p]]) * csp[row[p]]), 0 for fill.
__device__ __forceinline__ void refresh_entry(const RefineDev &R, const double *__restrict__ ax,
                                              float *__restrict__ lu, uint32_t p)
{
    const uint32_t s = R.lu_src[p];
    lu[p] = (s == 0xffffffffu) ? 0.f : (float)((ax[s] * R.rs_col[R.lu_col[p]]) * R.csp[R.row_idx[p]]);
}

__global__ void k_refresh(RefineDev R, const double *__restrict__ ax, float *__restrict__ lu, uint32_t nnz_lu,
                          int batch)
{
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (uint64_t)nnz_lu * batch) return;
    const uint32_t b = (uint32_t)(g / nnz_lu), p = (uint32_t)(g % nnz_lu);
    refresh_entry(R, ax + (uint64_t)b * R.nnz_a, lu + (uint64_t)b * nnz_lu, p);
}

__device__ __forceinline__ double wave_sum64(double v)
{
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// Entries [q0, q1) of row i of A (transpose = 0: the CSR of A) or of A'
// (transpose = 1: column i of the caller's CSC).
struct ARow {
    uint32_t q0, q1;
};

__device__ __forceinline__ ARow a_row(const RefineDev &R, uint32_t i, int transpose)
{
    return transpose ? ARow{R.a_cptr[i], R.a_cptr[i + 1]} : ARow{R.a_rptr[i], R.a_rptr[i + 1]};
}

// sum_q a_q * x[col_q] and sum_q |a_q| over entries [q0, q0 + step, ...) of row i
__device__ __forceinline__ void a_row_partial(const RefineDev &R, const double *__restrict__ ax,
                                              const double *__restrict__ x, ARow row, uint32_t first,
                                              uint32_t step, int transpose, double &dot, double &asum)
{
    dot = 0.0;
    asum = 0.0;
    for (uint32_t q = row.q0 + first; q < row.q1; q += step) {
        const double a = transpose ? ax[q] : ax[R.a_src[q]];
        if (x) dot = fma(a, x[transpose ? R.a_ridx[q] : R.a_cidx[q]], dot);
        asum += fabs(a);
    }
}

// Rows of A for one cooperating group: long rows one wave each (FP64 wave
// reductions), the rest one lane each.  f(i, dot, asum) is called once per
// row by one lane.
template <typename F>
__device__ __forceinline__ void for_rows_of_a(const RefineDev &R, const double *__restrict__ ax,
                                              const double *__restrict__ x, uint32_t n, int transpose,
                                              uint32_t tid, uint32_t nthreads, F f)
{
    const uint32_t lane = tid & (kWave - 1), w = tid / kWave, nw = nthreads / kWave;
    const uint32_t *rows = transpose ? R.a_cols : R.a_rows;
    const uint32_t nl = transpose ? R.a_nlong_t : R.a_nlong;
    for (uint32_t t = w; t < nl; t += nw) {
        const uint32_t i = rows[t];
        double dot, asum;
        a_row_partial(R, ax, x, a_row(R, i, transpose), lane, kWave, transpose, dot, asum);
        dot = wave_sum64(dot);
        asum = wave_sum64(asum);
        if (lane == 0) f(i, dot, asum);
    }
    for (uint32_t t = nl + tid; t < n; t += nthreads) {
        const uint32_t i = rows[t];
        double dot, asum;
        a_row_partial(R, ax, x, a_row(R, i, transpose), 0, 1, transpose, dot, asum);
        f(i, dot, asum);
    }
}

// for_rows_of_a for one workgroup, with the short rows taken in natural
// order (long ones skipped: their waves take them), four rows per lane in
// flight and four entries of each per step, so a lane keeps up to 16
// independent gathers outstanding.
template <typename F>
__device__ __forceinline__ void wg_rows_of_a(const RefineDev &R, const double *__restrict__ ax,
                                             const double *__restrict__ x, uint32_t n, F f)
{
    constexpr int RB = 2, EB = 4;
    const uint32_t tid = threadIdx.x, nt = blockDim.x;
    const uint32_t lane = tid & (kWave - 1), w = tid / kWave, nw = nt / kWave;
    for (uint32_t t = w; t < R.a_nlong; t += nw) {
        const uint32_t i = R.a_rows[t];
        double dot, asum;
        a_row_partial(R, ax, x, a_row(R, i, 0), lane, kWave, 0, dot, asum);
        dot = wave_sum64(dot);
        asum = wave_sum64(asum);
        if (lane == 0) f(i, dot, asum);
    }
    for (uint32_t i0 = tid; i0 < n; i0 += RB * nt) {
        uint32_t q[RB], qe[RB];
        double dot[RB], asum[RB];
#pragma unroll
        for (int r = 0; r < RB; ++r) {
            const uint32_t i = i0 + r * nt;
            q[r] = qe[r] = 0;
            dot[r] = asum[r] = 0.0;
            if (i < n) {
                q[r] = R.a_rptr[i];
                qe[r] = R.a_rptr[i + 1];
                if (qe[r] - q[r] > kLongRow) qe[r] = q[r];   // a wave's row
            }
        }
        for (;;) {
            bool more = false;
#pragma unroll
            for (int r = 0; r < RB; ++r) more |= q[r] < qe[r];
            if (!more) break;
            uint32_t src[RB][EB], col[RB][EB];
#pragma unroll
            for (int r = 0; r < RB; ++r)
#pragma unroll
                for (int e = 0; e < EB; ++e)
                    if (q[r] + e < qe[r]) {
                        src[r][e] = R.a_src[q[r] + e];
                        col[r][e] = R.a_cidx[q[r] + e];
                    }
            double av[RB][EB], xv[RB][EB];
#pragma unroll
            for (int r = 0; r < RB; ++r)
#pragma unroll
                for (int e = 0; e < EB; ++e)
                    if (q[r] + e < qe[r]) {
                        av[r][e] = ax[src[r][e]];
                        xv[r][e] = x ? x[col[r][e]] : 0.0;
                    }
#pragma unroll
            for (int r = 0; r < RB; ++r)
#pragma unroll
                for (int e = 0; e < EB; ++e)
                    if (q[r] + e < qe[r]) {
                        if (x) dot[r] = fma(av[r][e], xv[r][e], dot[r]);
                        asum[r] += fabs(av[r][e]);
                    }
#pragma unroll
            for (int r = 0; r < RB; ++r) q[r] = q[r] + EB < qe[r] ? q[r] + EB : qe[r];
        }
#pragma unroll
        for (int r = 0; r < RB; ++r) {
            const uint32_t i = i0 + r * nt;
            if (i < n && R.a_rptr[i + 1] - R.a_rptr[i] <= kLongRow) f(i, dot[r], asum[r]);
        }
    }
}

// norms[0] = |r|_inf, [1] = |x|_inf, [2] = |A|_inf, [3] = |b|_inf (as u64
// bits of non-negative doubles; NaN compares largest).
__device__ __forceinline__ void umax64(unsigned long long *p, double v)
{
    unsigned long long bits = (v == v) ? (unsigned long long)__double_as_longlong(v) : 0x7ff8000000000000ull;
    atomicMax(p, bits);
}

// r = b - A x (or A' x); max norms.  Grid: one wave per long row plus one
// lane per short row.
__global__ void __launch_bounds__(256) k_residual(RefineDev R, const double *__restrict__ ax,
                                                  const double *__restrict__ b, const double *__restrict__ x,
                                                  double *__restrict__ r, uint32_t n, int transpose, int with_anorm,
                                                  unsigned long long *__restrict__ norms)
{
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = g & (kWave - 1), w = g / kWave;
    const uint32_t *rows = transpose ? R.a_cols : R.a_rows;
    const uint32_t nl = transpose ? R.a_nlong_t : R.a_nlong;
    double rmax = 0.0, xmax = 0.0, amax = 0.0, bmax = 0.0;
    auto take = [&](uint32_t i, double dot, double asum) {
        const double ri = b[i] - dot;
        r[i] = ri;
        rmax = nmax(rmax, fabs(ri));
        xmax = nmax(xmax, fabs(x[i]));
        bmax = nmax(bmax, fabs(b[i]));
        amax = nmax(amax, asum);
    };
    if (w < nl) {
        const uint32_t i = rows[w];
        double dot, asum;
        a_row_partial(R, ax, x, a_row(R, i, transpose), lane, kWave, transpose, dot, asum);
        dot = wave_sum64(dot);
        asum = wave_sum64(asum);
        if (lane == 0) take(i, dot, asum);
    } else {
        const uint32_t t = nl + (g - nl * kWave);
        if (t < n) {
            const uint32_t i = rows[t];
            double dot, asum;
            a_row_partial(R, ax, x, a_row(R, i, transpose), 0, 1, transpose, dot, asum);
            take(i, dot, asum);
        }
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
    float *lucsr;       // df_nval scratch: the sweeps' value copy
    double *norms;      // 2 scratch: [1] = |b|_inf (factor -> refine kernel), [0] unused
    double *kry;        // (3 krylov + 1) n scratch: Arnoldi vectors V, Z = M V, W = A Z (null: plain refinement)
    int32_t *iters;
    double *eta;
};

// One triangular sweep of the refinement inside a workgroup (x may be LDS).
// Rows longer than kLongRow (listed first in each level) take a wave each
// and sum their products with a wave reduction; the rest take a lane each
// and accumulate sequentially.  This is the correction solve of iterative
// refinement, where the summation order is free; glu_solve keeps the
// reference's sequential order for every row.
__device__ __forceinline__ void wg_tri_refine(const TriDev &T, const float *__restrict__ lu, float *x)
{
    const uint32_t tid = threadIdx.x, nt = blockDim.x;
    const uint32_t lane = tid & (kWave - 1), w = tid / kWave, nw = nt / kWave;
    for (int lvl = 0; lvl < T.nlev; ++lvl) {
        const uint32_t r0 = T.lvl_ptr[lvl], r1 = T.lvl_ptr[lvl + 1], nl = T.lvl_long[lvl];
        for (uint32_t t = w; t < nl; t += nw) {
            const uint32_t r = T.rows[r0 + t];
            const uint32_t b = T.beg[r] + T.begoff, e = T.end[r];
            float sum = 0.f;
            for (uint32_t q = b + lane; q < e; q += kWave) {
                const uint32_t p = T.vpos ? T.vpos[q] : q;
                sum += lu[p] * x[T.idx[q]];
            }
            sum = wave_sum(sum);
            if (lane == 0) {
                float acc = x[r] - sum;
                if (T.dpos) acc = acc / lu[T.dpos[r]];
                x[r] = acc;
            }
        }
        for (uint32_t t = r0 + nl + tid; t < r1; t += nt) tri_row(T, lu, x, T.rows[t]);
        __syncthreads();
    }
}


__device__ __forceinline__ void globalize(FactorDev &F)
{
    F.col_ptr = glb(F.col_ptr);
    F.ctask_ptr = glb(F.ctask_ptr);
    F.cln = glb(F.cln);
    F.ctgt = glb(F.ctgt);
    F.cup = glb(F.cup);
    F.ment = glb(F.ment);
    F.ment_ptr = glb(F.ment_ptr);
    F.ment_long = glb(F.ment_long);
    F.msrc_ptr = glb(F.msrc_ptr);
    F.msrc = glb(F.msrc);
    F.mrec = glb(F.mrec);
    F.dvd = glb(F.dvd);
    F.dvd_ptr = glb(F.dvd_ptr);
    F.win = glb(F.win);
}

__device__ __forceinline__ void globalize(TriDev &T)
{
    T.beg = glb(T.beg);
    T.end = glb(T.end);
    T.idx = glb(T.idx);
    if (T.vpos) T.vpos = glb(T.vpos);
    if (T.dpos) T.dpos = glb(T.dpos);
    T.rows = glb(T.rows);
    T.lvl_ptr = glb(T.lvl_ptr);
    T.lvl_long = glb(T.lvl_long);
    T.dtask = glb(T.dtask);
}

__device__ __forceinline__ void globalize(DfDev &D)
{
    D.task = glb(D.task);
    D.rl = glb(D.rl);
    D.idx = glb(D.idx);
}

__device__ __forceinline__ void globalize(RefineDev &R)
{
    R.a_rptr = glb(R.a_rptr);
    R.a_cidx = glb(R.a_cidx);
    R.a_src = glb(R.a_src);
    R.a_cptr = glb(R.a_cptr);
    R.a_ridx = glb(R.a_ridx);
    R.lu_src = glb(R.lu_src);
    R.row_idx = glb(R.row_idx);
    R.lu_col = glb(R.lu_col);
    R.a_rows = glb(R.a_rows);
    R.a_cols = glb(R.a_cols);
    R.rs_col = glb(R.rs_col);
    R.csp = glb(R.csp);
    R.a_dst = glb(R.a_dst);
    R.a_rs = glb(R.a_rs);
    R.a_cs = glb(R.a_cs);
}

__device__ __forceinline__ void globalize(MatIO &io)
{
    io.ax = glb(io.ax);
    io.b = glb(io.b);
    io.x = glb(io.x);
    io.lu = glb(io.lu);
    io.t32 = glb(io.t32);
    io.r64 = glb(io.r64);
    if (io.lucsr) io.lucsr = glb(io.lucsr);
    io.norms = glb(io.norms);
    if (io.kry) io.kry = glb(io.kry);
    io.iters = glb(io.iters);
    io.eta = glb(io.eta);
}

// ---------------------------------------------------------------------------
// pipelined dataflow sweeps (refinement of the batch)
// ---------------------------------------------------------------------------
//
// The same dataflow rule as df_sweep, over the task-contiguous layout of
// DfDev, with the memory latency of each task hidden: a wave holds the
// descriptors of its next 64 tasks in registers (one per lane) and loads
// task j + 1's rows, columns and values while it runs task j.  Summation
// order within a row is free here (a correction solve).

// Sum of v over each group of gw = 16, 32 or 64 lanes, in every lane of the
// group (gw = 64: uniform).  Rows of 16 by DPP lane permutes (quad
// swaps, half-row and row mirrors), then one 32-lane swizzle or the four
// row sums.
__device__ __forceinline__ float group_sum(float v, uint32_t gw)
{
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xf, 0xf, false));   // quad [1,0,3,2]
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4E, 0xf, 0xf, false));   // quad [2,3,0,1]
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x141, 0xf, 0xf, false));  // row half mirror
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x140, 0xf, 0xf, false));  // row mirror
    if (gw == 32) v += __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x401F));        // lane ^ 16
    if (gw == 64)
        v = (__int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 15)) +
             __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 31))) +
            (__int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 47)) +
             __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63)));
    return v;
}

// The sweep vector as the dataflow sweeps see it: get(c), a finished row's
// value; in(r), row r's right-hand side; put(r, v), row r's result.
// XWhole: the whole vector in one place (LDS or global memory).
struct XWhole {
    float *x;
    __device__ __forceinline__ float get(uint32_t c) const { return x[c]; }
    __device__ __forceinline__ float in(uint32_t r) const { return x[r]; }
    __device__ __forceinline__ void put(uint32_t r, float v) const { x[r] = v; }
};

// XSplit: the vector in global memory, except that rows r >= r0 put their
// results into LDS (xl[r - r0]) and are read from there: the LDS the ready
// bits leave free holds the last rows' values.  The caller copies xl back
// once the sweep is done.
struct XSplit {
    float *x, *xl;
    uint32_t r0;
    __device__ __forceinline__ float get(uint32_t c) const { return c >= r0 ? xl[c - r0] : x[c]; }
    __device__ __forceinline__ float in(uint32_t r) const { return x[r]; }
    __device__ __forceinline__ void put(uint32_t r, float v) const
    {
        if (r >= r0) xl[r - r0] = v;
        else x[r] = v;
    }
};

struct DfTask {
    uint32_t base, p0, cnt, m, lg;   // lg: log2 lanes per row of a wave task
    bool wave;
    uint32_t row, len;
    uint32_t idx[kDfK];
    float val[kDfK];
    float dval;
};

__device__ __forceinline__ void df2_issue(const DfDev &D, const float *__restrict__ sv, uint4 d, uint32_t lane,
                                          DfTask &t)
{
    t.base = d.x;
    t.p0 = d.y;
    t.wave = (d.z & kLongTask) != 0;
    t.cnt = d.z & 0xffffffu;
    t.lg = (d.z >> 24) & 0x7fu;
    t.m = d.w;
    t.dval = 1.f;
    // All loads unconditional (the tables are padded past their last task
    // and row), so they need no branches; what a lane does not own is
    // masked after loading.  Entries past a row's length are never read.
    if (!t.wave) {
        const bool on = lane < t.cnt;
        const uint2 rl = D.rl[t.p0 + lane];
        t.row = on ? rl.x : 0u;
        t.len = on ? rl.y : 0u;
#pragma unroll
        for (uint32_t e = 0; e < kDfK; ++e) {
            t.idx[e] = D.idx[t.base + e * kWave + lane];
            t.val[e] = sv[t.base + e * kWave + lane];
        }
        const float dv = sv[D.nslots + t.p0 + lane];
        if (D.div) t.dval = dv;
    } else {
        // every lane's first slot exists (groups and 64-lane rows are padded)
        const uint32_t g = lane >> t.lg;
        const bool on = g < t.cnt;
        const uint2 rl = D.rl[t.p0 + g];
        t.row = on ? rl.x : 0u;
        t.len = on ? rl.y : 0u;
        t.idx[0] = D.idx[t.base + lane];
        t.val[0] = sv[t.base + lane];
        const float dv = sv[D.nslots + t.p0 + g];
        if (D.div) t.dval = dv;
    }
}

// A task of lane rows some longer than kDfK entries: each lane walks its
// row kDfK entries at a time (the slots of entry e at base + 64 e + lane),
// consuming ready entries in the row's order as the short-row path does.
// Polls are bounded (kSpinCap); a row still pending then completes with NaN.
template <typename XV>
__device__ __forceinline__ void df2_lane_long(const DfDev &D, const float *__restrict__ sv, const XV &x,
                                              uint32_t *flags, const DfTask &t, uint32_t lane)
{
    bool pending = lane < t.cnt;
    float acc = pending ? x.in(t.row) : 0.f;
    // the row's entries kDfK at a time: [c0, c0 + kDfK) in idx / val (the
    // first chunk read again from the slots the task's prefetch brought in:
    // taking it from t would push the short-row path's arrays to scratch)
    // The next chunk is loaded while this one is consumed (nidx / nval),
    // when the lane's row has entries there (so within the task's slots).
    uint32_t idx[kDfK], nidx[kDfK];
    float val[kDfK], nval[kDfK];
#pragma unroll
    for (uint32_t e = 0; e < kDfK; ++e) {
        idx[e] = D.idx[t.base + e * kWave + lane];
        val[e] = sv[t.base + e * kWave + lane];
    }
    if (pending && kDfK < t.len)
#pragma unroll
        for (uint32_t e = 0; e < kDfK; ++e) {
            nidx[e] = D.idx[t.base + (kDfK + e) * kWave + lane];
            nval[e] = sv[t.base + (kDfK + e) * kWave + lane];
        }
    uint32_t ci = 0, c0 = 0;   // entries consumed, first entry of the chunk
    for (uint32_t spins = 0; __ballot(pending) && spins < kSpinCap; ++spins) {
        if (pending) {
            if (ci == c0 + kDfK) {
                // chunk consumed, row not done: the next kDfK entries
                // (loaded ahead), and the chunk after them requested
                c0 += kDfK;
#pragma unroll
                for (uint32_t e = 0; e < kDfK; ++e) {
                    idx[e] = nidx[e];
                    val[e] = nval[e];
                }
                if (c0 + kDfK < t.len)
#pragma unroll
                    for (uint32_t e = 0; e < kDfK; ++e) {
                        const uint32_t sl = t.base + (c0 + kDfK + e) * kWave + lane;
                        nidx[e] = D.idx[sl];
                        nval[e] = sv[sl];
                    }
            }
            // every entry's ready bit and x value read without branches
            // (independent LDS reads, one wait each), then the ready
            // prefix consumed in order
            uint32_t fw[kDfK];
#pragma unroll
            for (uint32_t e = 0; e < kDfK; ++e)
                fw[e] = __hip_atomic_load(flags + (idx[e] >> 5), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            uint32_t mask = 0;
#pragma unroll
            for (uint32_t e = 0; e < kDfK; ++e)
                mask |= (((fw[e] >> (idx[e] & 31u)) & 1u) & (uint32_t)(c0 + e >= ci && c0 + e < t.len)) << e;
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            float xv[kDfK];
#pragma unroll
            for (uint32_t e = 0; e < kDfK; ++e) xv[e] = x.get(idx[e]);
#pragma unroll
            for (uint32_t e = 0; e < kDfK; ++e)
                if (c0 + e == ci && ((mask >> e) & 1u)) {
                    acc = acc - val[e] * xv[e];
                    ++ci;
                }
            if (ci == t.len) {
                if (D.div) acc = acc / t.dval;
                x.put(t.row, acc);
                df_set(flags, t.row);
                pending = false;
            }
        }
    }
    if (pending) {   // gave up: the row completes with NaN
        x.put(t.row, df_nan());
        df_set(flags, t.row);
    }
}

// Waits spin on the LDS ready bits without s_sleep: measured on the n=20k
// batch, the wake-up latency of a sleeping wave costs more (7.08 vs
// 6.92 ms per batch) than the LDS reads of the spinning ones.
// kExact: every row's products subtracted one at a time in the row's order
// (group rows through the wave's LDS slice `lds`, by the group's first
// lane) — the arithmetic of Symbolic_Matrix::solve, for glu_solve.
template <bool kExact = false, typename XV = XWhole>
__device__ __forceinline__ void df2_run(const DfDev &D, const float *__restrict__ sv, const XV &x, uint32_t *flags,
                                        const DfTask &t, uint32_t lane, float *lds = nullptr)
{
    if (!t.wave) {
        if (t.m > kDfK) {   // uniform: a task of long lane rows
            df2_lane_long(D, sv, x, flags, t, lane);
            return;
        }
        bool pending = lane < t.cnt;
        float acc = pending ? x.in(t.row) : 0.f;
        uint32_t ci = 0;
        while (__ballot(pending)) {   // unbounded: the plan's tables are checked on the host
            if (pending) {
                // every entry's ready bit and x value read without branches
                // (independent LDS reads, one wait each), then the ready
                // prefix consumed in order
                uint32_t fw[kDfK];
#pragma unroll
                for (uint32_t e = 0; e < kDfK; ++e)
                    fw[e] = __hip_atomic_load(flags + (t.idx[e] >> 5), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                uint32_t mask = 0;
#pragma unroll
                for (uint32_t e = 0; e < kDfK; ++e)
                    mask |= (((fw[e] >> (t.idx[e] & 31u)) & 1u) & (uint32_t)(e >= ci && e < t.len)) << e;
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                float xv[kDfK];
#pragma unroll
                for (uint32_t e = 0; e < kDfK; ++e) xv[e] = x.get(t.idx[e]);
#pragma unroll
                for (uint32_t e = 0; e < kDfK; ++e)
                    if (e == ci && ((mask >> e) & 1u)) {
                        acc = acc - t.val[e] * xv[e];
                        ++ci;
                    }
                if (ci == t.len) {
                    if (D.div) acc = acc / t.dval;
                    x.put(t.row, acc);
                    df_set(flags, t.row);
                    pending = false;
                }
            }
        }
        return;
    }
    // groups of gw lanes, a row each (no row depends on another of the
    // task); only a 64-lane row may need more than one pass, and then loads
    // four passes at a time
    const uint32_t gw = 1u << t.lg, j = lane & (gw - 1u);
    if (kExact) {
        // The products of a pass in parallel, then subtracted in the row's
        // order along the group's lanes: lane j takes lane j - 1's running
        // value (a DPP shift within each row of 16 lanes; across rows, the
        // previous row's lane 15) and subtracts its own product.  Lanes past
        // the row's end subtract +0, which changes no value.  The group's
        // last lane ends with acc - p[0] - p[1] - ... - p[len - 1].
        const bool on = (lane >> t.lg) < t.cnt;
        float carry = on ? x.in(t.row) : 0.f;   // the running value entering a pass, in lane 0 of the group
        float a = 0.f;
        for (uint32_t c0 = 0; c0 < t.len || c0 == 0; c0 += kWave) {   // a pass of gw (64: of 64) entries
            uint32_t col = t.idx[0];
            float v = t.val[0];
            if (c0 > 0) {   // gw == 64 here
                col = c0 + j < t.len ? D.idx[t.base + c0 + j] : 0u;
                v = c0 + j < t.len ? sv[t.base + c0 + j] : 0.f;
            }
            float pr = 0.f;
            if (c0 + j < t.len) {
                while (!df_ready(flags, col)) {}
                pr = v * x.get(col);
            }
            a = j == 0 ? carry - pr : pr;
            for (uint32_t r = 0; r < gw / 16u; ++r) {
                if (r > 0) {   // row_bcast:15
                    const float b = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(a), 0x142, 0xf, 0xf, false));
                    if (j == 16u * r) a = b - pr;
                }
#pragma unroll
                for (uint32_t k = 1; k < 16u; ++k) {   // row_shr:1
                    const float prev = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(a), 0x111, 0xf, 0xf, false));
                    if (j == 16u * r + k) a = prev - pr;
                }
            }
            if (gw == kWave) carry = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(a), kWave - 1));
        }
        if (on && j == gw - 1u) {
            if (D.div) a = a / t.dval;
            x.put(t.row, a);
            df_set(flags, t.row);
        }
        return;
    }
    float sum = 0.f;
    if (j < t.len) {
        while (!df_ready(flags, t.idx[0])) {}
        sum = t.val[0] * x.get(t.idx[0]);
    }
    for (uint32_t c0 = kWave; c0 < t.len; c0 += 4 * kWave) {   // gw == 64 here
        uint32_t col[4];
        float v[4];
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u)
            if (c0 + u * kWave + j < t.len) {
                col[u] = D.idx[t.base + c0 + u * kWave + j];
                v[u] = sv[t.base + c0 + u * kWave + j];
            }
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u)
            if (c0 + u * kWave + j < t.len) {
                while (!df_ready(flags, col[u])) {}
                sum += v[u] * x.get(col[u]);
            }
    }
    sum = group_sum(sum, gw);
    if (j == 0 && (lane >> t.lg) < t.cnt) {
        float acc = x.in(t.row) - sum;
        if (D.div) acc = acc / t.dval;
        x.put(t.row, acc);
        df_set(flags, t.row);
    }
}

// All tasks of sweep D for the calling workgroup; sv: this sweep's values.
// The caller has cleared every ready bit.  A wave runs tasks w, w + nw, ...
// Descriptors are scalar loads (uniform per wave; they count against the
// scalar/LDS wait counter, so waiting for one never waits for the vector
// loads in flight), fetched two tasks ahead; a task's rows, columns and
// values are vector loads issued one task ahead.
template <bool kExact = false, typename XV = XWhole>
__device__ __forceinline__ void df2_sweep(const DfDev &D, const float *__restrict__ sv, const XV &x, uint32_t *flags,
                                          float *lds = nullptr)
{
    const uint32_t lane = threadIdx.x & (kWave - 1), nw = blockDim.x / kWave;
    const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);   // uniform: scalar loads below
    const uint32_t mine = D.ntask > w ? (D.ntask - w + nw - 1) / nw : 0;   // tasks w, w + nw, ...
    if (mine == 0) return;
    // an empty task (nothing to load or run) past the wave's last one
    // through the constant address space: the index is uniform, so this is
    // a scalar load, which never waits for the vector loads in flight
    const __attribute__((address_space(4))) uint32_t *tasks =
        (const __attribute__((address_space(4))) uint32_t *)(uintptr_t)D.task;
    auto desc = [&](uint32_t j) {
        const uint32_t q = 4 * (w + (j < mine ? j : mine - 1) * nw);
        const uint4 d = make_uint4(tasks[q], tasks[q + 1], tasks[q + 2], tasks[q + 3]);
        return j < mine ? d : make_uint4(0u, 0u, 0u, 0u);
    };
    uint4 dn = desc(0), dn2 = desc(1);
    DfTask cur, nxt;
    df2_issue(D, sv, dn, lane, nxt);
    for (uint32_t j = 0; j < mine; ++j) {
        cur = nxt;
        // this task's loads (issued a task ago) complete before the next
        // task's are issued: the task's waits inside its polling loop then
        // cover nothing in flight, where a wait there would also wait for
        // the next task's loads
        __builtin_amdgcn_s_waitcnt(0xF70);   // vmcnt(0); expcnt, lgkmcnt untouched
        dn = dn2;
        dn2 = desc(j + 2);
        df2_issue(D, sv, dn, lane, nxt);
        df2_run<kExact, XV>(D, sv, x, flags, cur, lane, lds);
    }
}

// LDS placement of the batch pipeline's sweep state (dynamic LDS).
enum : int {
    kSweepLevels = 0,   // level-synchronous sweeps, correction vector in global memory
    kSweepFlags = 1,    // dataflow sweeps, ready bits in LDS, vector in global memory
    kSweepLds = 2,      // dataflow sweeps, ready bits and vector in LDS
};

// One workgroup runs one matrix: value refresh, GESP value, factorization
// and the copy of the factors for the sweeps (wg_factor), then, in a second
// kernel with its own register budget, refinement sweeps until eta <= tol
// or maxit (wg_refine).  Md is the matrix's descriptor in device memory:
// each phase takes the (globalized) tables it needs from it, so the
// descriptor is not held in registers as a whole.
//
// wg_values: steps 1-2 (the matrix's values, |b|_inf with x = 0 and
// r64 = b, the GESP value).
__device__ __forceinline__ void wg_values(const MatDesc *Md, const MatIO &io, int perturb, double *sh, double &bnorm,
                                          float &pert)
{
    const uint32_t n = Md->n, nnz = Md->nnz;
    const uint32_t tid = threadIdx.x, nt = blockDim.x;

    // 1. values: zero the pattern, then scatter the scaled entries of A
    //    (lu[a_dst[k]] = float((ax[k] * rs) * cs), the refactor order);
    //    |b|_inf (|A|_inf comes with the refinement's first product with A)
    for (uint32_t p = tid; p < nnz; p += nt) io.lu[p] = 0.f;
    __syncthreads();
    {
        RefineDev R = Md->r;
        globalize(R);
        constexpr int U = 4;
        const uint32_t na = R.nnz_a;
        for (uint32_t k0 = tid; k0 < na; k0 += U * nt) {
            double a[U], s1[U], s2[U];
            uint32_t dst[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t k = k0 + u * nt;
                if (k < na) {
                    a[u] = io.ax[k];
                    dst[u] = R.a_dst[k];
                    s1[u] = R.a_rs[k];
                    s2[u] = R.a_cs[k];
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (k0 + u * nt < na) io.lu[dst[u]] = (float)((a[u] * s1[u]) * s2[u]);
        }
        double bmax = 0.0;
        if (io.b)   // null: refactor only (k_batch_factor)
            for (uint32_t c = tid; c < n; c += nt) {
                bmax = nmax(bmax, fabs(io.b[c]));
                io.x[c] = 0.0;
                io.r64[c] = io.b[c];
            }
        bnorm = wg_max(bmax, sh);
    }

    // 2. GESP value from the freshly placed values
    pert = 0.f;
    if (perturb) {
        const uint32_t *col_ptr = glb(Md->f.col_ptr);
        double cm = 0.0;
        for (uint32_t c = tid; c < n; c += nt) {
            float s = 0.f;
            for (uint32_t p = col_ptr[c]; p < col_ptr[c + 1]; ++p) s += fabsf(io.lu[p]);
            if (s == s) cm = fmax(cm, (double)s);
        }
        pert = kPertFactor * (float)wg_max(cm, sh);
    }
}

// Steps 1-4 (l0 = 0), or, after the interleaved launches of steps [0, l0)
// (k_factor_grid_il; GESP value in io.norms[0]), |b|_inf into io.norms[1]
// with x = 0 and r = b, steps l0.. of step 3 and step 4.
template <bool kCopy>
__device__ __forceinline__ void wg_factor(const MatDesc *Md, const MatIO &io, int perturb, double *sh, float *wlds,
                                          int l0)
{
    const uint32_t tid = threadIdx.x, nt = blockDim.x;
    double bnorm = 0.0;
    float pert;
    if (l0 == 0) {
        wg_values(Md, io, perturb, sh, bnorm, pert);
    } else {
        pert = perturb ? (float)io.norms[0] : 0.f;
        double bmax = 0.0;
        if (io.b) {   // null: refactor only (k_batch_factor)
            for (uint32_t c = tid; c < Md->n; c += nt) {
                const double v = io.b[c];
                bmax = nmax(bmax, fabs(v));
                io.x[c] = 0.0;
                io.r64[c] = v;
            }
            bnorm = wg_max(bmax, sh);
        }
    }

    // 3. factorization, level by level; LDS windows with their values in LDS
    //    (the sweeps' LDS is free until step 5)
    {
        FactorDev F = Md->f;
        globalize(F);
        factor_all_wg(F, io.lu, g_lds_vec, pert, perturb, wlds, l0);
    }

    // 4. the factors copied into the sweeps' task-contiguous layout, built
    //    in LDS (free until step 5) one range of slots at a time: the
    //    range's sources read in increasing CSC position (few lines per
    //    wave) and placed in LDS, then the range written out whole
    //    (contiguous, whole-line stores); padding slots get 0
    if (kCopy) {
        const uint2 *cpy = glb(Md->s.df_cpy);
        const uint32_t *cptr = glb(Md->s.df_cpy_ptr);
        const uint32_t nval = Md->s.df_nval, range = Md->s.df_cpy_range;
        float *buf = g_lds_vec;
        for (uint32_t base = 0, r = 0; base < nval; base += range, ++r) {
            const uint32_t len = min(range, nval - base);
            for (uint32_t i = tid; i < len; i += nt) buf[i] = 0.f;
            __syncthreads();
            const uint32_t e1 = cptr[r + 1];
            constexpr int U = 8;
            for (uint32_t k0 = cptr[r] + tid; k0 < e1; k0 += U * nt) {
                uint2 e[U];
                float v[U];
                // loads without branches (past the end: the last entry
                // again), so that the eight table loads, then the eight
                // value loads, issue together
#pragma unroll
                for (int u = 0; u < U; ++u) e[u] = cpy[min(k0 + u * nt, e1 - 1)];
#pragma unroll
                for (int u = 0; u < U; ++u) v[u] = io.lu[e[u].x];
#pragma unroll
                for (int u = 0; u < U; ++u)
                    if (k0 + u * nt < e1) buf[e[u].y] = v[u];
            }
            __syncthreads();
            for (uint32_t i = tid; i < len; i += nt) io.lucsr[base + i] = buf[i];
            __syncthreads();
        }
    }
    if (tid == 0 && io.norms && (l0 == 0 || io.b)) io.norms[1] = bnorm;
}

// Workgroup sum in a fixed order (deterministic): each wave's butterfly,
// then the waves' sums in wave order.  Ends with a barrier.
__device__ double wg_sum(double v, double *sh)
{
    v = wave_sum64(v);
    const int nw = blockDim.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) sh[threadIdx.x / kWave] = v;
    __syncthreads();
    double r = sh[0];
    for (int i = 1; i < nw; ++i) r += sh[i];
    __syncthreads();
    return r;
}

// xs = (LU)^-1 applied to v in the factors' frame: xs[j] = v[col_perm[j]] *
// col_scale_perm[j] rounded to FP32, then the FP32 L and U sweeps in place
// (Symbolic_Matrix::solve's permutation and scaling, symbolic.cc:215-239).
// The caller maps xs back with row_perm_inv / row_scale.
template <int kMode, bool kT = false>
__device__ __forceinline__ void wg_precond(const MatDesc *Md, const MatIO &io, const double *v, float *xs,
                                           uint32_t *flags)
{
    const uint32_t n = Md->n;
    const uint32_t tid = threadIdx.x, nt = blockDim.x;
    const uint32_t nwords = (n + 31) / 32;
    if (!kT) {
        const uint32_t *cp = glb(Md->s.col_perm);
        const double *csp = glb(Md->s.col_scale_perm);
        for (uint32_t j = tid; j < n; j += nt) xs[j] = (float)(v[cp[j]] * csp[j]);
    } else {   // A' x = b: rows by row_perm, scaled by row_scale
        const uint32_t *rp = glb(Md->s.row_perm);
        const double *rs = glb(Md->s.row_scale);
        for (uint32_t j = tid; j < n; j += nt) {
            const uint32_t pj = rp[j];
            xs[j] = (float)(v[pj] * rs[pj]);
        }
    }
    if (kMode == kSweepLevels) {
        __syncthreads();
        TriDev lo = Md->s.lo, up = Md->s.up;
        globalize(lo);
        globalize(up);
        wg_tri_refine(lo, io.lu, xs);
        wg_tri_refine(up, io.lu, xs);
    } else {
        for (uint32_t i = tid; i < nwords; i += nt) flags[i] = 0u;
        __syncthreads();
        {
            DfDev lo = kT ? Md->s.tup_df : Md->s.lo_df;
            globalize(lo);
            df2_sweep(lo, io.lucsr + lo.voff, XWhole{xs}, flags);
        }
        __syncthreads();
        for (uint32_t i = tid; i < nwords; i += nt) flags[i] = 0u;
        __syncthreads();
        {
            DfDev up = kT ? Md->s.tlo_df : Md->s.up_df;
            globalize(up);
            df2_sweep(up, io.lucsr + up.voff, XWhole{xs}, flags);
        }
        __syncthreads();
    }
}

// r = b - A x in FP64 (into io.r64); returns the backward error
// eta = |r|_inf / (|A|_inf |x|_inf + |b|_inf) to every thread.
// Rows of A (kT: of A', the caller's CSC columns) for one workgroup.
template <bool kT, typename F>
__device__ __forceinline__ void wg_rows(const RefineDev &R, const double *__restrict__ ax,
                                        const double *__restrict__ x, uint32_t n, F f)
{
    if (!kT) wg_rows_of_a(R, ax, x, n, f);
    else for_rows_of_a(R, ax, x, n, 1, threadIdx.x, blockDim.x, f);
}

// anorm < 0: |A|_inf is taken from this product's row sums (and returned
// in anorm).
template <bool kT = false>
__device__ __forceinline__ double wg_residual(const MatDesc *Md, const MatIO &io, double &anorm, double bnorm,
                                              double *sh)
{
    const uint32_t n = Md->n;
    const uint32_t tid = threadIdx.x, nt = blockDim.x;
    double rmax = 0.0, xmax = 0.0, amax = 0.0;
    RefineDev R = Md->r;
    globalize(R);
    wg_rows<kT>(R, io.ax, io.x, n, [&](uint32_t i, double dot, double asum) {
        const double ri = io.b[i] - dot;
        io.r64[i] = ri;
        rmax = nmax(rmax, fabs(ri));
        amax = nmax(amax, asum);
    });
    for (uint32_t i = tid; i < n; i += nt) xmax = nmax(xmax, fabs(io.x[i]));
    if (anorm < 0.0) anorm = wg_max(amax, sh);
    const double rn = wg_max(rmax, sh);
    const double xn = wg_max(xmax, sh);
    return rn / (anorm * xn + bnorm);
}

// LDS of the Arnoldi cycle (doubles): the (kMaxKrylov + 1) x kMaxKrylov
// Hessenberg matrix H (rotated in place into R), the least-squares
// coefficients y, the rotated right-hand side g, the Givens cosines and
// sines, and the change of y since the previous step.
constexpr int kKrylovY = (kMaxKrylov + 1) * kMaxKrylov;
constexpr int kKrylovG = kKrylovY + kMaxKrylov;
constexpr int kKrylovC = kKrylovG + kMaxKrylov + 1;
constexpr int kKrylovS = kKrylovC + kMaxKrylov;
constexpr int kKrylovD = kKrylovS + kMaxKrylov;
constexpr int kKrylovLds = kKrylovD + kMaxKrylov;

// Refinement of one matrix by one workgroup until eta <= tol or maxit
// preconditioner solves, from x = 0, r = b.  M = (LU)^-1 is the FP32 sweep
// pair; every other quantity is FP64.
//  * Md->krylov = 0 (or no workspace): plain refinement, x += M r, then
//    r = b - A x and eta from it.
//  * Md->krylov = m > 0: restarted flexible GMRES(m), right-preconditioned
//    by M.  A cycle starts from the current residual, v_0 = r / |r|_2; step
//    j stores z_j = M v_j and w_j = A z_j, orthogonalizes w_j against V by
//    modified Gram-Schmidt into v_{j+1}, extends the Givens QR of the
//    Hessenberg matrix and solves min |beta e_1 - H y|_2; x and r move by
//    the change of y: x += sum dy_c z_c, r -= sum dy_c w_c (r stays b - A x
//    up to FP64 rounding, as every w_c is an FP64 product A z_c).  eta of
//    the updated r is checked after every solve: a matrix stops as soon as
//    it meets tol with room for the rounding drift between the updated and
//    the true residual (then the true eta <= tol); near tol, at maxit or on
//    NaN the true residual b - A x is formed and its eta decides and is
//    reported.  A restart cycle starts from the updated residual.
// Every workgroup-wide sum has a fixed order: results do not depend on
// timing.  xs is the FP32 sweep vector; flags the sweeps' ready bits.
template <int kMode, bool kT = false>
__device__ __forceinline__ void wg_refine(const MatDesc *Md, const MatIO &io, double tol, int maxit, double *sh,
                                          double *hs, float *xs, uint32_t *flags)
{
    const uint32_t n = Md->n;
    const uint32_t tid = threadIdx.x, nt = blockDim.x;
    const double bnorm = io.norms[1];
    double anorm = -1.0;   // |A|_inf: from the first product with A
    const int m = io.kry ? (int)Md->krylov : 0;
    // back from the factors' frame: x[i] += xs[bp[i]] * bs[kT ? bp[i] : i]
    const uint32_t *bp = glb(kT ? Md->s.col_perm_inv : Md->s.row_perm_inv);
    const double *bs = glb(kT ? Md->s.col_scale_perm : Md->s.row_scale);
    int it = 0;
    double eta = 0.0;
    if (m == 0) {
        for (;;) {
            wg_precond<kMode, kT>(Md, io, io.r64, xs, flags);
            ++it;
            for (uint32_t i = tid; i < n; i += nt) {
                const uint32_t pi = bp[i];
                io.x[i] = io.x[i] + (double)xs[pi] * bs[kT ? pi : i];
            }
            __syncthreads();
            eta = wg_residual<kT>(Md, io, anorm, bnorm, sh);
            if (!(eta > tol) || it >= maxit) break;   // also stops on NaN
        }
    } else {
        double *const V = io.kry;                            // m + 1 vectors
        double *const Z = V + (size_t)(m + 1) * n;           // m
        double *const W = Z + (size_t)m * n;                 // m
        double *const H = hs, *const y = hs + kKrylovY, *const g = hs + kKrylovG;
        double *const cs = hs + kKrylovC, *const sn = hs + kKrylovS, *const dy = hs + kKrylovD;
        bool done = false;
        while (!done) {
            double d = 0.0;
            for (uint32_t i = tid; i < n; i += nt) d = fma(io.r64[i], io.r64[i], d);
            const double beta = sqrt(wg_sum(d, sh));
            if (!(beta > 0.0)) {                  // r = 0 (x exact) or NaN
                eta = beta == 0.0 ? 0.0 : beta;
                break;
            }
            for (uint32_t i = tid; i < n; i += nt) V[i] = io.r64[i] / beta;
            if (tid == 0) {
                g[0] = beta;
                for (int c = 0; c < kMaxKrylov; ++c) y[c] = 0.0;
            }
            __syncthreads();
            for (int j = 0; j < m; ++j) {
                double *zj = Z + (size_t)j * n, *wj = W + (size_t)j * n, *vn = V + (size_t)(j + 1) * n;
                wg_precond<kMode, kT>(Md, io, V + (size_t)j * n, xs, flags);
                ++it;
                for (uint32_t i = tid; i < n; i += nt) {
                    const uint32_t pi = bp[i];
                    zj[i] = (double)xs[pi] * bs[kT ? pi : i];
                }
                __syncthreads();
                {
                    RefineDev R = Md->r;
                    globalize(R);
                    double amax = 0.0;
                    wg_rows<kT>(R, io.ax, zj, n, [&](uint32_t i, double dot, double asum) {
                        wj[i] = dot;
                        vn[i] = dot;
                        amax = nmax(amax, asum);
                    });
                    if (anorm < 0.0) anorm = wg_max(amax, sh);
                }
                __syncthreads();
                for (int q = 0; q <= j; ++q) {
                    const double *vq = V + (size_t)q * n;
                    double dq = 0.0;
                    for (uint32_t i = tid; i < n; i += nt) dq = fma(vn[i], vq[i], dq);
                    const double h = wg_sum(dq, sh);
                    for (uint32_t i = tid; i < n; i += nt) vn[i] = fma(-h, vq[i], vn[i]);
                    if (tid == 0) H[q * kMaxKrylov + j] = h;
                }
                double dn = 0.0;
                for (uint32_t i = tid; i < n; i += nt) dn = fma(vn[i], vn[i], dn);
                const double hn = sqrt(wg_sum(dn, sh));
                if (hn > 0.0)
                    for (uint32_t i = tid; i < n; i += nt) vn[i] = vn[i] / hn;
                if (tid == 0) {
                    // previous rotations on column j, a new one zeroing hn
                    for (int c = 0; c < j; ++c) {
                        const double t1 = H[c * kMaxKrylov + j], t2 = H[(c + 1) * kMaxKrylov + j];
                        H[c * kMaxKrylov + j] = cs[c] * t1 + sn[c] * t2;
                        H[(c + 1) * kMaxKrylov + j] = cs[c] * t2 - sn[c] * t1;
                    }
                    const double a = H[j * kMaxKrylov + j];
                    const double rr = hypot(a, hn);
                    cs[j] = rr > 0.0 ? a / rr : 1.0;
                    sn[j] = rr > 0.0 ? hn / rr : 0.0;
                    H[j * kMaxKrylov + j] = rr;
                    g[j + 1] = -sn[j] * g[j];
                    g[j] = cs[j] * g[j];
                    for (int c = j; c >= 0; --c) {
                        double t = g[c];
                        for (int cc = c + 1; cc <= j; ++cc) t -= H[c * kMaxKrylov + cc] * dy[cc];
                        dy[c] = t / H[c * kMaxKrylov + c];   // the new y, for now
                    }
                    for (int c = 0; c <= j; ++c) {
                        const double yn = dy[c];
                        dy[c] = yn - y[c];
                        y[c] = yn;
                    }
                }
                __syncthreads();
                double rmax = 0.0, xmax = 0.0;
                for (uint32_t i = tid; i < n; i += nt) {
                    double xi = io.x[i], ri = io.r64[i];
                    for (int c = 0; c <= j; ++c) {
                        const double dc = dy[c];
                        xi = fma(dc, Z[(size_t)c * n + i], xi);
                        ri = fma(-dc, W[(size_t)c * n + i], ri);
                    }
                    io.x[i] = xi;
                    io.r64[i] = ri;
                    rmax = nmax(rmax, fabs(ri));
                    xmax = nmax(xmax, fabs(xi));
                }
                const double rn = wg_max(rmax, sh);
                const double xn = wg_max(xmax, sh);
                eta = rn / (anorm * xn + bnorm);
                // The updated r drifts from b - A x by rounding only: each
                // step adds at most ~(2 k + m + 3) u (|A||x| + |b|) per entry
                // (k: most entries in a row of A), so the true eta is within
                // `drift` of this one.  Accept outright when that cannot
                // cross tol; otherwise (or at maxit, or NaN) decide on the
                // true residual.
                const double drift = (double)it * (2.0 * (double)Md->a_kmax + 8.0) * 0x1p-52;
                if (eta + drift <= tol) {
                    done = true;
                    break;
                }
                if (!(eta > tol) || it >= maxit) {
                    eta = wg_residual<kT>(Md, io, anorm, bnorm, sh);
                    if (!(eta > tol) || it >= maxit) done = true;
                    break;   // not done: restart from the true residual
                }
                if (!(hn > 0.0)) break;   // breakdown short of tol: restart
            }
        }
    }
    if (tid == 0) {
        *io.iters = it;
        *io.eta = eta;
    }
}

// Dynamic LDS of the batch kernels: ready bits (ceil(n/32) words), then,
// for kSweepLds, the n-float correction vector.
template <int kMode>
__device__ __forceinline__ void batch_lds(uint32_t n, float *t32, uint32_t *&flags, float *&xs)
{
    flags = kMode == kSweepLevels ? nullptr : reinterpret_cast<uint32_t *>(g_lds_vec);
    xs = kMode == kSweepLds ? g_lds_vec + (n + 31) / 32 : t32;
}

// Same pattern: matrix b = blockIdx.x uses desc[0] and contiguous buffers.
__device__ __forceinline__ MatIO batch_io(const MatDesc *desc, const double *ax, const double *b, double *x,
                                          float *lu, float *lucsr, float *t32, double *r64, double *norms,
                                          double *kry, int32_t *iters, double *eta)
{
    const uint32_t n = desc->n, nnz = desc->nnz, nnz_a = desc->r.nnz_a, nval = desc->s.df_nval;
    const uint64_t m = blockIdx.x;
    MatIO io;
    io.ax = ax + m * nnz_a;
    io.b = b + m * n;
    io.x = x + m * n;
    io.lu = lu + m * nnz;
    io.t32 = t32 + m * n;
    io.r64 = r64 + m * n;
    io.lucsr = lucsr ? lucsr + m * nval : nullptr;
    io.norms = norms + 2 * m;
    io.kry = kry ? kry + m * (3 * desc->krylov + 1) * n : nullptr;
    io.iters = iters + m;
    io.eta = eta + m;
    globalize(io);
    return io;
}

// Dynamic LDS: the largest LDS window's values.  l0 > 0: steps [0, l0) ran
// as interleaved grid launches (k_factor_grid_il).
template <bool kCopy>
__global__ void __launch_bounds__(1024) k_batch_same_factor(const MatDesc *__restrict__ desc,
                                                            const double *__restrict__ ax, const double *__restrict__ b,
                                                            double *__restrict__ x, float *__restrict__ lu,
                                                            float *__restrict__ lucsr, float *__restrict__ t32,
                                                            double *__restrict__ r64, double *__restrict__ norms,
                                                            double *__restrict__ kry, int32_t *__restrict__ iters,
                                                            double *__restrict__ eta, int perturb, int l0)
{
    __shared__ double sh[1024 / kWave];
    __shared__ __attribute__((aligned(16))) float wl[1024 / kWave][kWave];
    const MatIO io = batch_io(desc, ax, b, x, lu, lucsr, t32, r64, norms, kry, iters, eta);
    wg_factor<kCopy>(desc, io, perturb, sh, wl[threadIdx.x / kWave], l0);
}

// Refresh + factor only (glu_batch_factor): matrix blockIdx.x of a
// same-pattern batch, values ax + b nnz_a into lu + b nnz.  Dynamic LDS: the
// largest LDS window's values.  l0 > 0: steps [0, l0) ran as interleaved
// grid launches (k_factor_grid_il; GESP values in norms).
__global__ void __launch_bounds__(1024) k_batch_factor(const MatDesc *__restrict__ desc,
                                                       const double *__restrict__ ax, float *__restrict__ lu,
                                                       double *__restrict__ norms, int perturb, int l0)
{
    __shared__ double sh[1024 / kWave];
    __shared__ __attribute__((aligned(16))) float wl[1024 / kWave][kWave];
    MatIO io{};
    io.ax = glb(ax + (uint64_t)blockIdx.x * desc->r.nnz_a);
    io.lu = glb(lu + (uint64_t)blockIdx.x * desc->nnz);
    if (l0 > 0) io.norms = glb(norms + 2 * (uint64_t)blockIdx.x);
    wg_factor<false>(desc, io, perturb, sh, wl[threadIdx.x / kWave], l0);
}

// Dynamic LDS: the sweeps' ready bits (and correction vector).
template <int kMode>
__global__ void __launch_bounds__(1024) k_batch_same_refine(const MatDesc *__restrict__ desc,
                                                            const double *__restrict__ ax, const double *__restrict__ b,
                                                            double *__restrict__ x, float *__restrict__ lu,
                                                            float *__restrict__ lucsr, float *__restrict__ t32,
                                                            double *__restrict__ r64, double *__restrict__ norms,
                                                            double *__restrict__ kry, int32_t *__restrict__ iters,
                                                            double *__restrict__ eta, double tol, int maxit)
{
    __shared__ double sh[1024 / kWave];
    __shared__ double hs[kKrylovLds];
    const MatIO io = batch_io(desc, ax, b, x, lu, lucsr, t32, r64, norms, kry, iters, eta);
    uint32_t *flags;
    float *xs;
    batch_lds<kMode>(desc->n, io.t32, flags, xs);
    wg_refine<kMode>(desc, io, tol, maxit, sh, hs, xs, flags);
}

// k_batch_same_factor<true> and k_batch_same_refine<kMode> in one
// workgroup per matrix (each matrix's refinement starts as soon as its
// factors are copied).  Dynamic LDS: the larger of the two.
template <int kMode>
__global__ void __launch_bounds__(1024) k_batch_same_fused(const MatDesc *__restrict__ desc,
                                                           const double *__restrict__ ax, const double *__restrict__ b,
                                                           double *__restrict__ x, float *__restrict__ lu,
                                                           float *__restrict__ lucsr, float *__restrict__ t32,
                                                           double *__restrict__ r64, double *__restrict__ norms,
                                                           double *__restrict__ kry, int32_t *__restrict__ iters,
                                                           double *__restrict__ eta, int perturb, int l0, double tol,
                                                           int maxit)
{
    __shared__ double sh[1024 / kWave];
    __shared__ __attribute__((aligned(16))) float wl[1024 / kWave][kWave];
    static_assert(sizeof(double) * kKrylovLds <= sizeof(wl), "Arnoldi scalars exceed the wave slices");
    const MatIO io = batch_io(desc, ax, b, x, lu, lucsr, t32, r64, norms, kry, iters, eta);
    wg_factor<true>(desc, io, perturb, sh, wl[threadIdx.x / kWave], l0);
    __syncthreads();
    // the refinement's Arnoldi scalars take the factorization's wave slices
    // (static LDS no larger than k_batch_same_factor's)
    double *hs = reinterpret_cast<double *>(&wl[0][0]);
    uint32_t *flags;
    float *xs;
    batch_lds<kMode>(desc->n, io.t32, flags, xs);
    wg_refine<kMode>(desc, io, tol, maxit, sh, hs, xs, flags);
}

// One matrix's factors gathered into the sweeps' layout (grid, a thread per
// slot): lucsr[q] = lu[df_src[q]], 0 for padding slots.
template <bool kT>
__global__ void __launch_bounds__(256) k_copy_df(const MatDesc *__restrict__ desc, const float *__restrict__ lu,
                                                 float *__restrict__ lucsr)
{
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (kT ? desc->s.tdf_nval : desc->s.df_nval)) return;
    const uint32_t p = glb(kT ? desc->s.tdf_src : desc->s.df_src)[q];
    lucsr[q] = p != ~0u ? lu[p] : 0.f;
}

// Refinement of one factored matrix (glu_refine; kT: A' x = b): x = 0,
// r = b, |b|_inf, then wg_refine in this one workgroup.  Dynamic LDS as
// k_batch_same_refine.
template <int kMode, bool kT>
__global__ void __launch_bounds__(1024) k_refine_one(const MatDesc *__restrict__ desc,
                                                     const double *__restrict__ ax, const double *__restrict__ b,
                                                     double *__restrict__ x, const float *__restrict__ lu,
                                                     float *__restrict__ lucsr, float *__restrict__ t32,
                                                     double *__restrict__ r64, double *__restrict__ norms,
                                                     double *__restrict__ kry, int32_t *__restrict__ iters,
                                                     double *__restrict__ eta, double tol, int maxit)
{
    __shared__ double sh[1024 / kWave];
    __shared__ double hs[kKrylovLds];
    const MatIO io = batch_io(desc, ax, b, x, const_cast<float *>(lu), lucsr, t32, r64, norms, kry, iters, eta);
    const uint32_t n = desc->n;
    double bmax = 0.0;
    for (uint32_t c = threadIdx.x; c < n; c += blockDim.x) {
        bmax = nmax(bmax, fabs(io.b[c]));
        io.x[c] = 0.0;
        io.r64[c] = io.b[c];
    }
    const double bnorm = wg_max(bmax, sh);
    if (threadIdx.x == 0) io.norms[1] = bnorm;
    __syncthreads();
    uint32_t *flags;
    float *xs;
    batch_lds<kMode>(n, io.t32, flags, xs);
    wg_refine<kMode, kT>(desc, io, tol, maxit, sh, hs, xs, flags);
}

// The preconditioner's two sweeps alone, in one workgroup, over a
// correction vector t32 in global memory that grid kernels scaled before
// and read back after (glu_refine for orders whose vector does not fit in
// LDS: the FP64 vector work and the products with A run across the grid).
// kT: A' x = b (U' then L').  Dynamic LDS: the ready bits, ceil(n/32) words,
// then the results of rows r0..n-1 (XSplit), copied to t32 after each sweep.
template <bool kT, bool kSplit>
__global__ void __launch_bounds__(1024) k_sweeps_df(const MatDesc *__restrict__ desc,
                                                    const float *__restrict__ lucsr, float *__restrict__ t32,
                                                    uint32_t r0)
{
    const uint32_t n = desc->n, nwords = (n + 31) / 32, tid = threadIdx.x, nt = blockDim.x;
    uint32_t *flags = reinterpret_cast<uint32_t *>(g_lds_vec);
    for (int sweep = 0; sweep < 2; ++sweep) {
        for (uint32_t i = tid; i < nwords; i += nt) flags[i] = 0u;
        __syncthreads();
        DfDev d = sweep == 0 ? (kT ? desc->s.tup_df : desc->s.lo_df) : (kT ? desc->s.tlo_df : desc->s.up_df);
        globalize(d);
        if (kSplit) {
            const XSplit xv{t32, g_lds_vec + nwords, r0};
            df2_sweep(d, lucsr + d.voff, xv, flags);
            __syncthreads();
            for (uint32_t r = r0 + tid; r < n; r += nt) t32[r] = xv.xl[r - r0];   // the LDS rows back
        } else {
            df2_sweep(d, lucsr + d.voff, XWhole{t32}, flags);
        }
        __syncthreads();
    }
}

// glu_solve in one workgroup (A x = b): xs = float(b[col_perm] *
// col_scale_perm), the exact dataflow L and U sweeps over the sweep-layout
// copy of the factors, x = float(xs[row_perm_inv] * row_scale) — the
// arithmetic of Symbolic_Matrix::solve (src/symbolic.cc:215-247).  Dynamic
// LDS as k_batch_same_refine.
template <int kMode, bool kSplit = false>
__global__ void __launch_bounds__(1024) k_solve_one(const MatDesc *__restrict__ desc,
                                                    const float *__restrict__ lucsr, const float *__restrict__ b,
                                                    float *__restrict__ x, float *__restrict__ t32, uint32_t r0)
{
    __shared__ __attribute__((aligned(16))) float wl[1024 / kWave][kWave];
    const uint32_t n = desc->n, tid = threadIdx.x, nt = blockDim.x;
    uint32_t *flags;
    float *xs;
    batch_lds<kMode>(n, t32, flags, xs);
    {
        const uint32_t *cp = glb(desc->s.col_perm);
        const double *csp = glb(desc->s.col_scale_perm);
        for (uint32_t j = tid; j < n; j += nt) xs[j] = (float)((double)b[cp[j]] * csp[j]);
    }
    const uint32_t nwords = (n + 31) / 32;
    for (int sweep = 0; sweep < 2; ++sweep) {
        for (uint32_t i = tid; i < nwords; i += nt) flags[i] = 0u;
        __syncthreads();
        DfDev d = sweep == 0 ? desc->s.lo_df : desc->s.up_df;
        globalize(d);
        if (kMode == kSweepFlags && kSplit) {   // the vector in global memory but its last rows (r0..n-1) in LDS
            const XSplit xv{xs, g_lds_vec + nwords, r0};
            df2_sweep<true>(d, glb(lucsr) + d.voff, xv, flags, wl[tid / kWave]);
            __syncthreads();
            for (uint32_t r = r0 + tid; r < n; r += nt) xs[r] = xv.xl[r - r0];
        } else {
            df2_sweep<true>(d, glb(lucsr) + d.voff, XWhole{xs}, flags, wl[tid / kWave]);
        }
        __syncthreads();
    }
    const uint32_t *rpi = glb(desc->s.row_perm_inv);
    const double *rs = glb(desc->s.row_scale);
    for (uint32_t j = tid; j < n; j += nt) x[j] = (float)((double)xs[rpi[j]] * rs[j]);
}

// Distinct patterns: workgroup g runs matrix order[g] with its own tables.
template <bool kCopy>
__global__ void __launch_bounds__(1024) k_batch_group_factor(const MatDesc *__restrict__ descs,
                                                             const MatIO *__restrict__ ios,
                                                             const uint32_t *__restrict__ order, int perturb)
{
    __shared__ double sh[1024 / kWave];
    __shared__ __attribute__((aligned(16))) float wl[1024 / kWave][kWave];
    const uint32_t m = order[blockIdx.x];
    MatIO io = ios[m];
    globalize(io);
    wg_factor<kCopy>(descs + m, io, perturb, sh, wl[threadIdx.x / kWave], 0);
}

template <int kMode>
__global__ void __launch_bounds__(1024) k_batch_group_refine(const MatDesc *__restrict__ descs,
                                                             const MatIO *__restrict__ ios,
                                                             const uint32_t *__restrict__ order, double tol, int maxit)
{
    __shared__ double sh[1024 / kWave];
    __shared__ double hs[kKrylovLds];
    const uint32_t m = order[blockIdx.x];
    MatIO io = ios[m];
    globalize(io);
    uint32_t *flags;
    float *xs;
    batch_lds<kMode>(descs[m].n, io.t32, flags, xs);
    wg_refine<kMode>(descs + m, io, tol, maxit, sh, hs, xs, flags);
}

// Distinct patterns whose largest order leaves the vector out of LDS: each
// workgroup keeps its own matrix's vector in LDS when it fits the launch's
// dynamic LDS (lds_bytes), else only the ready bits.
__global__ void __launch_bounds__(1024) k_batch_group_refine_mixed(const MatDesc *__restrict__ descs,
                                                                   const MatIO *__restrict__ ios,
                                                                   const uint32_t *__restrict__ order, double tol,
                                                                   int maxit, uint32_t lds_bytes)
{
    __shared__ double sh[1024 / kWave];
    __shared__ double hs[kKrylovLds];
    const uint32_t m = order[blockIdx.x];
    MatIO io = ios[m];
    globalize(io);
    const uint32_t n = descs[m].n;
    uint32_t *flags;
    float *xs;
    if (4u * ((n + 31) / 32) + 4u * n <= lds_bytes) {
        batch_lds<kSweepLds>(n, io.t32, flags, xs);
        wg_refine<kSweepLds>(descs + m, io, tol, maxit, sh, hs, xs, flags);
    } else {
        batch_lds<kSweepFlags>(n, io.t32, flags, xs);
        wg_refine<kSweepFlags>(descs + m, io, tol, maxit, sh, hs, xs, flags);
    }
}

// ---------------------------------------------------------------------------
// ABFT column checksums (src/symbolic.cc:163-197), a batch of value sets
// ---------------------------------------------------------------------------
//
// One thread per column and matrix (blockIdx.y), each sum sequential in the
// reference's order, FP32, products rounded before they are added.

// CCA[i] = sum of column i of the values (before the factorization).
__global__ void __launch_bounds__(256) k_abft_cca(const uint32_t *__restrict__ col_ptr,
                                                  const float *__restrict__ val, float *__restrict__ cca,
                                                  uint32_t n, uint32_t nnz)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float *v = val + (size_t)blockIdx.y * nnz;
    float s = 0.f;
    for (uint32_t j = col_ptr[i]; j < col_ptr[i + 1]; ++j) s += v[j];
    cca[(size_t)blockIdx.y * n + i] = s;
}

// CCL[r] = 1 + sum of L(:, r); first_bad / max_diff armed for the check.
__global__ void __launch_bounds__(256) k_abft_ccl(const uint32_t *__restrict__ col_ptr,
                                                  const uint32_t *__restrict__ diag, const float *__restrict__ lu,
                                                  float *__restrict__ ccl, int32_t *__restrict__ first_bad,
                                                  unsigned int *__restrict__ max_bits, uint32_t n, uint32_t nnz)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        first_bad[blockIdx.y] = INT32_MAX;
        max_bits[blockIdx.y] = 0u;
    }
    if (i >= n) return;
    const float *v = lu + (size_t)blockIdx.y * nnz;
    float s = 1.f;
    for (uint32_t j = diag[i] + 1; j < col_ptr[i + 1]; ++j) s += v[j];
    ccl[(size_t)blockIdx.y * n + i] = s;
}

// CCA_ABFT[i] = sum over U(:, i) (diagonal included) of CCL[row] * U; a
// column fails when |CCA_ABFT - CCA| > tol (the reference's test: a NaN
// difference does not fail it, but max_diff becomes NaN).
__global__ void __launch_bounds__(256) k_abft_check(const uint32_t *__restrict__ col_ptr,
                                                    const uint32_t *__restrict__ row_idx,
                                                    const uint32_t *__restrict__ diag, const float *__restrict__ lu,
                                                    const float *__restrict__ ccl, const float *__restrict__ cca,
                                                    float tol, int32_t *__restrict__ first_bad,
                                                    unsigned int *__restrict__ max_bits, uint32_t n, uint32_t nnz)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float *v = lu + (size_t)blockIdx.y * nnz;
    const float *cl = ccl + (size_t)blockIdx.y * n;
    float s = 0.f;
    for (uint32_t j = col_ptr[i]; j <= diag[i]; ++j) s += cl[row_idx[j]] * v[j];
    const float d = fabsf(s - cca[(size_t)blockIdx.y * n + i]);
    if (d > tol) atomicMin(first_bad + blockIdx.y, (int32_t)i);
    // non-negative floats order like their bit patterns; NaN sorts above inf
    atomicMax(max_bits + blockIdx.y, d == d ? __float_as_uint(d) : 0x7fc00000u);
}

// first_bad: INT32_MAX (no failing column) -> -1.
__global__ void k_abft_finish(int32_t *__restrict__ first_bad, int batch)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < batch && first_bad[b] == INT32_MAX) first_bad[b] = -1;
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

// lanes: the lanes the level needs (column tasks and long merges 64 each)
void factor_grid(const FactorDev &F, float *val, const float *d_pert, int lvl, uint32_t lanes,
                 int perturb, int threads, hipStream_t st)
{
    k_factor_grid<<<blocks_for(lanes, (unsigned)threads), threads, 0, st>>>(F, val, d_pert, lvl, perturb);
}

void factor_fused(const FactorDev &F, float *val, const float *d_pert, int l0, int l1, int perturb,
                  int threads, hipStream_t st)
{
    k_factor_fused<<<1, threads, 0, st>>>(F, val, d_pert, l0, l1, perturb);
}

// ndiv: the division items (all steps).
void factor_divide(const FactorDev &F, float *val, const float *d_pert, uint32_t ndiv, int perturb, hipStream_t st)
{
    k_factor_divide<<<std::max(1u, std::min(blocks_for(ndiv, 256), 2048u)), 256, 0, st>>>(F, val, d_pert, perturb);
}

// lanes: long rows x 64 + short rows of the level
void tri_grid(const TriDev &T, const float *lu, float *x, int lvl, uint32_t lanes, hipStream_t st)
{
    k_tri_grid<<<blocks_for(lanes, 256), 256, 0, st>>>(T, lu, x, lvl);
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

void refresh(const RefineDev &R, const double *ax, float *lu, uint32_t nnz_lu, int batch, hipStream_t st)
{
    k_refresh<<<blocks_for((uint64_t)nnz_lu * batch, 256), 256, 0, st>>>(R, ax, lu, nnz_lu, batch);
}

// lanes: one wave per long row of A (or A') plus one lane per short row
void residual(const RefineDev &R, const double *ax, const double *b, const double *x, double *r, uint32_t n,
              int transpose, int with_anorm, unsigned long long *norms, uint32_t lanes, hipStream_t st)
{
    k_residual<<<blocks_for(lanes, 256), 256, 0, st>>>(R, ax, b, x, r, n, transpose, with_anorm, norms);
}

// Dynamic LDS bytes of a batch kernel in sweep mode `mode` for order n.
size_t batch_lds_bytes(int mode, uint32_t n)
{
    const size_t words = (n + 31) / 32;
    return mode == kSweepLevels ? 0 : 4 * words + (mode == kSweepLds ? 4 * (size_t)n : 0);
}

// Sweep mode kSweepLds: one launch, factor then refine in each workgroup
// (dynamic LDS: the larger of fact_lds and ref_lds; 4.67 -> 4.65 ms per
// step of 256 n=20k matrices against the two launches, the refinements no
// longer waiting for the slowest factorization).  Other modes: two
// launches, factor (dynamic LDS fact_lds: the largest window's values),
// then refine (dynamic LDS ref_lds).
void batch_same(const MatDesc *desc, int batch, const double *ax, const double *b, double *x, float *lu,
                float *lucsr, float *t32, double *r64, double *norms, double *kry, int32_t *iters, double *eta,
                double tol, int maxit, int perturb, int threads, int mode, size_t fact_lds, size_t ref_lds,
                int l0, hipStream_t st)
{
    if (mode == kSweepLds) {
        k_batch_same_fused<kSweepLds><<<batch, threads, std::max(fact_lds, ref_lds), st>>>(
            desc, ax, b, x, lu, lucsr, t32, r64, norms, kry, iters, eta, perturb, l0, tol, maxit);
        return;
    }
    if (mode == kSweepLevels)
        k_batch_same_factor<false><<<batch, threads, fact_lds, st>>>(desc, ax, b, x, lu, lucsr, t32, r64, norms,
                                                                      kry, iters, eta, perturb, l0);
    else
        k_batch_same_factor<true><<<batch, threads, fact_lds, st>>>(desc, ax, b, x, lu, lucsr, t32, r64, norms,
                                                                     kry, iters, eta, perturb, l0);
    if (mode == kSweepFlags)
        k_batch_same_refine<kSweepFlags><<<batch, threads, ref_lds, st>>>(desc, ax, b, x, lu, lucsr, t32, r64,
                                                                           norms, kry, iters, eta, tol, maxit);
    else
        k_batch_same_refine<kSweepLevels><<<batch, threads, ref_lds, st>>>(desc, ax, b, x, lu, lucsr, t32, r64,
                                                                            norms, kry, iters, eta, tol, maxit);
}

// glu_refine in one workgroup (orders whose sweep vector fits in LDS): the
// factors into the sweeps' layout, then the refinement.
void refine_one(const MatDesc *desc, uint32_t nval, const double *ax, const double *b, double *x,
                const float *lu, float *lucsr, float *t32, double *r64, double *norms, double *kry, int32_t *iters,
                double *eta, double tol, int maxit, int threads, size_t ref_lds, int transpose, hipStream_t st)
{
    const unsigned grid = blocks_for(nval, 256);
    if (transpose) {
        k_copy_df<true><<<grid, 256, 0, st>>>(desc, lu, lucsr);
        k_refine_one<kSweepLds, true><<<1, threads, ref_lds, st>>>(desc, ax, b, x, lu, lucsr, t32, r64, norms, kry,
                                                                   iters, eta, tol, maxit);
    } else {
        k_copy_df<false><<<grid, 256, 0, st>>>(desc, lu, lucsr);
        k_refine_one<kSweepLds, false><<<1, threads, ref_lds, st>>>(desc, ax, b, x, lu, lucsr, t32, r64, norms, kry,
                                                                    iters, eta, tol, maxit);
    }
}

// The factors into the sweeps' layout (A or A' tables), grid.
void copy_df(const MatDesc *desc, uint32_t nval, const float *lu, float *lucsr, int transpose, hipStream_t st)
{
    if (transpose) k_copy_df<true><<<blocks_for(nval, 256), 256, 0, st>>>(desc, lu, lucsr);
    else k_copy_df<false><<<blocks_for(nval, 256), 256, 0, st>>>(desc, lu, lucsr);
}

// Both sweeps of the preconditioner in one workgroup over t32 (global).
void sweeps_df(const MatDesc *desc, const float *lucsr, float *t32, uint32_t r0, int threads, size_t lds,
               int transpose, hipStream_t st)
{
    // r0 = UINT32_MAX: the whole vector in global memory
    if (r0 != UINT32_MAX) {
        if (transpose) k_sweeps_df<true, true><<<1, threads, lds, st>>>(desc, lucsr, t32, r0);
        else k_sweeps_df<false, true><<<1, threads, lds, st>>>(desc, lucsr, t32, r0);
    } else {
        if (transpose) k_sweeps_df<true, false><<<1, threads, lds, st>>>(desc, lucsr, t32, r0);
        else k_sweeps_df<false, false><<<1, threads, lds, st>>>(desc, lucsr, t32, r0);
    }
}

// glu_solve in one workgroup: the factors into the sweep layout, then the
// exact sweeps (mode: kSweepLds or kSweepFlags).
void solve_one(const MatDesc *desc, uint32_t nval, const float *lu, float *lucsr, const float *b, float *x,
               float *t32, uint32_t r0, int threads, int mode, size_t lds, hipStream_t st)
{
    k_copy_df<false><<<blocks_for(nval, 256), 256, 0, st>>>(desc, lu, lucsr);
    if (mode == kSweepLds) k_solve_one<kSweepLds><<<1, threads, lds, st>>>(desc, lucsr, b, x, t32, r0);
    else if (r0 != UINT32_MAX) k_solve_one<kSweepFlags, true><<<1, threads, lds, st>>>(desc, lucsr, b, x, t32, r0);
    else k_solve_one<kSweepFlags><<<1, threads, lds, st>>>(desc, lucsr, b, x, t32, r0);
}

void abft_cca(const FactorDev &F, const float *val, float *cca, int batch, hipStream_t st)
{
    k_abft_cca<<<dim3(blocks_for(F.n, 256), batch), 256, 0, st>>>(F.col_ptr, val, cca, F.n, F.nnz);
}

// max_diff: `batch` floats, filled from the bit maxima in place.
void abft_check(const FactorDev &F, const uint32_t *row_idx, const uint32_t *diag, const float *lu,
                const float *cca, float *ccl, float tol, int32_t *first_bad, float *max_diff, int batch,
                hipStream_t st)
{
    unsigned int *bits = reinterpret_cast<unsigned int *>(max_diff);
    const dim3 grid(blocks_for(F.n, 256), batch);
    k_abft_ccl<<<grid, 256, 0, st>>>(F.col_ptr, diag, lu, ccl, first_bad, bits, F.n, F.nnz);
    k_abft_check<<<grid, 256, 0, st>>>(F.col_ptr, row_idx, diag, lu, ccl, cca, tol, first_bad, bits, F.n, F.nnz);
    k_abft_finish<<<blocks_for(batch, 256), 256, 0, st>>>(first_bad, batch);
}

void batch_factor(const MatDesc *desc, int batch, const double *ax, float *lu, double *norms, int perturb,
                  int threads, size_t fact_lds, int l0, hipStream_t st)
{
    k_batch_factor<<<batch, threads, fact_lds, st>>>(desc, ax, lu, norms, perturb, l0);
}

// The interleaved copy of a same-pattern batch (groups of kIl matrices,
// S: groups kIl nnz floats): zero, the scaled entries of A and, with
// perturb, each matrix's GESP value into pertd[2 m] (bits: batch words,
// zero on entry and on return).
void il_values(const RefineDev &R, const FactorDev &F, const double *ax, float *S, unsigned int *bits,
               double *pertd, uint32_t nnz, uint32_t n, int batch, int perturb, hipStream_t st)
{
    const uint32_t groups = (uint32_t)((batch + kIl - 1) / kIl);
    const size_t count4 = (size_t)groups * nnz * kIl / 4;
    // whole groups: only the fill-in lines need zeros (the scatter writes
    // every other line whole; 4.64 -> 4.62 ms per step of 256 n=20k
    // matrices); with a partial group, whose unused lanes the scatter does
    // not write, every line is zeroed
    if (batch % kIl == 0) {
        if (R.nfill)
            k_il_zero_fill<<<dim3(blocks_for(4 * R.nfill, 256), groups), 256, 0, st>>>(
                R.lu_fill, R.nfill, reinterpret_cast<float4 *>(S), nnz);
    } else {
        k_il_zero<<<(unsigned)std::min<size_t>(blocks_for(count4, 256), 8192), 256, 0, st>>>(
            reinterpret_cast<float4 *>(S), count4);
    }
    k_il_scatter<<<dim3(blocks_for(R.nnz_a, 64), groups), 256, 0, st>>>(R, ax, S, nnz, (uint32_t)batch);
    if (perturb) {
        k_il_pert<<<dim3(blocks_for((uint64_t)n * kIl, 256), groups), 256, 0, st>>>(F.col_ptr, S, n, nnz,
                                                                                      (uint32_t)batch, bits);
        k_il_pert_finish<<<blocks_for(batch, 256), 256, 0, st>>>(bits, pertd, (uint32_t)batch);
    }
}

// waves: the step's waves for one group (long merges, then the column-task
// items and short merges kIlItems a wave).
void factor_grid_il(const FactorDev &F, float *S, const double *pertd, int lvl, uint32_t waves, int batch,
                    int perturb, hipStream_t st)
{
    const uint32_t groups = (uint32_t)((batch + kIl - 1) / kIl);
    k_factor_grid_il<<<blocks_for(waves, 256 / kWave) * groups, 256, 0, st>>>(F, S, pertd, lvl, perturb, groups,
                                                                               (uint32_t)batch);
}

void il_to_rows(const float *S, float *lu, uint32_t nnz, int batch, hipStream_t st)
{
    const uint32_t groups = (uint32_t)((batch + kIl - 1) / kIl);
    k_il_to_rows<<<dim3(blocks_for(nnz, 256), groups), 256, 0, st>>>(S, lu, nnz, (uint32_t)batch);
}

uint32_t il_items_per_wave() { return kIlItems * kIlPer; }
uint32_t il_group() { return kIl; }

void batch_group(const MatDesc *descs, const void *ios, const uint32_t *order, int count, double tol,
                 int maxit, int perturb, int threads, int mode, size_t fact_lds, size_t ref_lds, hipStream_t st)
{
    const MatIO *io = static_cast<const MatIO *>(ios);
    if (mode == kSweepLevels)
        k_batch_group_factor<false><<<count, threads, fact_lds, st>>>(descs, io, order, perturb);
    else
        k_batch_group_factor<true><<<count, threads, fact_lds, st>>>(descs, io, order, perturb);
    if (mode == kSweepLds)
        k_batch_group_refine<kSweepLds><<<count, threads, ref_lds, st>>>(descs, io, order, tol, maxit);
    else if (mode == kSweepFlags)   // per matrix: vector in LDS where it fits ref_lds
        k_batch_group_refine_mixed<<<count, threads, ref_lds, st>>>(descs, io, order, tol, maxit,
                                                                   (uint32_t)ref_lds);
    else
        k_batch_group_refine<kSweepLevels><<<count, threads, ref_lds, st>>>(descs, io, order, tol, maxit);
}

void factor_window(const FactorDev &F, int w, uint32_t nslot, float *val, const float *d_pert, int perturb,
                   int threads, hipStream_t st)
{
    k_factor_window<<<1, threads, 4 * (size_t)nslot, st>>>(F, val, d_pert, w, perturb);
}

// Dataflow sweep of fused levels [l0, l1) (tasks [k0, k1)) in one workgroup.
void tri_df(const TriDev &T, const float *lu, float *x, int l0, int l1, uint32_t k0, uint32_t k1, uint32_t n,
            int threads, hipStream_t st)
{
    k_tri_df<<<1, threads, 4 * (size_t)((n + 31) / 32), st>>>(T, lu, x, l0, l1, k0, k1, n);
}

// Raises the dynamic-LDS cap of the batch, dataflow and window kernels to the
// device's per-workgroup limit (less their static LDS) and returns that
// limit in bytes (0 if no device).  Launches then size their own dynamic LDS
// up to it.
size_t enable_lds()
{
    int dev = 0, lim = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    // once per device (the attributes stay set for the process); plans may
    // be created from several host threads
    static std::mutex mu;
    static int done_dev = -1;
    static size_t done_avail = 0;
    std::lock_guard<std::mutex> lock(mu);
    if (dev == done_dev) return done_avail;
    if (hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) return 0;
    const void *fns[] = {
        reinterpret_cast<const void *>(&k_batch_same_factor<true>),
        reinterpret_cast<const void *>(&k_batch_same_factor<false>),
        reinterpret_cast<const void *>(&k_batch_factor),
        reinterpret_cast<const void *>(&k_solve_one<kSweepLds>),
        reinterpret_cast<const void *>(&k_solve_one<kSweepFlags>),
        reinterpret_cast<const void *>(&k_refine_one<kSweepLds, false>),
        reinterpret_cast<const void *>(&k_refine_one<kSweepLds, true>),
        reinterpret_cast<const void *>(&k_sweeps_df<false, false>),
        reinterpret_cast<const void *>(&k_sweeps_df<true, false>),
        reinterpret_cast<const void *>(&k_sweeps_df<false, true>),
        reinterpret_cast<const void *>(&k_sweeps_df<true, true>),
        reinterpret_cast<const void *>(&k_solve_one<kSweepFlags, true>),
        reinterpret_cast<const void *>(&k_batch_same_fused<kSweepLds>),
        reinterpret_cast<const void *>(&k_batch_same_refine<kSweepFlags>),
        reinterpret_cast<const void *>(&k_batch_same_refine<kSweepLevels>),
        reinterpret_cast<const void *>(&k_batch_group_factor<true>),
        reinterpret_cast<const void *>(&k_batch_group_factor<false>),
        reinterpret_cast<const void *>(&k_batch_group_refine<kSweepLds>),
        reinterpret_cast<const void *>(&k_batch_group_refine<kSweepFlags>),
        reinterpret_cast<const void *>(&k_batch_group_refine<kSweepLevels>),
        reinterpret_cast<const void *>(&k_batch_group_refine_mixed),
        reinterpret_cast<const void *>(&k_tri_df),
        reinterpret_cast<const void *>(&k_factor_window),
    };
    // the limit less the largest static LDS among them (their wave slices
    // and reduction arrays, ~4 KiB); 8 KiB if the attributes are unavailable
    size_t stat = 0;
    for (const void *f : fns) {
        hipFuncAttributes a;
        if (hipFuncGetAttributes(&a, f) != hipSuccess) {
            stat = 8192;
            break;
        }
        stat = std::max(stat, (size_t)a.sharedSizeBytes);
    }
    const size_t avail = (size_t)lim > stat ? (size_t)lim - stat : 0;
    for (const void *f : fns) (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)avail);
    done_dev = dev;
    done_avail = avail;
    return avail;
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
