// Device-resident tables of one sparsity pattern (POD structs of device
// pointers, passed by value to kernels or stored in device memory for the
// fused multi-matrix launch).
//
// Layout in HBM (all 32-bit indices, one copy per pattern):
//   factor: per level, 16-byte update items (target, L, U and pivot
//           positions of one scalar update), merge entries with their source
//           lists, and 8-byte division items.
//   solves: row lists per level for the four triangular sweeps.
//   refine: CSR of the original A (row pointers, columns, source index into
//           the caller's CSC values), per-LU-position source index and the
//           frozen MC64 scale vectors.
// Values (FP32 L\U, FP64 A, FP64 x/b) are separate buffers indexed by
// matrix; a batch lays matrices out back to back (matrix-major).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace glu {

// An entry updated by several pivot columns of one level: one lane owns it
// and applies the updates in increasing column order (deterministic).
struct MergeSrc {
    uint32_t lpos;     // CSC position of L(i,k) (not yet divided)
    uint32_t upos;     // CSC position of U(k,j)
    uint32_t dpos;     // CSC position of the pivot U(k,k)
    uint32_t pad;
};

constexpr uint32_t kLongTask = 0x80000000u;
// Dataflow sweeps: rows with at most this many entries are taken a lane
// each, with all their entries prefetched; longer rows a wave each.
constexpr uint32_t kDfLaneRow = 16;

// Rows / merge entries longer than this are processed by a whole wave.
constexpr uint32_t kLongRow = 64;
constexpr int kMaxKrylov = 4;     // largest restart length of the batch refinement
constexpr uint32_t kLongMerge = 32;

// Flag of a division item: the first L entry of its column, whose lane also
// stores the GESP pivot.
constexpr uint32_t kFirstBit = 0x80000000u;

// Per-level items of the right-looking factorization:
//   column task  (a wave) A(i,j) += -(A(i,k) / A(k,k) * A(k,j)) for targets
//                exactly one column of the level updates;
//   merge        all updates of an entry several columns of the level
//                update, in increasing column order (a lane, or a wave);
//   divide       A(i,k) /= A(k,k) for the columns of the previous level
//                {lpos, dpos | kFirstBit} (deferred so a level needs one
//                barrier).

// One LDS window: steps [l0, l1) with the nslot values they touch staged in
// LDS.  Its items (cln / ctgt / cup / ment / msrc / dvd, indexed like the
// factor's tables) address LDS slots instead of CSC positions; pos[slot] is
// the slot's CSC position.
struct WinDev {
    int32_t l0, l1;
    uint32_t nslot, pad;
    const uint32_t *pos;
    const uint2 *cln;
    const uint32_t *ctgt;
    const uint32_t *cup;
    const uint32_t *ment;
    const MergeSrc *msrc;
    const uint4 *mrec;
    const uint2 *dvd;
};

struct FactorDev {
    const uint32_t *col_ptr;   // sym_c_ptr [n+1] (GESP column sums)
    // Column tasks, by level (tasks ctask_ptr[l] .. ctask_ptr[l+1]).  Lane
    // `lane` of task t owns one L entry of a pivot column k of the level and
    // up to four consecutive entries of U_k: cln[64 t + lane] = {L(i,k)
    // position, pivot position} ({0, 0} for an unused lane); its e-th update
    // reads U(k,j) at cup[256 t + 64 e + lane] and targets ctgt[same index]
    // (~0u: no update here — an unused slot, or a merge entry does it).
    const uint32_t *ctask_ptr; // [num_lev + 2]
    const uint2 *cln;
    const uint32_t *ctgt;
    const uint32_t *cup;
    const uint32_t *ment;      // merge entries (target CSC positions), by level,
                               // entries with > kLongMerge sources first
    const uint32_t *ment_ptr;  // [num_lev + 2]
    const uint32_t *ment_long; // long merge entries at the head of each level [num_lev + 1]
    const uint32_t *msrc_ptr;  // per merge entry [entries + 1]
    const MergeSrc *msrc;      // sources in increasing column order
    // per merge entry e, its first two sources in one place (every merge
    // entry has at least two): mrec[2e] = {target, L, U, pivot position of
    // source 0}, mrec[2e + 1] = {L, U, pivot position of source 1, source
    // count}; sources from the third on are read from msrc
    const uint4 *mrec;
    // division items {L position, pivot position | kFirstBit}: those of the
    // columns of each step before the batched kernels' last grid step, step
    // by step, then the rest; each step's, and the rest, in position order
    const uint2 *dvd;
    const uint32_t *dvd_ptr;   // [num_lev + 1]: items of the columns of steps < l
    int32_t num_lev;
    uint32_t n, nnz;
    // LDS windows: runs of steps that work on values staged in LDS
    const WinDev *win;         // [nwin], in step order
    int32_t nwin;
    uint32_t win_nslot;        // the largest window's values
};

// One level-scheduled triangular sweep, row-pull form:
//   x[r] = (x[r] - sum_{e in row r} V[vpos(e)] * x[idx[e]]) [/ V[dpos[r]]]
// with row r's entries [beg[r] + begoff, end[r]) taken in ascending order,
// or descending when desc != 0.
struct TriDev {
    const uint32_t *beg;
    const uint32_t *end;
    const uint32_t *idx;
    const uint32_t *vpos;      // nullptr: identity
    const uint32_t *dpos;      // nullptr: unit diagonal
    const uint32_t *rows;      // rows grouped by level, long rows first
    const uint32_t *lvl_ptr;   // [nlev + 1]
    const uint32_t *lvl_long;  // long rows at the head of each level [nlev]
    // Dataflow schedule: tasks in topological order, each a run of <= 64
    // short rows or one long row (start position | kLongTask); task k covers
    // positions [dtask[k] & ~kLongTask, dtask[k + 1] & ~kLongTask).  Tasks
    // never cross a fused-segment boundary; seg_task gives each fused
    // segment's first task.
    const uint32_t *dtask;     // [ntask + 1]
    uint32_t ntask;
    uint32_t begoff;
    int32_t desc;
    int32_t nlev;
    int32_t pad;
};

// A pipelined dataflow sweep over a per-matrix copy of the factor values
// laid out task by task (the batch refinement).  Task k = {base, p0, z, m}:
//  - z = rows (< 2^24), no kLongTask bit: rows one per lane: lane l takes
//    position p0 + l (l < rows), its entry e at slot base + 64 e + l (e < m,
//    m = the task's longest row; shorter rows leave padding slots); a row
//    longer than kDfK is walked kDfK entries at a time;
//  - z = kLongTask | log2(gw) << 24 | rows: rows of one level (so no row of
//    the task depends on another) taken gw = 16, 32 or 64 lanes each: row
//    g at position p0 + g, its entry j at slot base + gw g + j; a 64-lane
//    row (rows = 1) may have more than 64 entries, at base + j.
// Each position's {row, entries} is in rl; entries are stored in the order
// the row consumes them.  div: U sweeps, the pivot of position p is value
// nslots + p.
constexpr uint32_t kDfK = 6;
// Rows of kDfK < len <= kDfLaneMax entries are lane rows (walked kDfK
// entries at a time) in levels that hold at least kDfWideLevel of them.
constexpr uint32_t kDfLaneMax = 64;
constexpr uint32_t kDfWideLevel = 32;

struct DfDev {
    const uint4 *task;         // [ntask]
    const uint2 *rl;           // [n]
    const uint32_t *idx;       // column of each slot [nslots]
    uint32_t ntask, nslots;
    int32_t div;
    uint32_t voff;             // this sweep's first value in the per-matrix copy
};

struct SolveDev {
    TriDev lo, up;             // A x = b:   L then U
    TriDev tup, tlo;           // A' x = b:  U' then L'
    DfDev lo_df, up_df;        // L, U over the per-matrix value copy
    const uint32_t *df_src;    // value copy: CSC position of each slot (~0u: padding, holds 0) [df_nval]
    uint32_t df_nval;          // slots per matrix
    // the same copy staged in LDS, df_cpy_range slots at a time: per range
    // r, entries [df_cpy_ptr[r], df_cpy_ptr[r + 1]) of df_cpy, {CSC
    // position, slot - r df_cpy_range} in increasing position order
    const uint2 *df_cpy;
    const uint32_t *df_cpy_ptr;
    uint32_t df_cpy_range;
    // A' x = b over a second value copy (U' slots, U' pivots, L' slots),
    // the same kind of tables
    DfDev tup_df, tlo_df;
    const uint32_t *tdf_src;
    uint32_t tdf_nval;
    const uint32_t *row_perm, *row_perm_inv, *col_perm, *col_perm_inv;
    const double *row_scale, *col_scale_perm;
};

struct RefineDev {
    const uint32_t *a_rptr;    // CSR of A [n+1]
    const uint32_t *a_cidx;    // [nnz_a]
    const uint32_t *a_src;     // index into the caller's CSC values [nnz_a]
    const uint32_t *a_cptr;    // caller CSC column pointers [n+1]
    const uint32_t *a_ridx;    // caller CSC row indices [nnz_a]
    const uint32_t *lu_src;    // LU position -> caller value index, or ~0u [nnz_lu]
    const uint32_t *row_idx;   // sym_r_idx [nnz_lu]
    const uint32_t *lu_col;    // column of each LU position [nnz_lu]
    const uint32_t *a_rows;    // rows of A, rows with > kLongRow entries first [n]
    const uint32_t *a_cols;    // columns of A, long ones first [n]
    uint32_t a_nlong;          // number of long rows
    uint32_t a_nlong_t;        // number of long columns
    const double *rs_col;      // row_scale[row_perm[c]] per GLU column c [n]
    const double *csp;         // col_scale_perm [n]
    // per caller value k: its LU position and the two scales of its entry
    const uint32_t *a_dst;     // [nnz_a]
    const double *a_rs;        // rs_col of the entry's LU column [nnz_a]
    const double *a_cs;        // csp of the entry's LU row [nnz_a]
    const uint32_t *lu_fill;   // LU positions no caller value lands on (fill-in), ascending [nfill]
    uint32_t nnz_a, nfill;
};

struct MatDesc {
    FactorDev f;
    SolveDev s;
    RefineDev r;
    uint32_t n, nnz;
    uint32_t krylov;   // Arnoldi vectors per refinement cycle (0: plain refinement)
    uint32_t a_kmax;   // most entries in a row or a column of A (rounding bound of the refinement)
};

}  // namespace glu
