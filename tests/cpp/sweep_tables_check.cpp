// CPU check of the dataflow sweep tables (build_df in glu_plan.hip) on a
// pattern: every slot, row and pivot load the sweep kernel issues (task
// prefetch, chunk refills of long lane rows, wave rows) stays inside the
// tables; every dependency of a row is finished when the row runs in task
// order (lane rows of one task in lane order, wave rows of one task
// independent); every row runs once; and the emulated L and U sweeps equal
// the sequential row-order solve bit for bit.
//
//   sweep_tables_check <pattern file> <long lane rows: 0 | 1>
//
// The pattern file holds n, csr_r_ptr, csr_c_idx, csr_diag_ptr, each as a
// uint64 count followed by uint32 values.  Compiled together with the
// library's host objects (this file includes glu_plan.hip itself).
#include "glu_plan.hip"
#include <cstdio>
#include <random>

using namespace glu;

static std::vector<uint32_t> rd(FILE *f)
{
    uint64_t k;
    if (fread(&k, 8, 1, f) != 1) exit(2);
    std::vector<uint32_t> v(k);
    if (k && fread(v.data(), 4, k, f) != k) exit(2);
    return v;
}

static int emulate(const char *name, const DfHost &D, uint32_t n, const std::vector<float> &lu, bool div,
                   const std::vector<float> &b, std::vector<float> &x)
{
    const size_t nslots = D.idx.size();
    std::vector<uint8_t> done(n, 0);
    x = b;
    auto chk = [&](size_t s, size_t lim, const char *what) {
        if (s >= lim) {
            printf("%s: OOB %s %zu >= %zu\n", name, what, s, lim);
            exit(1);
        }
    };
    auto sval = [&](size_t s) { return D.vpos[s] == UINT32_MAX ? 0.f : lu[D.vpos[s]]; };
    size_t lane_tasks = 0, long_lane_tasks = 0, wave_tasks = 0, maxslot = 0;
    for (size_t k = 0; k < D.task.size(); ++k) {
        const uint4 d = D.task[k];
        const uint32_t base = d.x, p0 = d.y, z = d.z, m = d.w;
        if (!(z & kLongTask)) {
            ++lane_tasks;
            if (m > kDfK) ++long_lane_tasks;
            const uint32_t cnt = z;
            for (uint32_t lane = 0; lane < 64; ++lane) {
                chk(p0 + lane, D.rl.size(), "rl");
                for (uint32_t e = 0; e < kDfK; ++e) chk(base + e * 64 + lane, nslots, "slot prefetch");
                chk(p0 + lane, (size_t)n + 64, "pivot");
                if (lane >= cnt) continue;
                const uint32_t row = D.rl[p0 + lane].x, len = D.rl[p0 + lane].y;
                if (len > m) { printf("%s: len %u > m %u\n", name, len, m); return 1; }
                float acc = x[row];
                for (uint32_t c0 = 0; c0 < len; c0 += kDfK)
                    for (uint32_t e = 0; e < kDfK; ++e) {
                        const size_t s = base + (size_t)(c0 + e) * 64 + lane;
                        chk(s, nslots, "slot chunk");
                        maxslot = std::max(maxslot, s);
                        if (c0 + e >= len) continue;
                        const uint32_t dep = D.idx[s];
                        chk(dep, n, "dep");
                        if (!done[dep]) { printf("%s: task %zu lane %u row %u needs %u not done\n", name, k, lane, row, dep); return 1; }
                        acc = acc - sval(s) * x[dep];
                    }
                if (div) acc = acc / lu[D.pvpos[p0 + lane]];
                if (done[row]) { printf("%s: row %u twice\n", name, row); return 1; }
                x[row] = acc;
                done[row] = 1;
            }
        } else {
            ++wave_tasks;
            const uint32_t lg = (z >> 24) & 0x7f, cnt = z & 0xffffff, gw = 1u << lg;
            std::vector<uint32_t> mine;
            for (uint32_t g = 0; g < cnt; ++g) mine.push_back(D.rl[p0 + g].x);
            for (uint32_t lane = 0; lane < 64; ++lane) chk(base + lane, nslots, "wave slot0");
            for (uint32_t g = 0; g < cnt; ++g) {
                const uint32_t row = D.rl[p0 + g].x, len = D.rl[p0 + g].y;
                if (gw < 64 && len > gw) { printf("%s: group row %u len %u > gw %u\n", name, row, len, gw); return 1; }
                float acc = x[row];
                for (uint32_t j = 0; j < len; ++j) {
                    const size_t s = base + (size_t)gw * g + j;
                    chk(s, nslots, "wave slot");
                    const uint32_t dep = D.idx[s];
                    for (uint32_t r : mine)
                        if (r == dep) { printf("%s: wave task %zu row %u depends on its own task\n", name, k, row); return 1; }
                    if (!done[dep]) { printf("%s: wave task %zu row %u needs %u not done\n", name, k, row, dep); return 1; }
                    acc = acc - sval(s) * x[dep];
                }
                if (div) acc = acc / lu[D.pvpos[p0 + g]];
                if (done[row]) { printf("%s: row %u twice\n", name, row); return 1; }
                x[row] = acc;
                done[row] = 1;
            }
        }
    }
    for (uint32_t r = 0; r < n; ++r)
        if (!done[r]) { printf("%s: row %u never done\n", name, r); return 1; }
    printf("%s: tasks lane %zu (long %zu) wave %zu, slots %zu (max used %zu)\n", name, lane_tasks, long_lane_tasks,
           wave_tasks, nslots, maxslot);
    return 0;
}

int main(int argc, char **argv)
{
    if (argc < 3) return 2;
    const bool long_lanes = atoi(argv[2]) != 0;
    FILE *f = fopen(argv[1], "rb");
    if (!f) return 2;
    auto meta = rd(f);
    const uint32_t n = meta[0];
    auto rptr = rd(f), cidx = rd(f), dptr = rd(f);
    fclose(f);
    const size_t nnz = cidx.size();
    std::mt19937 g(7);
    std::uniform_real_distribution<float> U(-1.f, 1.f);
    std::vector<float> lu(nnz);
    for (auto &v : lu) v = U(g) * 0.1f;
    for (uint32_t r = 0; r < n; ++r) lu[dptr[r]] = 1.f + std::fabs(U(g));
    std::vector<float> b(n);
    for (auto &v : b) v = U(g);
    TriHost lo, up;
    tri_levels(n, +1, [&](uint32_t r, auto fn) {
        for (uint32_t q = rptr[r]; q < dptr[r]; ++q) fn(cidx[q]);
    }, [&](uint32_t r) { return dptr[r] - rptr[r]; }, lo);
    tri_levels(n, -1, [&](uint32_t r, auto fn) {
        for (uint32_t q = dptr[r] + 1; q < rptr[r + 1]; ++q) fn(cidx[q]);
    }, [&](uint32_t r) { return rptr[r + 1] - dptr[r] - 1; }, up);
    DfHost dlo, dup;
    build_df(lo, [&](uint32_t r, std::vector<uint2> &out) {
        for (uint32_t q = rptr[r]; q < dptr[r]; ++q) out.push_back(make_uint2(cidx[q], q));
    }, false, [&](uint32_t) { return 0u; }, long_lanes, dlo);
    build_df(up, [&](uint32_t r, std::vector<uint2> &out) {
        for (uint32_t q = rptr[r + 1]; q > dptr[r] + 1; --q) out.push_back(make_uint2(cidx[q - 1], q - 1));
    }, true, [&](uint32_t r) { return dptr[r]; }, long_lanes, dup);
    // the plan builder's own host check accepts these tables and refuses
    // planted mistakes: a task run early (swapped with its successor), a row
    // dropped, a dependency on a row that runs later
    std::string err;
    if (!check_df(dlo, n, (uint32_t)nnz, "L", err) || !check_df(dup, n, (uint32_t)nnz, "U", err)) {
        printf("check_df refused valid tables: %s\n", err.c_str());
        return 1;
    }
    int planted = 0, caught = 0;
    for (const DfHost *src : {&dlo, &dup}) {
        for (size_t k = 0; k + 1 < src->task.size(); ++k) {   // first swap that reorders dependent tasks
            DfHost bad = *src;
            std::swap(bad.task[k], bad.task[k + 1]);
            std::string e2;
            if (!check_df(bad, n, (uint32_t)nnz, "swap", e2)) {
                ++planted;
                ++caught;
                break;
            }
        }
        if (!src->task.empty()) {
            DfHost bad = *src;
            bad.task.pop_back();
            std::string e2;
            ++planted;
            caught += !check_df(bad, n, (uint32_t)nnz, "drop", e2);
        }
        for (size_t k = 0; k < src->task.size(); ++k) {   // a slot of a row made to read the task's last row
            const uint4 d = src->task[k];
            if ((d.z & kLongTask) || d.z < 2) continue;
            const uint32_t first_len = src->rl[d.y].y;
            if (first_len == 0) continue;
            DfHost bad = *src;
            bad.idx[d.x] = src->rl[d.y + d.z - 1].x;   // lane 0's first entry
            std::string e2;
            ++planted;
            caught += !check_df(bad, n, (uint32_t)nnz, "dep", e2);
            break;
        }
    }
    printf("check_df: planted %d mistakes, refused %d\n", planted, caught);
    if (caught != planted) return 1;

    std::vector<float> y, x;
    if (emulate("L", dlo, n, lu, false, b, y)) return 1;
    if (emulate("U", dup, n, lu, true, y, x)) return 1;
    // sequential reference
    std::vector<float> yr = b, xr;
    for (uint32_t r = 0; r < n; ++r) {
        float acc = yr[r];
        for (uint32_t q = rptr[r]; q < dptr[r]; ++q) acc = acc - lu[q] * yr[cidx[q]];
        yr[r] = acc;
    }
    xr = yr;
    for (uint32_t r = n; r-- > 0;) {
        float acc = xr[r];
        for (uint32_t q = rptr[r + 1]; q > dptr[r] + 1; --q) acc = acc - lu[q - 1] * xr[cidx[q - 1]];
        xr[r] = acc / lu[dptr[r]];
    }
    for (uint32_t r = 0; r < n; ++r)
        if (std::memcmp(&x[r], &xr[r], 4)) { printf("x[%u] differs %g %g\n", r, x[r], xr[r]); return 1; }
    printf("ok n=%u: emulated sweeps equal the sequential solve bit for bit\n", n);
    return 0;
}
