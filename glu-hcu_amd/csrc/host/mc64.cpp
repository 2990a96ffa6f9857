// MC64 (job 5): maximum-product bipartite matching with dual variables,
// the static-pivoting step of the host preprocess.
//
// Reproduces src/nicslu/source/nicslu_mc64.c (_I_NicsLU_MC64ad :12-100 and
// _I_NicsLU_MC64wd :362-841, heap helpers :103-359).  The matching is a
// shortest-augmenting-path (Dijkstra) search over costs
//     c(r, c) = log(max_k |a(r,k)|) - log|a(r,c)|,
// and which of several equal-cost candidates wins decides the permutation,
// so the cheap-assignment pass, the two-queue Dijkstra (a list "B" of
// minimum-distance columns in q[low..up) and a binary heap in q[0..qlen)),
// the shared use of the q array, and every comparison (< versus <=) are kept
// exactly as the reference has them.  Orientation is NICSLU's: ap indexes
// "rows" (caller columns), ai holds "columns" (caller rows).
//
// A derivative work of NICSLU, Copyright (c) 2011-2013 Tsinghua University,
// distributed under the GNU Lesser General Public License, version 2.1 or
// later (LICENSES/LGPL-2.1.txt and LICENSES/NOTICE in this directory).
#include "glu_host.h"

#include <cfloat>
#include <cmath>
#include <cstring>

namespace glu {
namespace {

using i32 = int32_t;
using u32 = uint32_t;

// Min-heap stored in q[0..qlen), key dist[col], position of col in loc[col].
// Sift-up of col from loc[col] (nicslu_mc64.c:103-156, iway=2).
void heap_sift_up(u32 col, u32 n, i32 *q, const double *dist, i32 *loc)
{
    double dc = dist[col];
    u32 pos = (u32)loc[col];
    for (u32 guard = 0; guard < n; ++guard) {
        if (pos == 0) break;
        u32 parent = (pos - 1) >> 1;
        i32 pc = q[parent];
        if (dc >= dist[pc]) break;
        q[pos] = pc;
        loc[pc] = (i32)pos;
        pos = parent;
    }
    q[pos] = (i32)col;
    loc[col] = (i32)pos;
}

// Remove the root: move the last element to the top and sift it down
// (nicslu_mc64.c:158-232, iway=2).
void heap_pop_root(i32 *qlen, u32 n, i32 *q, const double *dist, i32 *loc)
{
    i32 last = *qlen - 1;
    u32 moved = (u32)q[last];
    double dm = dist[moved];
    --(*qlen);
    i32 pos = 0;
    for (u32 guard = 0; guard < n; ++guard) {
        i32 child = ((pos + 1) << 1) - 1;
        if (child > last) break;
        double dk = dist[q[child]];
        if (child < last) {
            double dr = dist[q[child + 1]];
            if (dk > dr) { ++child; dk = dr; }
        }
        if (dm <= dk) break;
        q[pos] = q[child];
        loc[q[pos]] = pos;
        pos = child;
    }
    q[pos] = (i32)moved;
    loc[moved] = pos;
}

// Delete the element at position at: the last element takes its place, then
// sifts up and down (nicslu_mc64.c:234-359, iway=2).
void heap_delete_at(i32 at, i32 *qlen, u32 n, i32 *q, const double *dist, i32 *loc)
{
    i32 last = *qlen - 1;
    if (at == last) { --(*qlen); return; }
    u32 moved = (u32)q[last];
    double dm = dist[moved];
    --(*qlen);
    i32 pos = at;
    for (u32 guard = 0; guard < n; ++guard) {
        if (pos <= 0) break;
        i32 parent = (pos - 1) >> 1;
        i32 pc = q[parent];
        if (dm >= dist[pc]) break;
        q[pos] = pc;
        loc[pc] = pos;
        pos = parent;
    }
    q[pos] = (i32)moved;
    loc[moved] = pos;
    for (u32 guard = 0; guard < n; ++guard) {
        i32 child = ((pos + 1) << 1) - 1;
        if (child > last) break;
        double dk = dist[q[child]];
        if (child < last) {
            double dr = dist[q[child + 1]];
            if (dk > dr) { ++child; dk = dr; }
        }
        if (dm <= dk) break;
        i32 qc = q[child];
        q[pos] = qc;
        loc[qc] = pos;
        pos = child;
    }
    q[pos] = (i32)moved;
    loc[moved] = pos;
}

// The weighted matching proper (nicslu_mc64.c:362-841).  cost[k] per entry.
// Outputs: iperm[col] = matched row, u = column duals, d = row values.
// Returns the matching cardinality.
i32 weighted_match(u32 n, const u32 *ai, const u32 *ap, const double *cost, i32 *iperm,
                   i32 *jperm, i32 *out, i32 *pr, i32 *q, i32 *l, double *u, double *d)
{
    u32 num = 0;
    for (u32 c = 0; c < n; ++c) u[c] = DBL_MAX;
    for (u32 c = 0; c < n; ++c) { d[c] = 0.0; iperm[c] = -1; jperm[c] = -1; l[c] = -1; }
    for (u32 r = 0; r < n; ++r) pr[r] = (i32)ap[r];

    // Column minima; on ties the later row wins (":405 continue on >").
    for (u32 r = 0; r < n; ++r) {
        for (u32 k = ap[r]; k < ap[r + 1]; ++k) {
            u32 col = ai[k];
            if (cost[k] > u[col]) continue;
            u[col] = cost[k];
            iperm[col] = (i32)r;
            l[col] = (i32)k;
        }
    }
    // Cheap assignment: each column takes its minimising row if still free.
    for (u32 c = 0; c < n; ++c) {
        i32 r = iperm[c];
        if (r < 0) continue;
        iperm[c] = -1;
        if (jperm[r] >= 0) continue;
        ++num;
        iperm[c] = r;
        jperm[r] = l[c];
    }

    if (num != n) {
        // Second cheap pass over unassigned rows, with one-step augmentation
        // through the remaining part of an assigned row (:438-530).
        for (u32 r = 0; r < n; ++r) {
            if (jperm[r] >= 0) continue;
            u32 beg = ap[r], end = ap[r + 1];
            if (beg >= end) return kErrStructSingular;

            double vj = DBL_MAX;
            u32 i0 = 0, k0 = 0;
            for (u32 k = beg; k < end; ++k) {
                u32 col = ai[k];
                double di = cost[k] - u[col];
                if (di > vj) continue;
                bool take = (di < vj || di == DBL_MAX);
                if (!take && (iperm[col] >= 0 || iperm[i0] < 0)) continue;
                vj = di;
                i0 = col;
                k0 = k;
            }
            d[r] = vj;
            u32 k = k0, col = i0;
            bool assign = iperm[col] < 0;
            if (!assign) {
                bool found = false;
                i32 jj = -1;
                u32 kk = 0, col1 = 0;
                for (k = k0; k < end && !found; ++k) {
                    col = ai[k];
                    if (cost[k] - u[col] > vj) continue;
                    jj = iperm[col];
                    if (jj < 0) continue;  // unreachable for finite costs
                    u32 kk1 = (u32)pr[jj], kk2 = ap[jj + 1];
                    if (kk1 >= kk2) continue;
                    for (kk = kk1; kk < kk2; ++kk) {
                        col1 = ai[kk];
                        if (iperm[col1] >= 0) continue;
                        if (cost[kk] - u[col1] <= d[jj]) { found = true; break; }
                    }
                    if (found) break;
                    pr[jj] = (i32)kk2;
                }
                if (!found) continue;
                // Reassign row jj to col1, freeing col for row r.
                jperm[jj] = (i32)kk;
                iperm[col1] = jj;
                pr[jj] = (i32)(kk + 1);
                assign = true;
            }
            ++num;
            jperm[r] = (i32)k;
            iperm[col] = (i32)r;
            pr[r] = (i32)(k + 1);
        }
    }

    if (num != n) {
        for (u32 c = 0; c < n; ++c) d[c] = DBL_MAX;
        for (u32 c = 0; c < n; ++c) l[c] = -1;

        // Dijkstra from every unassigned row (:547-786).
        for (u32 root = 0; root < n; ++root) {
            if (jperm[root] >= 0) continue;
            double dmin = DBL_MAX, csp = DBL_MAX;
            i32 qlen = 0, low = (i32)n, up = (i32)n;
            u32 isp = 0, jsp = 0;
            u32 r = root;
            pr[r] = -2;

            for (u32 k = ap[r]; k < ap[r + 1]; ++k) {
                u32 col = ai[k];
                double dnew = cost[k] - u[col];
                if (dnew >= csp) continue;
                if (iperm[col] < 0) {
                    csp = dnew;
                    isp = k;
                    jsp = r;
                } else {
                    if (dnew < dmin) dmin = dnew;
                    d[col] = dnew;
                    q[qlen++] = (i32)k;
                }
            }

            // q[0..q0) holds entry positions; the heap and list B are built
            // in the same array as the positions are consumed.
            u32 q0 = (u32)qlen;
            qlen = 0;
            for (u32 t = 0; t < q0; ++t) {
                u32 k = (u32)q[t];
                u32 col = ai[k];
                if (d[col] >= csp) { d[col] = DBL_MAX; continue; }
                if (d[col] <= dmin) {
                    q[--low] = (i32)col;
                    l[col] = low;
                } else {
                    l[col] = qlen++;
                    heap_sift_up(col, n, q, d, l);
                }
                i32 jj = iperm[col];
                out[jj] = (i32)k;
                pr[jj] = (i32)r;
            }

            for (u32 step = 0; step < num; ++step) {
                if (low == up) {
                    // List B empty: move all minimum-distance heap entries.
                    if (qlen == 0) break;
                    u32 col = (u32)q[0];
                    if (d[col] >= csp) break;
                    dmin = d[col];
                    for (;;) {
                        heap_pop_root(&qlen, n, q, d, l);
                        q[--low] = (i32)col;
                        l[col] = low;
                        if (qlen == 0) break;
                        col = (u32)q[0];
                        if (d[col] > dmin) break;
                    }
                }
                u32 cq = (u32)q[up - 1];
                double dq = d[cq];
                if (dq >= csp) break;
                --up;
                u32 rr = (u32)iperm[cq];
                double vj = dq - cost[jperm[rr]] + u[cq];
                for (u32 k = ap[rr]; k < ap[rr + 1]; ++k) {
                    u32 col = ai[k];
                    if (l[col] >= up) continue;
                    double dnew = vj + cost[k] - u[col];
                    if (dnew >= csp) continue;
                    if (iperm[col] < 0) {
                        csp = dnew;
                        isp = k;
                        jsp = rr;
                    } else {
                        double di = d[col];
                        if (di <= dnew) continue;
                        if (l[col] >= low) continue;
                        d[col] = dnew;
                        if (dnew <= dmin) {
                            if (l[col] >= 0) heap_delete_at(l[col], &qlen, n, q, d, l);
                            q[--low] = (i32)col;
                            l[col] = low;
                        } else {
                            if (l[col] < 0) l[col] = qlen++;
                            heap_sift_up(col, n, q, d, l);
                        }
                        i32 jj = iperm[col];
                        out[jj] = (i32)k;
                        pr[jj] = (i32)rr;
                    }
                }
            }

            if (csp != DBL_MAX) {
                // Augment along the shortest path and update the duals.
                ++num;
                u32 col = ai[isp];
                iperm[col] = (i32)jsp;
                jperm[jsp] = (i32)isp;
                u32 rr = jsp;
                for (u32 step = 0; step < num; ++step) {
                    i32 jj = pr[rr];
                    if (jj == -2) break;
                    u32 k = (u32)out[rr];
                    u32 c2 = ai[k];
                    iperm[c2] = jj;
                    jperm[jj] = (i32)k;
                    rr = (u32)jj;
                }
                for (i32 t = up; t < (i32)n; ++t) {
                    u32 c2 = (u32)q[t];
                    u[c2] = u[c2] + d[c2] - csp;
                }
            }
            for (i32 t = low; t < (i32)n; ++t) {
                u32 c2 = (u32)q[t];
                d[c2] = DBL_MAX;
                l[c2] = -1;
            }
            for (i32 t = 0; t < qlen; ++t) {
                u32 c2 = (u32)q[t];
                d[c2] = DBL_MAX;
                l[c2] = -1;
            }
        }
    }

    // Final duals (:788-806).
    for (u32 r = 0; r < n; ++r) {
        i32 kk = jperm[r];
        d[r] = (kk >= 0) ? cost[kk] - u[ai[kk]] : 0.0;
        if (iperm[r] < 0) u[r] = 0.0;
    }
    return (i32)num;
}

}  // namespace

int64_t mc64_match(uint32_t n, const uint32_t *ai, const uint32_t *ap, const double *ax,
                   std::vector<int32_t> &match, std::vector<double> &u,
                   std::vector<double> &v)
{
    uint32_t nnz = ap[n];
    std::vector<double> cost(nnz), rowmax(n);
    const double rinf = DBL_MAX / n;
    // Costs relative to each row's largest magnitude (nicslu_mc64.c:29-68).
    for (uint32_t r = 0; r < n; ++r) {
        double mx = 0.0;
        for (uint32_t k = ap[r]; k < ap[r + 1]; ++k) {
            double a = ax[k];
            cost[k] = a > 0 ? a : -a;
            if (cost[k] > mx) mx = cost[k];
        }
        rowmax[r] = mx;
        if (mx == 0.0) return kErrStructSingular;
        double lmx = std::log(mx);
        for (uint32_t k = ap[r]; k < ap[r + 1]; ++k)
            cost[k] = (cost[k] != 0.0) ? lmx - std::log(cost[k]) : rinf;
    }

    match.assign(n, -1);
    u.assign(n, 0.0);
    v.assign(n, 0.0);
    std::vector<int32_t> jperm(n), out(n), pr(n), q(n), l(n);
    int32_t card = weighted_match(n, ai, ap, cost.data(), match.data(), jperm.data(), out.data(),
                                  pr.data(), q.data(), l.data(), u.data(), v.data());
    if (card < 0) return card;
    if ((uint32_t)card < n) return kErrStructSingular;
    // Row duals shifted by the row maxima (:84-97).
    for (uint32_t r = 0; r < n; ++r)
        v[r] = (rowmax[r] != 0.0) ? v[r] - std::log(rowmax[r]) : 0.0;
    return n;
}

}  // namespace glu
