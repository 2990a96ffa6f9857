// Approximate minimum degree ordering of A+A' (fill-reducing permutation).
//
// Reproduces src/nicslu/source/nicslu_amd.c: the A+A' construction
// (_I_NicsLU_AAT :59-140, _I_NicsLU_AAT2 :142-214), the AMD elimination
// (_I_NicsLU_AMD :216-830, itself NICSLU's edition of Davis' AMD with dense
// rows set aside, aggressive absorption and a hash for supervariable
// detection), and the post-order of the assembly tree (_I_NicsLU_PostOrder
// :832-899, _I_NicsLU_PostTree :901-931).  Degree-list tie-breaks, the
// garbage-collection trigger (pfree >= iwlen) and the 32-bit wrap of the
// supervariable hash all influence the ordering, so they are kept as is.
//
// A derivative work of NICSLU, Copyright (c) 2011-2013 Tsinghua University,
// distributed under the GNU Lesser General Public License, version 2.1 or
// later (LICENSES/LGPL-2.1.txt and LICENSES/NOTICE in this directory).
#include "glu_host.h"

#include <climits>
#include <cmath>
#include <cstring>

namespace glu {
namespace {

using i32 = int32_t;
using u32 = uint32_t;

inline i32 flip(i32 i) { return -i - 2; }

i32 reset_marks(i32 wflg, i32 wbig, i32 *w, i32 n)
{
    if (wflg < 2 || wflg >= wbig) {
        for (i32 x = 0; x < n; ++x)
            if (w[x] != 0) w[x] = 1;
        wflg = 2;
    }
    return wflg;
}

// Degrees of A+A' excluding the diagonal, using the "scan row k, then walk
// the transposed part of each earlier row j" traversal.  tp is scratch.
u32 aat_lengths(u32 n, const u32 *ai, const u32 *ap, i32 *len, i32 *tp)
{
    std::memset(len, 0, sizeof(i32) * n);
    for (u32 k = 0; k < n; ++k) {
        u32 p = ap[k], p2 = ap[k + 1];
        while (p < p2) {
            u32 j = ai[p];
            if (j < k) {
                ++len[j];
                ++len[k];
                ++p;
            } else {
                if (j == k) ++p;
                break;
            }
            u32 pj = (u32)tp[j], pj2 = ap[j + 1];
            while (pj < pj2) {
                u32 i = ai[pj];
                if (i < k) {
                    ++len[i];
                    ++len[j];
                    ++pj;
                } else {
                    if (i == k) ++pj;
                    break;
                }
            }
            tp[j] = (i32)pj;
        }
        tp[k] = (i32)p;
    }
    for (u32 j = 0; j < n; ++j)
        for (u32 pj = (u32)tp[j]; pj < ap[j + 1]; ++pj) {
            ++len[ai[pj]];
            ++len[j];
        }
    u32 total = 0;
    for (u32 k = 0; k < n; ++k) total += (u32)len[k];
    return total;
}

// Same traversal, writing the adjacency lists into iw at pe[]/sp[].
void aat_fill(u32 n, const u32 *ai, const u32 *ap, const i32 *len, i32 *pe, i32 *sp, i32 *iw,
              i32 *tp)
{
    u32 pfree = 0;
    for (u32 j = 0; j < n; ++j) {
        pe[j] = sp[j] = (i32)pfree;
        pfree += (u32)len[j];
    }
    for (u32 k = 0; k < n; ++k) {
        u32 p = ap[k], p2 = ap[k + 1];
        while (p < p2) {
            u32 j = ai[p];
            if (j < k) {
                iw[sp[j]++] = (i32)k;
                iw[sp[k]++] = (i32)j;
                ++p;
            } else {
                if (j == k) ++p;
                break;
            }
            u32 pj = (u32)tp[j], pj2 = ap[j + 1];
            while (pj < pj2) {
                u32 i = ai[pj];
                if (i < k) {
                    iw[sp[i]++] = (i32)j;
                    iw[sp[j]++] = (i32)i;
                    ++pj;
                } else {
                    if (i == k) ++pj;
                    break;
                }
            }
            tp[j] = (i32)pj;
        }
        tp[k] = (i32)p;
    }
    for (u32 j = 0; j < n; ++j)
        for (u32 pj = (u32)tp[j]; pj < ap[j + 1]; ++pj) {
            u32 i = ai[pj];
            iw[sp[i]++] = (i32)j;
            iw[sp[j]++] = (i32)i;
        }
}

// Depth-first post-order of one tree of the assembly forest.
i32 post_tree(i32 root, i32 k, i32 *child, const i32 *sibling, i32 *order, i32 *stack)
{
    i32 head = 0;
    stack[0] = root;
    while (head >= 0) {
        i32 i = stack[head];
        if (child[i] != -1) {
            for (i32 f = child[i]; f != -1; f = sibling[f]) ++head;
            i32 h = head;
            for (i32 f = child[i]; f != -1; f = sibling[f]) stack[h--] = f;
            child[i] = -1;
        } else {
            --head;
            order[i] = k++;
        }
    }
    return k;
}

// Post-order the assembly tree, visiting the child with the largest front
// last (:832-899).
void post_order(i32 n, const i32 *parent, const i32 *nv, const i32 *fsize, i32 *order,
                i32 *child, i32 *sibling, i32 *stack)
{
    for (i32 j = 0; j < n; ++j) { child[j] = -1; sibling[j] = -1; }
    for (i32 j = n - 1; j >= 0; --j) {
        if (nv[j] > 0) {
            i32 par = parent[j];
            if (par != -1) {
                sibling[j] = child[par];
                child[par] = j;
            }
        }
    }
    for (i32 i = 0; i < n; ++i) {
        if (nv[i] > 0 && child[i] != -1) {
            i32 fprev = -1, maxfr = -1, bigfp = -1, bigf = -1;
            for (i32 f = child[i]; f != -1; f = sibling[f]) {
                i32 frsize = fsize[f];
                if (frsize >= maxfr) {
                    maxfr = frsize;
                    bigfp = fprev;
                    bigf = f;
                }
                fprev = f;
            }
            i32 fnext = sibling[bigf];
            if (fnext != -1) {
                if (bigfp == -1) child[i] = fnext;
                else sibling[bigfp] = fnext;
                sibling[bigf] = -1;
                sibling[fprev] = bigf;
            }
        }
    }
    for (i32 j = 0; j < n; ++j) order[j] = -1;
    i32 k = 0;
    for (i32 i = 0; i < n; ++i)
        if (parent[i] == -1 && nv[i] > 0) k = post_tree(i, k, child, sibling, order, stack);
}

// The minimum-degree elimination on the quotient graph held in iw.
// pe/len/iw: adjacency lists (pfree = first free slot, iwlen = capacity).
// Outputs last[k] (k-th pivot) and next[i] (position of i).
void amd_core(i32 n, i32 pfree, i32 iwlen, i32 *pe, i32 *iw, i32 *len, i32 *nv, i32 *head,
              i32 *elen, i32 *degree, i32 *w, i32 *last, i32 *next, double alpha, i32 aggr)
{
    i32 me = -1, mindeg = 0, nel = 0, lemax = 0;

    i32 dense = (i32)((alpha < 0.0) ? (n - 2) : (alpha * std::sqrt((double)n)));
    if (dense < 16) dense = 16;
    if (dense > n) dense = n;

    for (i32 i = 0; i < n; ++i) {
        last[i] = -1;
        head[i] = -1;
        next[i] = -1;
        elen[i] = 0;
        degree[i] = len[i];
        nv[i] = 1;
        w[i] = 1;
    }
    const i32 wbig = INT_MAX - n;
    i32 wflg = reset_marks(0, wbig, w, n);
    i32 ndense = 0;

    for (i32 i = 0; i < n; ++i) {
        i32 deg = degree[i];
        if (deg == 0) {
            elen[i] = flip(1);
            ++nel;
            pe[i] = -1;
            w[i] = 0;
        } else if (deg > dense) {
            ++ndense;
            nv[i] = 0;
            elen[i] = -1;
            ++nel;
            pe[i] = -1;
        } else {
            i32 inext = head[deg];
            if (inext != -1) last[inext] = i;
            next[i] = inext;
            head[deg] = i;
        }
    }

    while (nel < n) {
        // Pivot: first supervariable in the lowest non-empty degree list.
        i32 deg;
        for (deg = mindeg; deg < n; ++deg) {
            me = head[deg];
            if (me != -1) break;
        }
        mindeg = deg;
        {
            i32 inext = next[me];
            if (inext != -1) last[inext] = -1;
            head[deg] = inext;
        }
        i32 elenme = elen[me];
        i32 nvpiv = nv[me];
        nel += nvpiv;
        nv[me] = -nvpiv;
        i32 degme = 0;
        i32 pme1, pme2;

        if (elenme == 0) {
            // New element built in place over me's own variable list.
            pme1 = pe[me];
            pme2 = pme1 - 1;
            i32 pme3 = pme1 + len[me];
            for (i32 p = pme1; p < pme3; ++p) {
                i32 i = iw[p];
                i32 nvi = nv[i];
                if (nvi > 0) {
                    degme += nvi;
                    nv[i] = -nvi;
                    iw[++pme2] = i;
                    i32 ilast = last[i], inext = next[i];
                    if (inext != -1) last[inext] = ilast;
                    if (ilast != -1) next[ilast] = inext;
                    else head[degree[i]] = inext;
                }
            }
        } else {
            // New element built in free space, absorbing me's elements.
            i32 p = pe[me];
            pme1 = pfree;
            i32 slenme = len[me] - elenme;
            for (i32 knt1 = 1; knt1 <= elenme + 1; ++knt1) {
                i32 e, pj, ln;
                if (knt1 > elenme) {
                    e = me;
                    pj = p;
                    ln = slenme;
                } else {
                    e = iw[p++];
                    pj = pe[e];
                    ln = len[e];
                }
                for (i32 knt2 = 1; knt2 <= ln; ++knt2) {
                    i32 i = iw[pj++];
                    i32 nvi = nv[i];
                    if (nvi <= 0) continue;
                    if (pfree >= iwlen) {
                        // Compress iw: keep every live list, then the
                        // partially built element at the end.
                        pe[me] = p;
                        len[me] -= knt1;
                        if (len[me] == 0) pe[me] = -1;
                        pe[e] = pj;
                        len[e] = ln - knt2;
                        if (len[e] == 0) pe[e] = -1;
                        for (i32 j = 0; j < n; ++j) {
                            i32 pn = pe[j];
                            if (pn >= 0) {
                                pe[j] = iw[pn];
                                iw[pn] = flip(j);
                            }
                        }
                        i32 psrc = 0, pdst = 0, pend = pme1 - 1;
                        while (psrc <= pend) {
                            i32 j = flip(iw[psrc++]);
                            if (j >= 0) {
                                iw[pdst] = pe[j];
                                pe[j] = pdst++;
                                i32 lenj = len[j];
                                for (i32 knt3 = 0; knt3 < lenj - 1; ++knt3) iw[pdst++] = iw[psrc++];
                            }
                        }
                        i32 p1 = pdst;
                        for (psrc = pme1; psrc < pfree; ++psrc) iw[pdst++] = iw[psrc];
                        pme1 = p1;
                        pfree = pdst;
                        pj = pe[e];
                        p = pe[me];
                    }
                    degme += nvi;
                    nv[i] = -nvi;
                    iw[pfree++] = i;
                    i32 ilast = last[i], inext = next[i];
                    if (inext != -1) last[inext] = ilast;
                    if (ilast != -1) next[ilast] = inext;
                    else head[degree[i]] = inext;
                }
                if (e != me) {
                    pe[e] = flip(me);
                    w[e] = 0;
                }
            }
            pme2 = pfree - 1;
        }

        degree[me] = degme;
        pe[me] = pme1;
        len[me] = pme2 - pme1 + 1;
        elen[me] = flip(nvpiv + degme);
        wflg = reset_marks(wflg, wbig, w, n);

        // |Le \ Lme| for every element e adjacent to a variable of me.
        for (i32 pme = pme1; pme <= pme2; ++pme) {
            i32 i = iw[pme];
            i32 eln = elen[i];
            if (eln <= 0) continue;
            i32 nvi = -nv[i];
            i32 wnvi = wflg - nvi;
            i32 pend2 = pe[i] + eln;
            for (i32 p = pe[i]; p < pend2; ++p) {
                i32 e = iw[p];
                i32 we = w[e];
                if (we >= wflg) we -= nvi;
                else if (we != 0) we = degree[e] + wnvi;
                w[e] = we;
            }
        }

        // Approximate degrees, aggressive absorption, mass elimination,
        // and hashing of the variables of me.
        for (i32 pme = pme1; pme <= pme2; ++pme) {
            i32 i = iw[pme];
            i32 p1 = pe[i];
            i32 p2 = p1 + elen[i] - 1;
            i32 pn = p1;
            u32 hash = 0;
            i32 dg = 0;
            for (i32 p = p1; p <= p2; ++p) {
                i32 e = iw[p];
                i32 we = w[e];
                if (we == 0) continue;
                i32 dext = we - wflg;
                if (aggr && dext <= 0) {
                    pe[e] = flip(me);
                    w[e] = 0;
                    continue;
                }
                dg += dext;
                iw[pn++] = e;
                hash += (u32)e;
            }
            elen[i] = pn - p1 + 1;
            i32 p3 = pn;
            i32 p4 = p1 + len[i];
            for (i32 p = p2 + 1; p < p4; ++p) {
                i32 j = iw[p];
                i32 nvj = nv[j];
                if (nvj > 0) {
                    dg += nvj;
                    iw[pn++] = j;
                    hash += (u32)j;
                }
            }
            if (elen[i] == 1 && p3 == pn) {
                // Mass elimination: i is adjacent only to me.
                pe[i] = flip(me);
                i32 nvi = -nv[i];
                degme -= nvi;
                nvpiv += nvi;
                nel += nvi;
                nv[i] = 0;
                elen[i] = -1;
            } else {
                if (dg < degree[i]) degree[i] = dg;
                iw[pn] = iw[p3];
                iw[p3] = iw[p1];
                iw[p1] = me;
                len[i] = pn - p1 + 1;
                hash = hash % (u32)n;
                i32 j = head[hash];
                if (j <= -1) {
                    next[i] = flip(j);
                    head[hash] = flip(i);
                } else {
                    next[i] = last[j];
                    last[j] = i;
                }
                last[i] = (i32)hash;
            }
        }

        degree[me] = degme;
        if (degme > lemax) lemax = degme;
        wflg += lemax;
        wflg = reset_marks(wflg, wbig, w, n);

        // Supervariable detection within each hash bucket.
        for (i32 pme = pme1; pme <= pme2; ++pme) {
            i32 i = iw[pme];
            if (nv[i] >= 0) continue;
            u32 hash = (u32)last[i];
            i32 j = head[hash];
            if (j == -1) {
                i = -1;
            } else if (j < -1) {
                i = flip(j);
                head[hash] = -1;
            } else {
                i = last[j];
                last[j] = -1;
            }
            while (i != -1 && next[i] != -1) {
                i32 ln = len[i], eln = elen[i];
                for (i32 p = pe[i] + 1; p < pe[i] + ln; ++p) w[iw[p]] = wflg;
                i32 jlast = i;
                j = next[i];
                while (j != -1) {
                    bool ok = (len[j] == ln && elen[j] == eln);
                    for (i32 p = pe[j] + 1; ok && p < pe[j] + ln; ++p)
                        if (w[iw[p]] != wflg) ok = false;
                    if (ok) {
                        pe[j] = flip(i);
                        nv[i] += nv[j];
                        nv[j] = 0;
                        elen[j] = -1;
                        j = next[j];
                        next[jlast] = j;
                    } else {
                        jlast = j;
                        j = next[j];
                    }
                }
                ++wflg;
                i = next[i];
            }
        }

        // Restore degree lists; drop non-principal variables from me.
        i32 p = pme1;
        i32 nleft = n - nel;
        for (i32 pme = pme1; pme <= pme2; ++pme) {
            i32 i = iw[pme];
            i32 nvi = -nv[i];
            if (nvi <= 0) continue;
            nv[i] = nvi;
            i32 dg = degree[i] + degme - nvi;
            if (nleft - nvi < dg) dg = nleft - nvi;
            i32 inext = head[dg];
            if (inext != -1) last[inext] = i;
            next[i] = inext;
            last[i] = -1;
            head[dg] = i;
            if (dg < mindeg) mindeg = dg;
            degree[i] = dg;
            iw[p++] = i;
        }
        nv[me] = nvpiv;
        len[me] = p - pme1;
        if (len[me] == 0) {
            pe[me] = -1;
            w[me] = 0;
        }
        if (elenme != 0) pfree = p;
    }

    // Assembly tree: parents of non-principal variables point at the
    // element that absorbed them (:748-781).
    for (i32 i = 0; i < n; ++i) {
        pe[i] = flip(pe[i]);
        elen[i] = flip(elen[i]);
    }
    for (i32 i = 0; i < n; ++i) {
        if (nv[i] != 0) continue;
        i32 j = pe[i];
        if (j == -1) continue;
        while (nv[j] == 0) j = pe[j];
        i32 e = j;
        j = i;
        while (nv[j] == 0) {
            i32 jnext = pe[j];
            pe[j] = e;
            j = jnext;
        }
    }

    post_order(n, pe, nv, elen, w, head, next, last);

    for (i32 i = 0; i < n; ++i) { head[i] = -1; next[i] = -1; }
    for (i32 e = 0; e < n; ++e) {
        i32 k = w[e];
        if (k != -1) head[k] = e;
    }
    nel = 0;
    for (i32 k = 0; k < n; ++k) {
        i32 e = head[k];
        if (e == -1) break;
        next[e] = nel;
        nel += nv[e];
    }
    for (i32 i = 0; i < n; ++i) {
        if (nv[i] != 0) continue;
        i32 e = pe[i];
        if (e != -1) {
            next[i] = next[e];
            ++next[e];
        } else {
            next[i] = nel++;
        }
    }
    for (i32 i = 0; i < n; ++i) last[next[i]] = i;
}

}  // namespace

void amd_order(uint32_t n, const uint32_t *ai, const uint32_t *ap, double alpha, int aggressive,
               std::vector<int32_t> &perm, std::vector<int32_t> &pinv)
{
    std::vector<i32> len(n), scratch(n);
    u32 nzaat = aat_lengths(n, ai, ap, len.data(), scratch.data());
    // Same workspace budget as the reference: iw holds n + nzaat*1.2 slots,
    // which fixes when the garbage collection runs.
    i32 iwlen = (i32)(n + nzaat + nzaat / 5);
    std::vector<i32> pe(n), sp(n), tp(n), iw((size_t)iwlen);
    aat_fill(n, ai, ap, len.data(), pe.data(), sp.data(), iw.data(), tp.data());
    std::vector<i32> nv(n), head(n), elen(n), degree(n), w(n);
    perm.assign(n, 0);
    pinv.assign(n, 0);
    amd_core((i32)n, (i32)nzaat, iwlen, pe.data(), iw.data(), len.data(), nv.data(), head.data(),
             elen.data(), degree.data(), w.data(), perm.data(), pinv.data(), alpha, aggressive);
}

}  // namespace glu
