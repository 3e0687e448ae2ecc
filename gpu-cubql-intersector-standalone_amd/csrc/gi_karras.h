// gi_karras.h -- LBVH topology over sorted Morton keys (Karras 2012,
// "Maximizing Parallelism in the Construction of BVHs, Octrees, and k-d
// Trees"), as k_build_chunks runs it per node; host and device, so that the
// CPU tests check the kernel's own node logic against a plain Karras build.
//
// Leaves are the sorted keys 0..n-1 (node n-1+k), internal node i one of
// n-1; ties between equal keys break on the index (composite key).
#pragma once

#include <stdint.h>

#include "gi_device.h"
#include "gi_int.h"

namespace gi {

// Sorted keys (Morton codes, or mesh id << 3 axis_bits | Morton code for a
// forest), the chunk's slice read from LDS.
template <class K>
struct KeySrc {
    const K* __restrict__ g;
    const K* s;
    int cs, ce, n;
    GI_HD K at(int j) const { return (j >= cs && j < ce) ? s[j - cs] : g[j]; }
};

GI_HD int clz_key(uint32_t x) { return x ? __builtin_clz(x) : 32; }
GI_HD int clz_key(uint64_t x) { return x ? __builtin_clzll(x) : 64; }

// common-prefix length of the composite keys (key, index). In a forest the
// mesh id bits lead, so the leaves of one mesh share a longer prefix with
// each other than with any other leaf and form one subtree.
template <class K>
GI_HD int delta_keys(K a, int i, K b, int j) {
    if (a != b) return clz_key(a ^ b);
    return 8 * (int)sizeof(K) + clz_key((uint32_t)i ^ (uint32_t)j);
}
template <class K>
GI_HD int delta(const KeySrc<K>& k, int i, int j) {
    if (j < 0 || j >= k.n) return -1;
    return delta_keys(k.at(i), i, k.at(j), j);
}

// The node visited right after every subtree whose last leaf is m - 1: the
// right child that starts at leaf m. In Karras' numbering a right child
// covering [m, x] is internal node m, which is the case iff internal node m
// extends to the right; otherwise it is leaf m.
template <class K>
GI_HD uint32_t chain_top(const KeySrc<K>& k, int m) {
    if (m < k.n - 1 && delta(k, m, m + 1) > delta(k, m, m - 1)) return (uint32_t)m;
    return (uint32_t)(k.n - 1 + m);
}
// the same for m >= 1 from the keys m - 1, m and m + 1 already at hand
// (kp1 is used only when m < n - 1)
template <class K>
GI_HD uint32_t chain_top_of(int n, int m, K km1, K km, K kp1) {
    if (m < n - 1 && delta_keys(km, m, kp1, m + 1) > delta_keys(km, m, km1, m - 1)) return (uint32_t)m;
    return (uint32_t)(n - 1 + m);
}

// Children and rope of internal node i from its range end j and split gamma.
GI_HD void karras_children(int n, int lo_i, int hi_i, int gamma, uint32_t& left, uint32_t& right) {
    left = lo_i == gamma ? (uint32_t)(n - 1 + gamma) : (uint32_t)gamma;
    right = hi_i == gamma + 1 ? (uint32_t)(n - 1 + gamma + 1) : (uint32_t)(gamma + 1);
}
template <class K>
GI_HD void karras_links(const KeySrc<K>& KS, int i, int j, int gamma, int& lo_i, int& hi_i, uint32_t& left,
                        uint32_t& right, uint32_t& rope) {
    const int n = KS.n;
    lo_i = i < j ? i : j;
    hi_i = i < j ? j : i;
    karras_children(n, lo_i, hi_i, gamma, left, right);
    rope = hi_i == n - 1 ? kSentinel : chain_top(KS, hi_i + 1);
}

// ------------------------------------------------- in-chunk nodes by table
//
// For sorted keys the common prefix of keys i < j is the least of the
// adjacent ones between them: delta(i, j) = min A[i .. j - 1] with A[k] =
// delta(k, k + 1). A chunk [cs, ce) keeps A[cs - 1 .. ce - 1] at local
// positions q = 0 .. ce - cs (A[k] at q = k - cs + 1; -1 before the first
// key and from the last key on) and its minima over windows of 2^L
// positions, M[L][q] = min A-local[q .. q + 2^L - 1] (entries whose window
// passes the end are never read). Karras' two searches then become binary
// lifting over M: the range end is the first adjacent prefix at or below
// delta(i, i - d) on the node's side, and the split the position of the
// least adjacent prefix inside the range (unique for distinct composite
// keys). Prefix lengths fit a byte (-1 .. 96 for 64-bit keys).

template <int C>
struct DeltaLevels {
    static constexpr int kLevels = C == 64 ? 7 : (C == 128 ? 8 : (C == 256 ? 9 : (C == 512 ? 10 : 11)));
};

// Internal node i of the chunk from the table M (rows of C + 1 entries):
// returns false when its range leaves the chunk (a top node, for
// karras_top_node16); else its range, children and rope, those of Karras'
// searches.
template <int C, class K>
GI_HD bool karras_table_node(const KeySrc<K>& KS, const int8_t (*M)[C + 1], int i, int& lo_i, int& hi_i,
                             uint32_t& left, uint32_t& right, uint32_t& rope) {
    constexpr int top_level = DeltaLevels<C>::kLevels - 1;
    const int cs = KS.cs, qmax = KS.ce - cs;  // local position of A[ce - 1]
    const int q0 = i - cs + 1;                // A[i]; A[i - 1] at q0 - 1
    const int ai = M[0][q0], aim = M[0][q0 - 1];
    const int d = ai > aim ? 1 : -1;
    const int dmin = d > 0 ? aim : ai;
    int p = q0, dn = 127, j;
    if (d > 0) {  // first A[k], k >= i, at or below dmin: the range is [i, k]
        for (int L = top_level; L >= 0; --L) {
            const int w = 1 << L;
            if (p + w - 1 <= qmax) {
                const int v = M[L][p];
                if (v > dmin) {
                    dn = v < dn ? v : dn;
                    p += w;
                }
            }
        }
        if (p > qmax) return false;
        j = cs - 1 + p;
    } else {  // last A[k], k < i, at or below dmin: the range is [k + 1, i]
        for (int L = top_level; L >= 0; --L) {
            const int w = 1 << L;
            if (p - w >= 0) {
                const int v = M[L][p - w];
                if (v > dmin) {
                    dn = v < dn ? v : dn;
                    p -= w;
                }
            }
        }
        if (p == 0) return false;
        j = cs + p - 1;
    }
    // dn = delta(i, j), the least adjacent prefix of the range; the split
    // is where it sits (first A at or below it from the range's start)
    const int lo = i < j ? i : j;
    p = lo - cs + 1;
    for (int L = top_level; L >= 0; --L) {
        const int w = 1 << L;
        if (p + w - 1 <= qmax && M[L][p] > dn) p += w;
    }
    karras_links(KS, i, j, cs - 1 + p, lo_i, hi_i, left, right, rope);
    return true;
}

// ------------------------------------------------ top nodes, 16-way search
//
// A top node's range leaves its chunk, so its two searches run over the
// whole key array. They take 16 probes per round (the 16 lanes of a
// quarter-wave on the GPU, one key load each), which cuts the candidate
// interval 16-fold per round: a few dependent rounds instead of a binary
// search's ~40 dependent loads. The range search and the split search are
// separate calls, so that a kernel can issue the loads that need only the
// range (the rope's keys, the box windows) before the split search.

// probe k (0 .. 15) of a round over [lo, hi], lo < hi: nondecreasing in k,
// the last one at hi
GI_HD int probe16(int lo, int hi, int k) { return lo + (int)(((int64_t)(hi - lo) * (k + 1) + 15) / 16); }

// the interval left after a round in which the first c probes held (a
// monotone predicate, true at lo): from the last probe that held to just
// before the first that did not
GI_HD void narrow16(int& lo, int& hi, int c) {
    const int l2 = c ? probe16(lo, hi, c - 1) : lo;
    const int h2 = c < 16 ? probe16(lo, hi, c) - 1 : hi;
    lo = l2;
    hi = h2;
}

// probe k of a range search's first round: exponential steps lo + 4^k - 1
// (at most hi), since most ranges that leave a chunk end near its edge
GI_HD int probe_exp16(int lo, int hi, int k) {
    const int64_t x = (int64_t)lo + ((int64_t)1 << (2 * k)) - 1;
    return x < hi ? (int)x : hi;
}

// Top node i (key ki) whose range runs from i in direction d past its
// chunk's edge, at least lmin leaves away (delta(i, i + lmin d) > dmin =
// delta(i, i - d)): the range's length l, Karras' largest l with
// delta(i, i + l d) > dmin. count(pos, pred) returns how many of a round's
// 16 probes pos(0 .. 15) satisfy pred (a prefix of them: pos and pred are
// monotone).
template <class K, class Count>
GI_HD int karras_top_range16(const KeySrc<K>& KS, int i, K ki, int d, int dmin, int lmin, Count count) {
    int lo = lmin, hi = d > 0 ? KS.n - 1 - i : i;
    auto above = [&](int l) {
        const int p = i + l * d;
        return delta_keys(ki, i, KS.at(p), p) > dmin;
    };
    if (lo < hi) {  // probe 0 is lmin, which holds
        const int a = lo, b = hi;
        const int c = count([&](int k) { return probe_exp16(a, b, k); }, above);
        lo = probe_exp16(a, b, c - 1);
        hi = c < 16 ? probe_exp16(a, b, c) - 1 : b;
    }
    while (lo < hi) {
        const int a = lo, b = hi;
        narrow16(lo, hi, count([&](int k) { return probe16(a, b, k); }, above));
    }
    return lo;
}

// Its split: Karras' largest s in [0, l - 1] with delta(i, i + s d) > dn =
// delta(i, i + l d); returns the split position gamma.
template <class K, class Count>
GI_HD int karras_top_split16(const KeySrc<K>& KS, int i, K ki, int d, int l, int dn, Count count) {
    auto above = [&](int s) {
        const int p = i + s * d;
        return delta_keys(ki, i, KS.at(p), p) > dn;
    };
    int lo = 0, hi = l - 1;
    while (lo < hi) {
        const int a = lo, b = hi;
        narrow16(lo, hi, count([&](int k) { return probe16(a, b, k); }, above));
    }
    return i + lo * d + (d < 0 ? -1 : 0);
}

// Both searches and the links: range, children and rope of the top node.
template <class K, class Count>
GI_HD void karras_top_node16(const KeySrc<K>& KS, int i, K ki, int d, int dmin, int lmin, Count count, int& lo_i,
                             int& hi_i, uint32_t& left, uint32_t& right, uint32_t& rope) {
    const int l = karras_top_range16(KS, i, ki, d, dmin, lmin, count);
    const int j = i + l * d;
    karras_links(KS, i, j, karras_top_split16(KS, i, ki, d, l, delta_keys(ki, i, KS.at(j), j), count), lo_i, hi_i,
                 left, right, rope);
}

}  // namespace gi
