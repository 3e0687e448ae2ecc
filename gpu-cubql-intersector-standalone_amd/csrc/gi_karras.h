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

// Children and rope of internal node i from its range end j and split gamma.
template <class K>
GI_HD void karras_links(const KeySrc<K>& KS, int i, int j, int gamma, int& lo_i, int& hi_i, uint32_t& left,
                        uint32_t& right, uint32_t& rope) {
    const int n = KS.n;
    lo_i = i < j ? i : j;
    hi_i = i < j ? j : i;
    left = lo_i == gamma ? (uint32_t)(n - 1 + gamma) : (uint32_t)gamma;
    right = hi_i == gamma + 1 ? (uint32_t)(n - 1 + gamma + 1) : (uint32_t)(gamma + 1);
    rope = hi_i == n - 1 ? kSentinel : chain_top(KS, hi_i + 1);
}

// Karras' two searches from leaf i in direction d, both "the largest l in
// [0, L] with delta(i, i + l d) > thr" (true at l = 0, and monotone: sorted
// keys). One code path for every lane and both directions, so the probe
// chains of a wave's lanes overlap. The search first runs over every
// 2^sh-th key (sh > 0: the sort's key samples samp[q] = key at q << sh; sh
// = 0: the keys themselves), by galloping (the range end, usually near i) or
// by bisection (the split, inside a known range), and then bisects the keys
// inside the one stride left.
template <class K>
GI_HD int karras_extent(const KeySrc<K>& KS, const K* __restrict__ samp, int sh, int i, K ki, int d, int L, int thr,
                        bool gallop) {
    const int S = 1 << sh;
    const int q0 = d > 0 ? i >> sh : (i + S - 1) >> sh;  // stride point at or behind i
    const int e = i + L * d;
    const int kmax = d > 0 ? (e >> sh) - q0 : q0 - ((e + S - 1) >> sh);  // stride points past i, up to e
    auto above = [&](int k) {
        const int q = q0 + k * d;
        return delta_keys(ki, i, sh ? samp[q] : KS.at(q), q << sh) > thr;
    };
    int lo = 0, hi = kmax;
    if (gallop) {
        int g = 1;
        while (g <= kmax && above(g)) g <<= 1;
        lo = g >> 1;
        hi = g - 1 < kmax ? g - 1 : kmax;
    }
    while (lo < hi) {
        const int m = (lo + hi + 1) >> 1;
        if (above(m)) lo = m;
        else hi = m - 1;
    }
    // the answer lies between stride point lo and the next one
    int llo = lo > 0 ? (q0 + lo * d) * S - i : 0;
    llo = llo < 0 ? -llo : llo;
    int next = (q0 + d) * S - i;
    next = (next < 0 ? -next : next) - 1;
    int lhi = lo > 0 ? llo + S - 1 : next;
    lhi = lhi < L ? lhi : L;
    while (llo < lhi) {
        const int m = (llo + lhi + 1) >> 1;
        const int p = i + m * d;
        if (delta_keys(ki, i, KS.at(p), p) > thr) llo = m;
        else lhi = m - 1;
    }
    return llo;
}

// Internal node i of a build chunk (the keys [KS.cs, KS.ce) of KS in LDS, ki
// = key i): range, children and rope. Returns whether it is a "top" node,
// one whose range leaves the chunk: one probe past the chunk edge tells
// (sorted keys: the common prefix only shrinks with distance). An in-chunk
// node searches the chunk's keys up to its edge; a top node the key samples
// (samp: every 2^samp_shift-th sorted key; null: none) and then the keys of
// one stride.
template <class K>
GI_HD bool karras_chunk_node(const KeySrc<K>& KS, const K* __restrict__ samp, int samp_shift, int i, K ki, int& lo_i,
                             int& hi_i, uint32_t& left, uint32_t& right, uint32_t& rope) {
    const int d = delta(KS, i, i + 1) > delta(KS, i, i - 1) ? 1 : -1;
    const int dmin = delta(KS, i, i - d);
    const int edge = d > 0 ? KS.ce : KS.cs - 1;
    const bool top = edge >= 0 && edge < KS.n && delta(KS, i, edge) > dmin;
    const int sh = top && samp != nullptr ? samp_shift : 0;
    const int reach = top ? (d > 0 ? KS.n - 1 - i : i) : (d > 0 ? KS.ce - 1 - i : i - KS.cs);
    const int l = karras_extent(KS, samp, sh, i, ki, d, reach, dmin, true);
    const int j = i + l * d;
    const int s = karras_extent(KS, samp, sh, i, ki, d, l - 1, delta_keys(ki, i, KS.at(j), j), false);
    karras_links(KS, i, j, i + s * d + (d < 0 ? -1 : 0), lo_i, hi_i, left, right, rope);
    return top;
}

}  // namespace gi
