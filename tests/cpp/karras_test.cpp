// Host check of the BVH topology that k_build_chunks and k_top_boxes compute
// per node (gi_karras.h, compiled here for the CPU): for every internal node
// of every chunk, at several chunk sizes, the range, children and rope must
// equal those of a plain Karras build written out below (in-chunk nodes by
// binary lifting over the chunk's adjacent-prefix table, top nodes by the
// 16-way searches over all keys), and the whole tree must be well formed:
// each of the 2n - 1 nodes reached once from the root, children splitting
// their parent's leaf range, every rope the next node in pre-order after the
// subtree. Prints one line per case and exits non-zero on the first failure.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

#include "gi_karras.h"

using namespace gi;

namespace {

// ---- plain Karras (2012), one node at a time, all keys in one array
template <class K>
struct Plain {
    const std::vector<K>& k;
    int n;
    int delta(int i, int j) const {
        if (j < 0 || j >= n) return -1;
        if (k[i] != k[j]) return clz_key((K)(k[i] ^ k[j]));
        return 8 * (int)sizeof(K) + clz_key((uint32_t)(i ^ j));
    }
    void node(int i, int& lo, int& hi, int& gamma) const {
        const int d = delta(i, i + 1) > delta(i, i - 1) ? 1 : -1;
        const int dmin = delta(i, i - d);
        int lmax = 2;
        while (delta(i, i + lmax * d) > dmin) lmax *= 2;
        int l = 0;
        for (int t = lmax / 2; t >= 1; t /= 2)
            if (delta(i, i + (l + t) * d) > dmin) l += t;
        const int j = i + l * d, dn = delta(i, j);
        int s = 0, t = l;
        do {
            t = (t + 1) / 2;
            if (delta(i, i + (s + t) * d) > dn) s += t;
        } while (t > 1);
        gamma = i + s * d + (d < 0 ? -1 : 0);
        lo = std::min(i, j);
        hi = std::max(i, j);
    }
};

struct Tree {
    std::vector<uint32_t> left, right, rope;  // internal nodes 0..n-2; rope also for leaves (index n-1+k)
    std::vector<int> lo, hi;
};

int fail(const char* what, const char* set, int chunk, long at) {
    std::printf("FAIL %s: set %s chunk %d at %ld\n", what, set, chunk, at);
    return 1;
}

// Adjacent-prefix table of the chunk [cs, ce) as k_build_chunks builds it,
// written here with plain loops: A[k] = delta(k, k + 1) at q = k - cs + 1,
// -1 outside [0, n - 2]; M[L][q] = min over q .. q + 2^L - 1.
template <int C, class K>
void delta_table(const std::vector<K>& keys, int cs, int8_t (*M)[C + 1]) {
    const int n = (int)keys.size(), ce = std::min(n, cs + C);
    for (int q = 0; q <= C; ++q) {
        const int k = cs - 1 + q;
        int v = -1;
        if (k >= 0 && k + 1 < n && k <= ce - 1) {
            v = keys[k] != keys[k + 1] ? clz_key((K)(keys[k] ^ keys[k + 1]))
                                       : 8 * (int)sizeof(K) + clz_key((uint32_t)(k ^ (k + 1)));
        }
        M[0][q] = (int8_t)v;
    }
    for (int L = 1; L < DeltaLevels<C>::kLevels; ++L)
        for (int q = 0; q <= C; ++q)
            M[L][q] = q + (1 << L) - 1 <= C ? std::min(M[L - 1][q], M[L - 1][q + (1 << (L - 1))]) : (int8_t)99;
}

// The kernels' routines over all chunks of one chunk size C: in-chunk nodes
// from the chunk's adjacent-prefix table (k_build_chunks), top nodes from
// the 16-way searches over all keys (k_top_boxes, its 16 lanes' probes
// counted here one by one). Each node must equal plain Karras; the tree must
// be well formed.
template <int C, class K>
int check(const char* set, const std::vector<K>& keys, long& top_nodes) {
    const int n = (int)keys.size();
    if (n < 2) return 0;
    Plain<K> P{keys, n};
    Tree T;
    T.left.resize(n - 1); T.right.resize(n - 1); T.rope.resize(2 * n - 1);
    T.lo.resize(n - 1); T.hi.resize(n - 1);
    std::vector<int8_t> buf((size_t)DeltaLevels<C>::kLevels * (C + 1));
    int8_t (*M)[C + 1] = reinterpret_cast<int8_t (*)[C + 1]>(buf.data());
    const KeySrc<K> G{keys.data(), nullptr, 0, 0, n};  // all keys, read as the top-node kernel reads them
    auto count = [](auto pos, auto pred) {
        int c = 0;
        for (int k = 0; k < 16; ++k) c += pred(pos(k)) ? 1 : 0;
        return c;
    };
    for (int cs = 0; cs < n; cs += C) {
        const int ce = std::min(n, cs + C);
        delta_table<C, K>(keys, cs, M);
        const KeySrc<K> KS{keys.data(), keys.data() + cs, cs, ce, n};
        for (int i = cs; i < ce; ++i) {
            T.rope[n - 1 + i] = i == n - 1 ? kSentinel : chain_top(KS, i + 1);  // leaf rope, as the kernel writes it
            if (i + 1 < n &&  // the top-node kernel's form, from three keys
                chain_top_of(n, i + 1, keys[i], keys[i + 1], i + 2 < n ? keys[i + 2] : (K)0) != T.rope[n - 1 + i])
                return fail("chain_top_of differs from chain_top", set, C, i);
            if (i == n - 1) break;
            int lo, hi;
            uint32_t l, r, rope;
            if (!karras_table_node<C, K>(KS, M, i, lo, hi, l, r, rope)) {
                ++top_nodes;
                const int dn = delta(G, i, i + 1), dp = delta(G, i, i - 1);
                const int d = dn > dp ? 1 : -1, dmin = d > 0 ? dp : dn;
                if (delta(G, i, d > 0 ? ce : cs - 1) <= dmin)
                    return fail("a node the table calls top ends inside its chunk", set, C, i);
                karras_top_node16(G, i, keys[i], d, dmin, d > 0 ? ce - i : i - cs + 1, count, lo, hi, l, r, rope);
            }
            int plo, phi, pg;
            P.node(i, plo, phi, pg);
            const uint32_t pl = plo == pg ? (uint32_t)(n - 1 + pg) : (uint32_t)pg;
            const uint32_t pr = phi == pg + 1 ? (uint32_t)(n - 1 + pg + 1) : (uint32_t)(pg + 1);
            if (lo != plo || hi != phi || l != pl || r != pr)
                return fail("node differs from plain Karras", set, C, i);
            T.left[i] = l; T.right[i] = r; T.rope[i] = rope; T.lo[i] = lo; T.hi[i] = hi;
        }
    }
    // structure: pre-order walk from the root (internal node 0)
    std::vector<uint32_t> order;
    std::vector<char> seen(2 * n - 1, 0);
    std::vector<uint32_t> stack{0u};
    while (!stack.empty()) {
        const uint32_t v = stack.back();
        stack.pop_back();
        if (v >= (uint32_t)(2 * n - 1) || seen[v]) return fail("node reached twice or out of range", set, C, v);
        seen[v] = 1;
        order.push_back(v);
        if (v < (uint32_t)(n - 1)) {
            const uint32_t l = T.left[v], r = T.right[v];
            const int llo = l < (uint32_t)(n - 1) ? T.lo[l] : (int)(l - (n - 1));
            const int lhi = l < (uint32_t)(n - 1) ? T.hi[l] : (int)(l - (n - 1));
            const int rlo = r < (uint32_t)(n - 1) ? T.lo[r] : (int)(r - (n - 1));
            const int rhi = r < (uint32_t)(n - 1) ? T.hi[r] : (int)(r - (n - 1));
            if (llo != T.lo[v] || lhi + 1 != rlo || rhi != T.hi[v])
                return fail("children do not split the parent range", set, C, v);
            stack.push_back(r);
            stack.push_back(l);
        }
    }
    if ((int)order.size() != 2 * n - 1) return fail("nodes not reached from the root", set, C, (long)order.size());
    // rope of v = first node after v's subtree in pre-order
    std::vector<int> pos(2 * n - 1);
    for (int p = 0; p < 2 * n - 1; ++p) pos[order[p]] = p;
    for (int v = 0; v < 2 * n - 1; ++v) {
        const int subtree = v < n - 1 ? 2 * (T.hi[v] - T.lo[v] + 1) - 1 : 1;
        const int next = pos[v] + subtree;
        const uint32_t want = next < 2 * n - 1 ? order[next] : kSentinel;
        if (T.rope[v] != want) return fail("rope is not the pre-order successor", set, C, v);
    }
    return 0;
}

uint32_t spread3(uint32_t x) {  // 10 bits -> every third bit
    x &= 0x3ffu;
    x = (x | (x << 16)) & 0x030000ffu;
    x = (x | (x << 8)) & 0x0300f00fu;
    x = (x | (x << 4)) & 0x030c30c3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

template <class K>
int run_set(const char* name, std::vector<K> keys) {
    std::sort(keys.begin(), keys.end());
    for (int chunk : {64, 256, 512}) {
        long top = 0;
        const int bad = chunk == 64    ? check<64, K>(name, keys, top)
                        : chunk == 256 ? check<256, K>(name, keys, top)
                                       : check<512, K>(name, keys, top);
        if (bad) return 1;
        std::printf("ok %s n=%zu chunk=%d top_nodes=%ld\n", name, keys.size(), chunk, top);
    }
    return 0;
}

}  // namespace

int main() {
    std::mt19937_64 rng(20261203);
    int bad = 0;
    {  // Morton codes (8 bits per axis) of points on a sphere: the BVH workload's kind of keys
        const int n = 1 << 18;
        std::normal_distribution<double> g;
        std::vector<uint32_t> k(n);
        for (int i = 0; i < n; ++i) {
            double x = g(rng), y = g(rng), z = g(rng), r = std::sqrt(x * x + y * y + z * z);
            auto q = [&](double c) { return (uint32_t)std::min(255.0, std::max(0.0, (c / r + 1.0) * 128.0)); };
            k[i] = (spread3(q(x)) << 2) | (spread3(q(y)) << 1) | spread3(q(z));
        }
        bad |= run_set("sphere_morton", k);
    }
    {
        std::vector<uint32_t> k(100003);
        for (auto& v : k) v = (uint32_t)(rng() & 0xffffffu);
        bad |= run_set("uniform", k);
    }
    {  // long runs of equal keys: the index breaks the ties
        std::vector<uint32_t> k(70001);
        for (auto& v : k) v = (uint32_t)(rng() % 300);
        bad |= run_set("duplicates", k);
    }
    {  // a dense cluster and a sparse tail
        std::vector<uint32_t> k;
        for (int i = 0; i < 60000; ++i) k.push_back((uint32_t)(rng() & 0x3ffu));
        for (int i = 0; i < 700; ++i) k.push_back((uint32_t)((1u << 23) + (rng() & 0x7fffffu)));
        bad |= run_set("clustered", k);
    }
    for (int n : {2, 3, 4, 5, 17, 64, 65, 4099}) {
        std::vector<uint32_t> k(n);
        for (auto& v : k) v = (uint32_t)(rng() & 0xffu);
        bad |= run_set("small", k);
    }
    {  // batch keys as the library builds them: mesh id above 18 bits of Morton code, within 32 bits
        std::vector<uint32_t> k;
        for (uint32_t m = 0; m < 300; ++m) {
            const int nm = 1 + (int)(rng() % 900);
            for (int i = 0; i < nm; ++i) k.push_back((m << 18) | (uint32_t)(rng() & 0x3ffffu));
        }
        bad |= run_set("forest32", k);
    }
    {  // 64-bit forest keys: mesh id above 30 bits of Morton code
        std::vector<uint64_t> k;
        for (uint64_t m = 0; m < 40; ++m) {
            const int nm = 1 + (int)(rng() % 3000);
            for (int i = 0; i < nm; ++i) k.push_back((m << 30) | (rng() & 0x3fffffffu));
        }
        bad |= run_set("forest64", k);
    }
    std::printf(bad ? "FAILED\n" : "ALL OK\n");
    return bad;
}
