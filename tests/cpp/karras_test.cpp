// Host check of the BVH topology that k_build_chunks computes per node
// (gi_karras.h, compiled here for the CPU): for every internal node of every
// chunk, at several chunk sizes and key-sample strides, the range, children
// and rope must equal those of a plain Karras build written out below, and
// the whole tree must be well formed: each of the 2n - 1 nodes reached once
// from the root, children splitting their parent's leaf range, every rope
// the next node in pre-order after the subtree. Prints one line per case and
// exits non-zero on the first failure.
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

int fail(const char* what, const char* set, int chunk, int sh, long at) {
    std::printf("FAIL %s: set %s chunk %d shift %d at %ld\n", what, set, chunk, sh, at);
    return 1;
}

// the kernel's per-node routine over all chunks of one chunk size
template <class K>
int check(const char* set, const std::vector<K>& keys, int chunk, int sh, long& top_nodes) {
    const int n = (int)keys.size();
    if (n < 2) return 0;
    std::vector<K> samp;
    if (sh > 0)
        for (int p = 0; p < n; p += 1 << sh) samp.push_back(keys[p]);
    Plain<K> P{keys, n};
    Tree T;
    T.left.resize(n - 1); T.right.resize(n - 1); T.rope.resize(2 * n - 1);
    T.lo.resize(n - 1); T.hi.resize(n - 1);
    for (int i = 0; i < n; ++i) {
        const int cs = i / chunk * chunk, ce = std::min(n, cs + chunk);
        const KeySrc<K> KS{keys.data(), keys.data() + cs, cs, ce, n};
        T.rope[n - 1 + i] = i == n - 1 ? kSentinel : chain_top(KS, i + 1);  // leaf rope, as the kernel writes it
        if (i == n - 1) break;
        int lo, hi;
        uint32_t l, r, rope;
        top_nodes += karras_chunk_node(KS, sh > 0 ? samp.data() : (const K*)nullptr, sh, i, keys[i], lo, hi, l, r, rope);
        int plo, phi, pg;
        P.node(i, plo, phi, pg);
        const uint32_t pl = plo == pg ? (uint32_t)(n - 1 + pg) : (uint32_t)pg;
        const uint32_t pr = phi == pg + 1 ? (uint32_t)(n - 1 + pg + 1) : (uint32_t)(pg + 1);
        if (lo != plo || hi != phi || l != pl || r != pr) return fail("node differs from plain Karras", set, chunk, sh, i);
        T.left[i] = l; T.right[i] = r; T.rope[i] = rope; T.lo[i] = lo; T.hi[i] = hi;
    }
    // structure: pre-order walk from the root (internal node 0)
    std::vector<uint32_t> order;
    std::vector<char> seen(2 * n - 1, 0);
    std::vector<uint32_t> stack{0u};
    while (!stack.empty()) {
        const uint32_t v = stack.back();
        stack.pop_back();
        if (v >= (uint32_t)(2 * n - 1) || seen[v]) return fail("node reached twice or out of range", set, chunk, sh, v);
        seen[v] = 1;
        order.push_back(v);
        if (v < (uint32_t)(n - 1)) {
            const uint32_t l = T.left[v], r = T.right[v];
            const int llo = l < (uint32_t)(n - 1) ? T.lo[l] : (int)(l - (n - 1));
            const int lhi = l < (uint32_t)(n - 1) ? T.hi[l] : (int)(l - (n - 1));
            const int rlo = r < (uint32_t)(n - 1) ? T.lo[r] : (int)(r - (n - 1));
            const int rhi = r < (uint32_t)(n - 1) ? T.hi[r] : (int)(r - (n - 1));
            if (llo != T.lo[v] || lhi + 1 != rlo || rhi != T.hi[v])
                return fail("children do not split the parent range", set, chunk, sh, v);
            stack.push_back(r);
            stack.push_back(l);
        }
    }
    if ((int)order.size() != 2 * n - 1) return fail("nodes not reached from the root", set, chunk, sh, (long)order.size());
    // rope of v = first node after v's subtree in pre-order
    std::vector<int> pos(2 * n - 1);
    for (int p = 0; p < 2 * n - 1; ++p) pos[order[p]] = p;
    for (int v = 0; v < 2 * n - 1; ++v) {
        const int subtree = v < n - 1 ? 2 * (T.hi[v] - T.lo[v] + 1) - 1 : 1;
        const int next = pos[v] + subtree;
        const uint32_t want = next < 2 * n - 1 ? order[next] : kSentinel;
        if (T.rope[v] != want) return fail("rope is not the pre-order successor", set, chunk, sh, v);
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
int run_set(const char* name, std::vector<K> keys, std::initializer_list<int> shifts) {
    std::sort(keys.begin(), keys.end());
    for (int chunk : {64, 256, 512})
        for (int sh : shifts) {
            long top = 0;
            if (check(name, keys, chunk, sh, top)) return 1;
            std::printf("ok %s n=%zu chunk=%d shift=%d top_nodes=%ld\n", name, keys.size(), chunk, sh, top);
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
        bad |= run_set("sphere_morton", k, {0, 6, 8, 10});
    }
    {
        std::vector<uint32_t> k(100003);
        for (auto& v : k) v = (uint32_t)(rng() & 0xffffffu);
        bad |= run_set("uniform", k, {0, 4, 8});
    }
    {  // long runs of equal keys: the index breaks the ties
        std::vector<uint32_t> k(70001);
        for (auto& v : k) v = (uint32_t)(rng() % 300);
        bad |= run_set("duplicates", k, {0, 5, 8});
    }
    {  // a dense cluster and a sparse tail
        std::vector<uint32_t> k;
        for (int i = 0; i < 60000; ++i) k.push_back((uint32_t)(rng() & 0x3ffu));
        for (int i = 0; i < 700; ++i) k.push_back((uint32_t)((1u << 23) + (rng() & 0x7fffffu)));
        bad |= run_set("clustered", k, {0, 8});
    }
    for (int n : {2, 3, 4, 5, 17, 64, 65, 4099}) {
        std::vector<uint32_t> k(n);
        for (auto& v : k) v = (uint32_t)(rng() & 0xffu);
        bad |= run_set("small", k, {0, 1, 2});
    }
    {  // batch keys as the library builds them: mesh id above 18 bits of Morton code, within 32 bits
        std::vector<uint32_t> k;
        for (uint32_t m = 0; m < 300; ++m) {
            const int nm = 1 + (int)(rng() % 900);
            for (int i = 0; i < nm; ++i) k.push_back((m << 18) | (uint32_t)(rng() & 0x3ffffu));
        }
        bad |= run_set("forest32", k, {0, 8});
    }
    {  // 64-bit forest keys: mesh id above 30 bits of Morton code, searched without samples
        std::vector<uint64_t> k;
        for (uint64_t m = 0; m < 40; ++m) {
            const int nm = 1 + (int)(rng() % 3000);
            for (int i = 0; i < nm; ++i) k.push_back((m << 30) | (rng() & 0x3fffffffu));
        }
        bad |= run_set("forest64", k, {0});
    }
    std::printf(bad ? "FAILED\n" : "ALL OK\n");
    return bad;
}
