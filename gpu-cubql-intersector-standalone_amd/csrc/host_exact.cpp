// host_exact.cpp -- the host residual stage (see host_exact.h).
//
// Every coordinate of a pair is a dyadic rational, so one common power of two
// 2^e turns the 18 coordinates into integers; the exact algorithm of
// exact_tritri.h then runs on arbitrary-precision integers (BigNum below).
#include "host_exact.h"

#include <math.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <thread>

#include "exact_tritri.h"

namespace gi {

namespace {

typedef unsigned __int128 u128h;

struct Overflow {};

// Sign-magnitude integer with up to kLimbs 64-bit limbs.
template <int kLimbs>
struct BigNum {
    int sign = 0;  // -1, 0, +1
    int n = 0;
    uint64_t d[kLimbs];

    BigNum() {}
    BigNum(int64_t v) {
        if (v == 0) return;
        sign = v < 0 ? -1 : 1;
        d[0] = v < 0 ? (uint64_t)0 - (uint64_t)v : (uint64_t)v;
        n = 1;
    }
    void norm() {
        while (n > 0 && d[n - 1] == 0) --n;
        if (n == 0) sign = 0;
    }
    static int cmpMag(const BigNum& a, const BigNum& b) {
        if (a.n != b.n) return a.n < b.n ? -1 : 1;
        for (int i = a.n - 1; i >= 0; --i)
            if (a.d[i] != b.d[i]) return a.d[i] < b.d[i] ? -1 : 1;
        return 0;
    }
    static BigNum addMag(const BigNum& a, const BigNum& b, int sign) {
        const BigNum& x = a.n >= b.n ? a : b;
        const BigNum& y = a.n >= b.n ? b : a;
        BigNum r;
        uint64_t carry = 0;
        for (int i = 0; i < x.n; ++i) {
            u128h s = (u128h)x.d[i] + (i < y.n ? y.d[i] : 0) + carry;
            r.d[i] = (uint64_t)s;
            carry = (uint64_t)(s >> 64);
        }
        r.n = x.n;
        if (carry) {
            if (r.n >= kLimbs) throw Overflow();
            r.d[r.n++] = carry;
        }
        r.sign = sign;
        r.norm();
        return r;
    }
    // |a| - |b| with |a| >= |b|
    static BigNum subMag(const BigNum& a, const BigNum& b, int sign) {
        BigNum r;
        uint64_t borrow = 0;
        for (int i = 0; i < a.n; ++i) {
            u128h bi = (u128h)(i < b.n ? b.d[i] : 0) + borrow;
            u128h ai = a.d[i];
            if (ai >= bi) { r.d[i] = (uint64_t)(ai - bi); borrow = 0; }
            else { r.d[i] = (uint64_t)(((u128h)1 << 64) + ai - bi); borrow = 1; }
        }
        r.n = a.n;
        r.sign = sign;
        r.norm();
        return r;
    }
    friend BigNum operator+(const BigNum& a, const BigNum& b) {
        if (b.sign == 0) return a;
        if (a.sign == 0) return b;
        if (a.sign == b.sign) return addMag(a, b, a.sign);
        int c = cmpMag(a, b);
        if (c == 0) return BigNum();
        return c > 0 ? subMag(a, b, a.sign) : subMag(b, a, b.sign);
    }
    friend BigNum operator-(const BigNum& a) {
        BigNum r = a;
        r.sign = -r.sign;
        return r;
    }
    friend BigNum operator-(const BigNum& a, const BigNum& b) { return a + (-b); }
    friend BigNum operator*(const BigNum& a, const BigNum& b) {
        BigNum r;
        if (a.sign == 0 || b.sign == 0) return r;
        if (a.n + b.n > kLimbs) throw Overflow();
        memset(r.d, 0, sizeof(uint64_t) * (size_t)(a.n + b.n));
        for (int i = 0; i < a.n; ++i) {
            uint64_t carry = 0;
            for (int j = 0; j < b.n; ++j) {
                u128h p = (u128h)a.d[i] * b.d[j] + r.d[i + j] + carry;
                r.d[i + j] = (uint64_t)p;
                carry = (uint64_t)(p >> 64);
            }
            r.d[i + b.n] = carry;
        }
        r.n = a.n + b.n;
        r.sign = a.sign * b.sign;
        r.norm();
        return r;
    }
    static BigNum shl(const BigNum& a, int k) {
        if (a.sign == 0 || k == 0) return a;
        int limbs = k / 64, bits = k % 64;
        if (a.n + limbs + 1 > kLimbs) throw Overflow();
        BigNum r;
        memset(r.d, 0, sizeof(uint64_t) * (size_t)(a.n + limbs + 1));
        for (int i = 0; i < a.n; ++i) {
            r.d[i + limbs] |= a.d[i] << bits;
            if (bits) r.d[i + limbs + 1] |= a.d[i] >> (64 - bits);
        }
        r.n = a.n + limbs + 1;
        r.sign = a.sign;
        r.norm();
        return r;
    }
    int bits() const { return n == 0 ? 0 : 64 * (n - 1) + (64 - __builtin_clzll(d[n - 1])); }
    int exportLimbs() const {
        int b = bits() + 1;
        int l = (b + 63) / 64;
        return l < 1 ? 1 : l;
    }
    void exportTo(uint64_t* out, int limbs) const {
        for (int i = 0; i < limbs; ++i) out[i] = i < n ? d[i] : 0;
        if (sign < 0) {
            uint64_t carry = 1;
            for (int i = 0; i < limbs; ++i) {
                u128h s = (u128h)(~out[i]) + carry;
                out[i] = (uint64_t)s;
                carry = (uint64_t)(s >> 64);
            }
        }
    }
};

template <int K>
struct BigPolicy {
    typedef BigNum<K> N;
    typedef N C;
    typedef N B;
    typedef N H;
    static N toB(const N& x) { return x; }
    static N mulCC(const N& a, const N& b) { return a * b; }
    static N mulBC(const N& a, const N& b) { return a * b; }
    static N mulBB(const N& a, const N& b) { return a * b; }
    static int cmpH(const N& a, const N& b) { return (a - b).sign; }
    static int sgnC(const N& x) { return x.sign; }
    static int sgnB(const N& x) { return x.sign; }
    static int cmpAbsC(const N& a, const N& b) { return N::cmpMag(a, b); }
    static int cmpAbsB(const N& a, const N& b) { return N::cmpMag(a, b); }
    static const N& pick3(int k, const N& a, const N& b, const N& c) { return k == 0 ? a : (k == 1 ? b : c); }
};

// Capacities: 64 limbs serve every pair whose coordinates span up to ~560
// bits of exponent; the largest products of any two finite doubles' pair
// (coordinates of up to 2,098 bits: 2^971 down to 2^-1074) need 14,701 bits,
// within 256 limbs.
constexpr int kLimbsFast = 64;
constexpr int kLimbsAll = 256;

// lowest-set-bit exponent of a non-zero finite double and its odd mantissa
void dyadic(double x, int64_t& odd, int& elow) {
    int e;
    double m = frexp(x, &e);
    int64_t M = (int64_t)ldexp(m, 53);
    uint64_t um = M < 0 ? (uint64_t)0 - (uint64_t)M : (uint64_t)M;
    int t = __builtin_ctzll(um);
    odd = M / ((int64_t)1 << t);
    elow = e - 53 + t;
}

// One pair on K-limb integers; false with *fits = false when they overflow.
template <int K>
bool exact_pair_k(const double* a, const double* b, HostHit& out, bool* fits, int* npred) {
    typedef BigNum<K> BigN;
    typedef ExactTT<BigPolicy<K>> BigExact;
    *fits = true;
    out.nseg = 0;
    out.coplanar = out.degenerate = 0;
    const double* src[2] = {a, b};
    int emin = 0;
    for (int s = 0; s < 2; ++s)
        for (int i = 0; i < 9; ++i) {
            double x = src[s][i];
            if (x == 0.0) continue;
            int64_t odd;
            int el;
            dyadic(x, odd, el);
            if (el < emin) emin = el;
        }
    try {
        typename BigExact::V3 A[3], B[3];
        for (int s = 0; s < 2; ++s)
            for (int i = 0; i < 9; ++i) {
                double x = src[s][i];
                BigN& dst = (s == 0 ? A : B)[i / 3].v[i % 3];
                if (x == 0.0) { dst = BigN(); continue; }
                int64_t odd;
                int el;
                dyadic(x, odd, el);
                dst = BigN::shl(BigN(odd), el - emin);
            }
        typename BigExact::Result r;
        BigExact::intersect(A, B, r);
        if (npred) *npred = r.npred;
        if (!r.hit) return false;
        out.coplanar = r.coplanar;
        out.degenerate = r.degenerate;
        out.nseg = r.nseg;
        out.exp2 = emin;
        typename BigExact::HP P[GI_SEG_MAX][2];
        int limbs = 1;
        for (int i = 0; i < r.nseg; ++i) {
            BigExact::segment(A, B, r, i, P[i][0], P[i][1]);
            for (int e = 0; e < 2; ++e) {
                for (int k = 0; k < 3; ++k) limbs = std::max(limbs, P[i][e].h[k].exportLimbs());
                limbs = std::max(limbs, P[i][e].w.exportLimbs());
            }
        }
        out.limbs = limbs;
        out.data.assign((size_t)r.nseg * 8 * limbs, 0);
        for (int i = 0; i < r.nseg; ++i)
            for (int e = 0; e < 2; ++e) {
                for (int k = 0; k < 3; ++k) P[i][e].h[k].exportTo(&out.data[((size_t)i * 8 + 4 * e + k) * limbs], limbs);
                P[i][e].w.exportTo(&out.data[((size_t)i * 8 + 4 * e + 3) * limbs], limbs);
            }
        return true;
    } catch (const Overflow&) {
        *fits = false;
        return false;
    }
}

}  // namespace

bool host_exact_pair(const double* a, const double* b, HostHit& out, bool* ok, int* npred) {
    *ok = true;
    for (int i = 0; i < 9; ++i)
        if (!isfinite(a[i]) || !isfinite(b[i])) {
            *ok = false;
            return false;
        }
    bool fits;
    bool hit = exact_pair_k<kLimbsFast>(a, b, out, &fits, npred);
    if (fits) return hit;
    hit = exact_pair_k<kLimbsAll>(a, b, out, &fits, npred);
    if (!fits) *ok = false;  // unreachable for finite doubles
    return hit;
}

int host_exact_batch(const double* vA, const uint32_t* triA, const double* vB, const uint32_t* triB,
                     const uint32_t* list, uint32_t n, int nthreads, std::vector<HostHit>& hits, uint64_t* yellow_yes,
                     uint64_t* predicates) {
    if (nthreads < 1) nthreads = 1;
    std::vector<std::vector<HostHit>> local((size_t)nthreads);
    std::atomic<uint64_t> yes{0}, npred{0};
    std::atomic<int> unsound{0}, nonfinite{0};
    std::atomic<uint32_t> next{0};
    auto work = [&](int t) {
        const uint32_t chunk = 64;
        for (;;) {
            uint32_t start = next.fetch_add(chunk);
            if (start >= n) break;
            uint32_t end = std::min(n, start + chunk);
            for (uint32_t i = start; i < end; ++i) {
                uint32_t a = list[2 * (size_t)i], bw = list[2 * (size_t)i + 1];
                uint32_t b = bw & ((1u << 30) - 1u), status = bw >> 30;
                double pa[9], pb[9];
                for (int v = 0; v < 3; ++v)
                    for (int k = 0; k < 3; ++k) {
                        pa[3 * v + k] = vA[4 * (size_t)triA[3 * (size_t)a + v] + k];
                        pb[3 * v + k] = vB[4 * (size_t)triB[3 * (size_t)b + v] + k];
                    }
                HostHit h;
                bool ok;
                int np = 0;
                bool hit = host_exact_pair(pa, pb, h, &ok, &np);
                npred += (uint64_t)np;
                if (!ok) { nonfinite++; continue; }
                if (status == 1 && !hit) { unsound++; continue; }  // certified GREEN must intersect
                if (!hit) continue;
                if (status == 2) yes++;
                h.a = a;
                h.b = b;
                local[(size_t)t].push_back(std::move(h));
            }
        }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nthreads; ++t) th.emplace_back(work, t);
    work(0);
    for (auto& x : th) x.join();
    hits.clear();
    for (auto& v : local)
        for (auto& h : v) hits.push_back(std::move(h));
    std::sort(hits.begin(), hits.end(), [](const HostHit& x, const HostHit& y) {
        return x.a != y.a ? x.a < y.a : x.b < y.b;
    });
    if (yellow_yes) *yellow_yes = yes.load();
    if (predicates) *predicates = npred.load();
    if (nonfinite.load()) return kHostNonFinite;
    return unsound.load() ? kHostUnsound : 0;
}

}  // namespace gi
