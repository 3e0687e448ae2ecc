// exact_wide.h -- integer frame of one triangle pair given as doubles, for
// the GPU exact stage on meshes off the integer grid.
//
// Every finite double is an odd integer times a power of two. With e the
// lowest such exponent among the pair's 18 coordinates, each coordinate is
// the integer X = x / 2^e exactly. When every |X| < 2^61 the pair runs in
// that frame; otherwise the frame is translated by the integer copy X0 of
// vertex 0 of triangle A (coordinates X - X0, exact in int128 because both
// stay below 2^126), which removes the common leading bits of a pair far
// from the origin. The exact intersection is invariant under this
// translation and scaling, and a constructed point h/w of the frame is the
// point ((h + X0 w) / w) 2^e of the pair. Pairs whose coordinates do not fit
// (non-finite values, spreads of more than 61 bits after translation, or an
// absolute endpoint beyond 255 bits) are left to the host residual stage.
//
// Replaces the reference's anchor-relative conversion in its exact stage
// (src/GPU_exact_predicates/ExactBigIntPredicate.cu:381-391), which subtracts
// the anchor in double arithmetic and so is exact only when that subtraction
// is; here the subtraction is done on the exact integers.
#pragma once

#include "exact_tritri.h"

namespace gi {

typedef ExactTT<WidePolicy> WideExact;

constexpr int kWideCoordBits = 61;  // |coordinate| < 2^61 in the frame

struct WideFrame {
    int exp2;     // the frame's unit is 2^exp2
    i128 x0[3];   // translation in frame units (zero when none is needed)
    int shifted;  // x0 is not zero
};

// Odd mantissa and exponent of a finite double (m = 0 for zero). Returns
// false for infinities and NaNs.
GI_HD bool dyadic_parts(double x, int64_t& m, int& q) {
    const uint64_t b = __builtin_bit_cast(uint64_t, x);
    const int ef = (int)((b >> 52) & 0x7FF);
    if (ef == 0x7FF) return false;
    uint64_t mant = b & ((1ull << 52) - 1);
    int e = ef == 0 ? -1074 : ef - 1075;
    if (ef != 0) mant |= 1ull << 52;
    if (mant == 0) {
        m = 0;
        q = 0;
        return true;
    }
    const int t = __builtin_ctzll(mant);
    mant >>= t;
    e += t;
    m = (b >> 63) ? -(int64_t)mant : (int64_t)mant;
    q = e;
    return true;
}

GI_HD int bits128(i128 x) {
    const u128 u = uabs128(x);
    const uint64_t hi = (uint64_t)(u >> 64), lo = (uint64_t)u;
    return hi ? 128 - __builtin_clzll(hi) : (lo ? 64 - __builtin_clzll(lo) : 0);
}

// The pair's 18 doubles (a, b: 9 each, vertex-major) in the integer frame.
// Returns false when the pair does not fit.
GI_HD bool wide_frame(const double* a, const double* b, WideExact::V3* A, WideExact::V3* B, WideFrame& f) {
    int64_t m[18];
    int q[18];
    int emin = 0x7FFFFFFF;
#pragma unroll
    for (int i = 0; i < 18; ++i) {
        if (!dyadic_parts(i < 9 ? a[i] : b[i - 9], m[i], q[i])) return false;
        if (m[i] != 0 && q[i] < emin) emin = q[i];
    }
    if (emin == 0x7FFFFFFF) emin = 0;  // all zero
    i128 X[18];
    int kmax = 0;
#pragma unroll
    for (int i = 0; i < 18; ++i) {
        if (m[i] == 0) {
            X[i] = 0;
            continue;
        }
        const uint64_t um = m[i] < 0 ? (uint64_t)0 - (uint64_t)m[i] : (uint64_t)m[i];
        const int top = 63 - __builtin_clzll(um) + q[i] - emin;  // position of the top bit of |X|
        if (top > 125) return false;
        X[i] = (i128)m[i] * ((i128)1 << (q[i] - emin));
        kmax = top + 1 > kmax ? top + 1 : kmax;
    }
    f.exp2 = emin;
    f.shifted = kmax > kWideCoordBits;
#pragma unroll
    for (int k = 0; k < 3; ++k) f.x0[k] = f.shifted ? X[k] : (i128)0;
    int k1 = 0;
#pragma unroll
    for (int i = 0; i < 18; ++i) {
        const i128 d = X[i] - f.x0[i % 3];
        const int bt = bits128(d);
        k1 = bt > k1 ? bt : k1;
        WideExact::V3* T = i < 9 ? A : B;
        T[(i % 9) / 3].v[i % 3] = d;
    }
    return k1 <= kWideCoordBits;
}

// A constructed point of the frame as the pair's point: h_k + x0_k w over w,
// unit 2^exp2. Returns false when a numerator would leave 256 bits.
GI_HD bool wide_absolute(const WideFrame& f, WideExact::HP& p) {
    if (!f.shifted) return true;
    const int bw = w256_bits(p.w);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int bx = bits128(f.x0[k]);
        if (bx == 0) continue;
        if (bx + bw > 253 || w256_bits(p.h[k]) > 253) return false;
        p.h[k] = p.h[k] + w256_mul(p.w, f.x0[k]);
    }
    return true;
}

}  // namespace gi
