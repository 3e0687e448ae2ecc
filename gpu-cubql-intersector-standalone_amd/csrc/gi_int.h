// gi_int.h -- fixed-width integer helpers shared by host and device code.
//
// The exact stage works on integer vertex coordinates with |x| <= 2^26
// (the "grid" the meshes are snapped to). With that bound:
//   coordinate differences          < 2^27   (int64)
//   cross products, 2D orientations < 2^56   (int64)
//   orient3d determinants           < 2^84   (int128)
//   homogeneous crossing points     < 2^111  (int128)
//   cross-multiplied comparisons    < 2^196  (I256 below)
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gi {

typedef __int128 i128;
typedef unsigned __int128 u128;

#define GI_HD __host__ __device__ __forceinline__

GI_HD int sgn64(int64_t x) { return (x > 0) - (x < 0); }
GI_HD int sgn128(i128 x) { return (x > 0) - (x < 0); }

// 256-bit sign-magnitude value, only ever produced by a 128x128 product and
// only ever compared.
struct I256 {
    int sign;        // -1, 0, +1
    uint64_t m[4];   // magnitude, little endian
};

GI_HD u128 uabs128(i128 x) { return x < 0 ? (u128)0 - (u128)x : (u128)x; }

GI_HD void mul64x64(uint64_t a, uint64_t b, uint64_t& lo, uint64_t& hi) {
    u128 p = (u128)a * (u128)b;
    lo = (uint64_t)p;
    hi = (uint64_t)(p >> 64);
}

GI_HD I256 mul128(i128 a, i128 b) {
    I256 r;
    int s = sgn128(a) * sgn128(b);
    r.sign = s;
    u128 ua = uabs128(a), ub = uabs128(b);
    uint64_t a0 = (uint64_t)ua, a1 = (uint64_t)(ua >> 64);
    uint64_t b0 = (uint64_t)ub, b1 = (uint64_t)(ub >> 64);
    uint64_t p00l, p00h, p01l, p01h, p10l, p10h, p11l, p11h;
    mul64x64(a0, b0, p00l, p00h);
    mul64x64(a0, b1, p01l, p01h);
    mul64x64(a1, b0, p10l, p10h);
    mul64x64(a1, b1, p11l, p11h);
    r.m[0] = p00l;
    u128 mid = (u128)p00h + p01l + p10l;
    r.m[1] = (uint64_t)mid;
    u128 hi = (mid >> 64) + (u128)p01h + p10h + p11l;
    r.m[2] = (uint64_t)hi;
    r.m[3] = (uint64_t)(hi >> 64) + p11h;
    return r;
}

GI_HD int cmp256(const I256& x, const I256& y) {
    if (x.sign != y.sign) return x.sign < y.sign ? -1 : 1;
    if (x.sign == 0) return 0;
    int c = 0;
    for (int i = 3; i >= 0; --i) {
        if (x.m[i] != y.m[i]) { c = x.m[i] < y.m[i] ? -1 : 1; break; }
    }
    return x.sign > 0 ? c : -c;
}

// An atomic add of `add` to a 32-bit counter that held `old` wrapped past
// 2^32 (the stored count is then meaningless: the caller raises a flag).
GI_HD bool u32_add_wraps(uint32_t old, uint32_t add) { return (uint32_t)(old + add) < old; }

// ------------------------------------------------------------------ wide
// The exact stage for general doubles works on integer coordinates with
// |x| < 2^61 (a pair's doubles scaled by their common power of two):
//   coordinate differences          < 2^62
//   cross products, 2D orientations < 2^125, their differences < 2^126 (i128)
//   orient3d determinants           < 2^190
//   homogeneous crossing points     < 2^251 (W256 below, two's complement)
//   cross-multiplied comparisons    < 2^441 (M512, sign-magnitude)

// n x m limb product, the low r limbs (schoolbook; r <= n + m)
template <int N, int M, int R>
GI_HD void mul_limbs(const uint64_t* a, const uint64_t* b, uint64_t* r) {
#pragma unroll
    for (int k = 0; k < R; ++k) r[k] = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        uint64_t carry = 0;
#pragma unroll
        for (int j = 0; j < M; ++j) {
            if (i + j >= R) break;
            u128 p = (u128)a[i] * b[j] + r[i + j] + carry;
            r[i + j] = (uint64_t)p;
            carry = (uint64_t)(p >> 64);
        }
        if (i + M < R) r[i + M] = carry;
    }
}

// 256-bit two's complement integer, little-endian limbs.
struct W256 {
    uint64_t w[4];
};

GI_HD W256 w256_from(i128 x) {
    W256 r;
    r.w[0] = (uint64_t)x;
    r.w[1] = (uint64_t)((u128)x >> 64);
    r.w[2] = r.w[3] = x < 0 ? ~0ull : 0ull;
    return r;
}
GI_HD W256 operator+(const W256& a, const W256& b) {
    W256 r;
    uint64_t c = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        u128 s = (u128)a.w[i] + b.w[i] + c;
        r.w[i] = (uint64_t)s;
        c = (uint64_t)(s >> 64);
    }
    return r;
}
GI_HD W256 operator-(const W256& a) {
    W256 r;
    uint64_t c = 1;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        u128 s = (u128)(~a.w[i]) + c;
        r.w[i] = (uint64_t)s;
        c = (uint64_t)(s >> 64);
    }
    return r;
}
GI_HD W256 operator-(const W256& a, const W256& b) {
    W256 r;
    uint64_t borrow = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const uint64_t d = a.w[i] - b.w[i];
        const uint64_t nb = (a.w[i] < b.w[i]) | (d < borrow);
        r.w[i] = d - borrow;
        borrow = nb;
    }
    return r;
}
GI_HD bool w256_neg(const W256& a) { return (a.w[3] >> 63) != 0; }
GI_HD int w256_sgn(const W256& a) {
    if (w256_neg(a)) return -1;
    return (a.w[0] | a.w[1] | a.w[2] | a.w[3]) != 0;
}
GI_HD W256 w256_abs(const W256& a) { return w256_neg(a) ? -a : a; }
GI_HD int cmp_limbs4(const uint64_t* a, const uint64_t* b) {
#pragma unroll
    for (int i = 3; i >= 0; --i)
        if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
    return 0;
}
// |a| vs |b|
GI_HD int w256_cmp_abs(const W256& a, const W256& b) {
    const W256 x = w256_abs(a), y = w256_abs(b);
    return cmp_limbs4(x.w, y.w);
}
// bits of |a| (0 for zero)
GI_HD int w256_bits(const W256& a) {
    const W256 x = w256_abs(a);
#pragma unroll
    for (int i = 3; i >= 0; --i)
        if (x.w[i]) return 64 * i + 64 - __builtin_clzll(x.w[i]);
    return 0;
}
// exact product of two i128 (|a|, |b| < 2^127)
GI_HD W256 w256_mul(i128 a, i128 b) {
    const u128 ua = uabs128(a), ub = uabs128(b);
    const uint64_t x[2] = {(uint64_t)ua, (uint64_t)(ua >> 64)}, y[2] = {(uint64_t)ub, (uint64_t)(ub >> 64)};
    W256 r;
    mul_limbs<2, 2, 4>(x, y, r.w);
    return ((a < 0) != (b < 0)) ? -r : r;
}
// product of a W256 and an i128, when it lies within 256 bits
GI_HD W256 w256_mul(const W256& a, i128 b) {
    const W256 ua = w256_abs(a);
    const u128 ub = uabs128(b);
    const uint64_t y[2] = {(uint64_t)ub, (uint64_t)(ub >> 64)};
    W256 r;
    mul_limbs<4, 2, 4>(ua.w, y, r.w);
    return (w256_neg(a) != (b < 0)) ? -r : r;
}

// 512-bit sign-magnitude value: only produced by a W256 x W256 product and
// only compared.
struct M512 {
    int sign;
    uint64_t m[8];
};
GI_HD M512 m512_mul(const W256& a, const W256& b) {
    M512 r;
    const int sa = w256_sgn(a), sb = w256_sgn(b);
    r.sign = sa * sb;
    const W256 ua = w256_abs(a), ub = w256_abs(b);
    mul_limbs<4, 4, 8>(ua.w, ub.w, r.m);
    return r;
}
GI_HD int m512_cmp(const M512& x, const M512& y) {
    if (x.sign != y.sign) return x.sign < y.sign ? -1 : 1;
    if (x.sign == 0) return 0;
    int c = 0;
#pragma unroll
    for (int i = 7; i >= 0; --i)
        if (x.m[i] != y.m[i]) { c = x.m[i] < y.m[i] ? -1 : 1; break; }
    return x.sign > 0 ? c : -c;
}

}  // namespace gi
