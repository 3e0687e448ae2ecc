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

}  // namespace gi
