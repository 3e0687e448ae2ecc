/*
 * obig.h -- signed fixed-capacity big integers for the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY. The oracle is the checker for the HIP product
 * path; nothing in the product links or calls it.
 *
 * The reference evaluates its exact stage with 384-bit signed-magnitude
 * integers built on __int128 limbs (src/GPU_exact_predicates/
 * ExactBigIntPredicate.cu:11-148). The oracle needs more headroom, because it
 * never caps the exponent spread of the input doubles (the reference returns
 * YELLOW above a spread of 60, ExactBigIntPredicate.cu:419-420), so numbers
 * here are sign-magnitude with up to OB_LIMBS 64-bit limbs and abort loudly on
 * overflow instead of wrapping.
 */
#ifndef GI_ORACLE_OBIG_H
#define GI_ORACLE_OBIG_H

#include <stdint.h>

#define OB_LIMBS 40 /* 2560 bits */

typedef struct {
    int32_t sign; /* -1, 0, +1 */
    int32_t n;    /* used limbs; 0 iff value is zero */
    uint64_t d[OB_LIMBS];
} ob_t;

void ob_zero(ob_t* r);
void ob_from_i64(ob_t* r, int64_t v);
/* r = m * 2^k for a non-negative shift k (m may be negative). */
void ob_from_i64_shl(ob_t* r, int64_t m, int k);
void ob_copy(ob_t* r, const ob_t* a);
void ob_neg(ob_t* r, const ob_t* a);
void ob_add(ob_t* r, const ob_t* a, const ob_t* b);
void ob_sub(ob_t* r, const ob_t* a, const ob_t* b);
void ob_mul(ob_t* r, const ob_t* a, const ob_t* b);
int ob_sgn(const ob_t* a);
int ob_cmp(const ob_t* a, const ob_t* b);
int ob_cmpabs(const ob_t* a, const ob_t* b);
int ob_is_zero(const ob_t* a);
int ob_bits(const ob_t* a);
double ob_to_double(const ob_t* a);
/* Export as two's complement little-endian 64-bit limbs; returns 0 if it fits. */
int ob_export(const ob_t* a, uint64_t* out, int nlimbs);
/* Minimum number of two's complement 64-bit limbs needed for a. */
int ob_export_limbs(const ob_t* a);

#endif
