/*
 * obig.c -- sign-magnitude big integers for the CPU oracle (test
 * infrastructure only; see obig.h).
 */
#include "obig.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

static void ob_overflow(void) {
    fprintf(stderr, "oracle obig: overflow beyond %d limbs\n", OB_LIMBS);
    abort();
}

static void ob_norm(ob_t* r) {
    while (r->n > 0 && r->d[r->n - 1] == 0) r->n--;
    if (r->n == 0) r->sign = 0;
}

void ob_zero(ob_t* r) {
    r->sign = 0;
    r->n = 0;
}

void ob_from_i64(ob_t* r, int64_t v) {
    if (v == 0) { ob_zero(r); return; }
    r->sign = v < 0 ? -1 : 1;
    r->d[0] = v < 0 ? (uint64_t)(-(v + 1)) + 1u : (uint64_t)v;
    r->n = 1;
}

void ob_from_i64_shl(ob_t* r, int64_t m, int k) {
    ob_from_i64(r, m);
    if (r->n == 0 || k == 0) return;
    int limbs = k / 64, bits = k % 64;
    uint64_t lo = r->d[0];
    if (limbs + 2 > OB_LIMBS) ob_overflow();
    for (int i = 0; i < limbs; ++i) r->d[i] = 0;
    r->d[limbs] = lo << bits;
    r->d[limbs + 1] = bits ? (lo >> (64 - bits)) : 0;
    r->n = limbs + 2;
    ob_norm(r);
}

void ob_copy(ob_t* r, const ob_t* a) {
    if (r == a) return;
    r->sign = a->sign;
    r->n = a->n;
    memcpy(r->d, a->d, sizeof(uint64_t) * (size_t)a->n);
}

void ob_neg(ob_t* r, const ob_t* a) {
    ob_copy(r, a);
    r->sign = -r->sign;
}

static int mag_cmp(const uint64_t* a, int na, const uint64_t* b, int nb) {
    if (na != nb) return na < nb ? -1 : 1;
    for (int i = na - 1; i >= 0; --i)
        if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
    return 0;
}

/* out = a + b (magnitudes); out may alias a or b. Returns limb count. */
static int mag_add(uint64_t* out, const uint64_t* a, int na, const uint64_t* b, int nb) {
    if (na < nb) { const uint64_t* t = a; a = b; b = t; int tn = na; na = nb; nb = tn; }
    uint64_t carry = 0;
    for (int i = 0; i < na; ++i) {
        u128 s = (u128)a[i] + (i < nb ? b[i] : 0) + carry;
        out[i] = (uint64_t)s;
        carry = (uint64_t)(s >> 64);
    }
    if (carry) {
        if (na >= OB_LIMBS) ob_overflow();
        out[na++] = carry;
    }
    return na;
}

/* out = a - b with |a| >= |b|; out may alias a. */
static int mag_sub(uint64_t* out, const uint64_t* a, int na, const uint64_t* b, int nb) {
    uint64_t borrow = 0;
    for (int i = 0; i < na; ++i) {
        uint64_t bi = i < nb ? b[i] : 0;
        uint64_t d = a[i] - bi - borrow;
        borrow = (a[i] < bi + borrow) || (bi == UINT64_MAX && borrow) ? 1 : 0;
        out[i] = d;
    }
    while (na > 0 && out[na - 1] == 0) na--;
    return na;
}

static void ob_addsub(ob_t* r, const ob_t* a, const ob_t* b, int bsign) {
    if (b->sign == 0) { ob_copy(r, a); return; }
    if (a->sign == 0) { ob_copy(r, b); r->sign = bsign; return; }
    ob_t t;
    if (a->sign == bsign) {
        t.n = mag_add(t.d, a->d, a->n, b->d, b->n);
        t.sign = a->sign;
    } else {
        int c = mag_cmp(a->d, a->n, b->d, b->n);
        if (c == 0) { ob_zero(r); return; }
        if (c > 0) { t.n = mag_sub(t.d, a->d, a->n, b->d, b->n); t.sign = a->sign; }
        else { t.n = mag_sub(t.d, b->d, b->n, a->d, a->n); t.sign = bsign; }
    }
    ob_norm(&t);
    ob_copy(r, &t);
}

void ob_add(ob_t* r, const ob_t* a, const ob_t* b) { ob_addsub(r, a, b, b->sign); }
void ob_sub(ob_t* r, const ob_t* a, const ob_t* b) { ob_addsub(r, a, b, -b->sign); }

void ob_mul(ob_t* r, const ob_t* a, const ob_t* b) {
    if (a->sign == 0 || b->sign == 0) { ob_zero(r); return; }
    ob_t t;
    int n = a->n + b->n;
    if (n > OB_LIMBS) ob_overflow();
    memset(t.d, 0, sizeof(uint64_t) * (size_t)n);
    for (int i = 0; i < a->n; ++i) {
        uint64_t carry = 0;
        for (int j = 0; j < b->n; ++j) {
            u128 p = (u128)a->d[i] * b->d[j] + t.d[i + j] + carry;
            t.d[i + j] = (uint64_t)p;
            carry = (uint64_t)(p >> 64);
        }
        t.d[i + b->n] = carry;
    }
    t.n = n;
    t.sign = a->sign * b->sign;
    ob_norm(&t);
    ob_copy(r, &t);
}

int ob_sgn(const ob_t* a) { return a->sign; }
int ob_is_zero(const ob_t* a) { return a->sign == 0; }

int ob_cmpabs(const ob_t* a, const ob_t* b) { return mag_cmp(a->d, a->n, b->d, b->n); }

int ob_cmp(const ob_t* a, const ob_t* b) {
    if (a->sign != b->sign) return a->sign < b->sign ? -1 : 1;
    if (a->sign == 0) return 0;
    int c = mag_cmp(a->d, a->n, b->d, b->n);
    return a->sign > 0 ? c : -c;
}

int ob_bits(const ob_t* a) {
    if (a->n == 0) return 0;
    return 64 * (a->n - 1) + (64 - __builtin_clzll(a->d[a->n - 1]));
}

double ob_to_double(const ob_t* a) {
    double v = 0.0;
    for (int i = a->n - 1; i >= 0; --i) v = v * 18446744073709551616.0 + (double)a->d[i];
    return a->sign < 0 ? -v : v;
}

int ob_export_limbs(const ob_t* a) {
    /* two's complement needs bits+1 (sign bit) */
    int b = ob_bits(a) + 1;
    int n = (b + 63) / 64;
    return n < 1 ? 1 : n;
}

int ob_export(const ob_t* a, uint64_t* out, int nlimbs) {
    if (ob_export_limbs(a) > nlimbs) return -1;
    for (int i = 0; i < nlimbs; ++i) out[i] = i < a->n ? a->d[i] : 0;
    if (a->sign < 0) {
        /* negate in place: invert and add one */
        uint64_t carry = 1;
        for (int i = 0; i < nlimbs; ++i) {
            u128 s = (u128)(~out[i]) + carry;
            out[i] = (uint64_t)s;
            carry = (uint64_t)(s >> 64);
        }
    }
    return 0;
}
