/*
 * exact.c -- exact triangle/triangle decision and intersection for the CPU
 * oracle (test infrastructure only; see oracle.h for what it restates).
 *
 * Every coordinate of a triangle pair is turned into an integer by one common
 * power-of-two scale (all doubles are dyadic rationals), so every predicate
 * below is an exact integer determinant, and every constructed point is an
 * exact homogeneous rational (x, y, z, w) with w > 0.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "obig.h"
#include "oracle.h"

typedef struct { ob_t c[3]; } ov3;
typedef struct { ob_t c[3]; ob_t w; } ohp;

enum { PAIR_NO = 0, PAIR_GREEN = 1, PAIR_YELLOW = 2 };

/* ------------------------------------------------------------------ vectors */

static void v_sub(ov3* r, const ov3* a, const ov3* b) {
    for (int k = 0; k < 3; ++k) ob_sub(&r->c[k], &a->c[k], &b->c[k]);
}

static void v_cross(ov3* r, const ov3* a, const ov3* b) {
    ov3 t;
    ob_t p, q;
    for (int k = 0; k < 3; ++k) {
        int i = (k + 1) % 3, j = (k + 2) % 3;
        ob_mul(&p, &a->c[i], &b->c[j]);
        ob_mul(&q, &a->c[j], &b->c[i]);
        ob_sub(&t.c[k], &p, &q);
    }
    *r = t;
}

static void v_dot(ob_t* r, const ov3* a, const ov3* b) {
    ob_t acc, p;
    ob_mul(&acc, &a->c[0], &b->c[0]);
    ob_mul(&p, &a->c[1], &b->c[1]);
    ob_add(&acc, &acc, &p);
    ob_mul(&p, &a->c[2], &b->c[2]);
    ob_add(r, &acc, &p);
}

static int v_is_zero(const ov3* a) {
    return ob_is_zero(&a->c[0]) && ob_is_zero(&a->c[1]) && ob_is_zero(&a->c[2]);
}

static int v_eq(const ov3* a, const ov3* b) {
    for (int k = 0; k < 3; ++k)
        if (ob_cmp(&a->c[k], &b->c[k]) != 0) return 0;
    return 1;
}

/* orient3d(a,b,c,d) = (a-d) . ((b-d) x (c-d)), the determinant form of
 * src/GPU_exact_predicates/ExactBigIntPredicate.cu:191-210. */
static void orient3d_val(ob_t* r, const ov3* a, const ov3* b, const ov3* c, const ov3* d) {
    ov3 ad, bd, cd, x;
    v_sub(&ad, a, d);
    v_sub(&bd, b, d);
    v_sub(&cd, c, d);
    v_cross(&x, &bd, &cd);
    v_dot(r, &ad, &x);
}

static int orient3d_sgn(const ov3* a, const ov3* b, const ov3* c, const ov3* d) {
    ob_t r;
    orient3d_val(&r, a, b, c, d);
    return ob_sgn(&r);
}

/* Projection used by the reference's coplanar tests: drop the axis of the
 * largest |normal| component, preferring x then y
 * (ExactBigIntPredicate.cu:244-245). */
static int dominant_axis(const ov3* n) {
    if (ob_cmpabs(&n->c[0], &n->c[1]) >= 0 && ob_cmpabs(&n->c[0], &n->c[2]) >= 0) return 0;
    if (ob_cmpabs(&n->c[1], &n->c[2]) >= 0) return 1;
    return 2;
}

static void kept_axes(int drop, int* ia, int* ib) {
    if (drop == 0) { *ia = 1; *ib = 2; }
    else if (drop == 1) { *ia = 0; *ib = 2; }
    else { *ia = 0; *ib = 1; }
}

static void orient2d_val(ob_t* r, const ov3* p, const ov3* q, const ov3* s, int ia, int ib) {
    ob_t qa, qb, sa, sb, t1, t2;
    ob_sub(&qa, &q->c[ia], &p->c[ia]);
    ob_sub(&qb, &q->c[ib], &p->c[ib]);
    ob_sub(&sa, &s->c[ia], &p->c[ia]);
    ob_sub(&sb, &s->c[ib], &p->c[ib]);
    ob_mul(&t1, &qa, &sb);
    ob_mul(&t2, &qb, &sa);
    ob_sub(r, &t1, &t2);
}

static int orient2d_sgn(const ov3* p, const ov3* q, const ov3* s, int ia, int ib) {
    ob_t r;
    orient2d_val(&r, p, q, s, ia, ib);
    return ob_sgn(&r);
}

/* ------------------------------------------------- reference decision tree */

/* all_outside_edge_V2, ExactBigIntPredicate.cu:217-234 */
static int all_outside_edge(const ov3 E[3], const ov3 O[3], int i, int j, int k, int ia, int ib) {
    int oref = orient2d_sgn(&E[i], &E[j], &E[k], ia, ib);
    if (oref == 0) return 0;
    for (int m = 0; m < 3; ++m) {
        int o = orient2d_sgn(&E[i], &E[j], &O[m], ia, ib);
        int outside = oref > 0 ? (o < 0) : (o > 0);
        if (!outside) return 0;
    }
    return 1;
}

/* coplanar_tri_tri_exact_V2, ExactBigIntPredicate.cu:236-274 */
static int coplanar_tritri(const ov3 A[3], const ov3 B[3]) {
    ov3 d1, d2, n;
    v_sub(&d1, &A[1], &A[0]);
    v_sub(&d2, &A[2], &A[0]);
    v_cross(&n, &d1, &d2);
    int ia, ib;
    kept_axes(dominant_axis(&n), &ia, &ib);
    for (int i = 0; i < 3; ++i)
        if (all_outside_edge(A, B, i, (i + 1) % 3, (i + 2) % 3, ia, ib)) return PAIR_NO;
    for (int i = 0; i < 3; ++i)
        if (all_outside_edge(B, A, i, (i + 1) % 3, (i + 2) % 3, ia, ib)) return PAIR_NO;
    return PAIR_GREEN;
}

/* point_in_or_on_tri_2d_V2, ExactBigIntPredicate.cu:276-290 */
static int point_in_or_on_tri_2d(const ov3* p, const ov3 T[3], int ia, int ib) {
    for (int i = 0; i < 3; ++i) {
        int j = (i + 1) % 3, k = (i + 2) % 3;
        int oref = orient2d_sgn(&T[i], &T[j], &T[k], ia, ib);
        if (oref == 0) continue;
        int o = orient2d_sgn(&T[i], &T[j], p, ia, ib);
        if (oref > 0 ? (o < 0) : (o > 0)) return 0;
    }
    return 1;
}

/* on_segment_collinear_V2, ExactBigIntPredicate.cu:292-300 */
static int on_segment_collinear(const ov3* p, const ov3* q, const ov3* r, int ia, int ib) {
    int axes[2] = {ia, ib};
    for (int t = 0; t < 2; ++t) {
        int k = axes[t];
        const ob_t* lo = ob_cmp(&p->c[k], &q->c[k]) <= 0 ? &p->c[k] : &q->c[k];
        const ob_t* hi = ob_cmp(&p->c[k], &q->c[k]) <= 0 ? &q->c[k] : &p->c[k];
        if (ob_cmp(&r->c[k], lo) < 0 || ob_cmp(&r->c[k], hi) > 0) return 0;
    }
    return 1;
}

/* segments_intersect_2d_V2, ExactBigIntPredicate.cu:302-319 */
static int segments_intersect_2d(const ov3* a, const ov3* b, const ov3* c, const ov3* d, int ia, int ib) {
    int o1 = orient2d_sgn(a, b, c, ia, ib);
    int o2 = orient2d_sgn(a, b, d, ia, ib);
    int o3 = orient2d_sgn(c, d, a, ia, ib);
    int o4 = orient2d_sgn(c, d, b, ia, ib);
    if (o1 != 0 && o2 != 0 && o3 != 0 && o4 != 0 && o1 != o2 && o3 != o4) return 1;
    if (o1 == 0 && on_segment_collinear(a, b, c, ia, ib)) return 1;
    if (o2 == 0 && on_segment_collinear(a, b, d, ia, ib)) return 1;
    if (o3 == 0 && on_segment_collinear(c, d, a, ia, ib)) return 1;
    if (o4 == 0 && on_segment_collinear(c, d, b, ia, ib)) return 1;
    return 0;
}

/* edgeTri_coplanar_exact_V2, ExactBigIntPredicate.cu:321-354 */
static int edgetri_coplanar(const ov3* a, const ov3* b, const ov3 T[3]) {
    ov3 d1, d2, n;
    v_sub(&d1, &T[1], &T[0]);
    v_sub(&d2, &T[2], &T[0]);
    v_cross(&n, &d1, &d2);
    int ia, ib;
    kept_axes(dominant_axis(&n), &ia, &ib);
    if (point_in_or_on_tri_2d(a, T, ia, ib)) return PAIR_GREEN;
    if (point_in_or_on_tri_2d(b, T, ia, ib)) return PAIR_GREEN;
    for (int i = 0; i < 3; ++i)
        if (segments_intersect_2d(a, b, &T[i], &T[(i + 1) % 3], ia, ib)) return PAIR_GREEN;
    return PAIR_NO;
}

/* edgeTri_exact_V2, ExactBigIntPredicate.cu:356-376 */
static int edgetri(const ov3* a, const ov3* b, const ov3 T[3]) {
    int s0 = orient3d_sgn(&T[0], &T[1], &T[2], a);
    int s1 = orient3d_sgn(&T[0], &T[1], &T[2], b);
    if ((s0 > 0 && s1 > 0) || (s0 < 0 && s1 < 0)) return PAIR_NO;
    if (s0 == 0 && s1 == 0) return edgetri_coplanar(a, b, T);
    int e0 = orient3d_sgn(a, b, &T[0], &T[1]);
    int e1 = orient3d_sgn(a, b, &T[1], &T[2]);
    int e2 = orient3d_sgn(a, b, &T[2], &T[0]);
    if ((e0 >= 0 && e1 >= 0 && e2 >= 0) || (e0 <= 0 && e1 <= 0 && e2 <= 0)) return PAIR_GREEN;
    return PAIR_NO;
}

/* --------------------------------------------------- homogeneous points */

static void hp_from_int(ohp* r, const ov3* p) {
    for (int k = 0; k < 3; ++k) ob_copy(&r->c[k], &p->c[k]);
    ob_from_i64(&r->w, 1);
}

static void hp_normalize(ohp* r) {
    if (ob_sgn(&r->w) < 0) {
        for (int k = 0; k < 3; ++k) ob_neg(&r->c[k], &r->c[k]);
        ob_neg(&r->w, &r->w);
    }
}

/* Point where segment (u,v) meets the zero set of an affine function with
 * values fu at u and fv at v (fu != fv): (fu*v - fv*u) / (fu - fv). */
static void hp_crossing_int(ohp* r, const ov3* u, const ov3* v, const ob_t* fu, const ob_t* fv) {
    ob_t t1, t2;
    for (int k = 0; k < 3; ++k) {
        ob_mul(&t1, fu, &v->c[k]);
        ob_mul(&t2, fv, &u->c[k]);
        ob_sub(&r->c[k], &t1, &t2);
    }
    ob_sub(&r->w, fu, fv);
    hp_normalize(r);
}

static void hp_crossing(ohp* r, const ohp* P, const ohp* Q, const ob_t* fp, const ob_t* fq) {
    ohp t;
    ob_t t1, t2;
    for (int k = 0; k < 3; ++k) {
        ob_mul(&t1, fp, &Q->c[k]);
        ob_mul(&t2, fq, &P->c[k]);
        ob_sub(&t.c[k], &t1, &t2);
    }
    ob_mul(&t1, fp, &Q->w);
    ob_mul(&t2, fq, &P->w);
    ob_sub(&t.w, &t1, &t2);
    hp_normalize(&t);
    *r = t;
}

/* sign(P_k - Q_k) for homogeneous points with positive w */
static int hp_cmp_axis(const ohp* P, const ohp* Q, int k) {
    ob_t l, r;
    ob_mul(&l, &P->c[k], &Q->w);
    ob_mul(&r, &Q->c[k], &P->w);
    return ob_cmp(&l, &r);
}

static int hp_eq(const ohp* P, const ohp* Q) {
    for (int k = 0; k < 3; ++k)
        if (hp_cmp_axis(P, Q, k) != 0) return 0;
    return 1;
}

/* sign of orient2d of three homogeneous points in the (ia, ib) projection */
static int hp_orient2d(const ohp* P, const ohp* Q, const ohp* R, int ia, int ib) {
    /* det [[Pa Pb Pw] [Qa Qb Qw] [Ra Rb Rw]] */
    ob_t m1, m2, m3, t1, t2, acc;
    ob_mul(&t1, &Q->c[ib], &R->w); ob_mul(&t2, &Q->w, &R->c[ib]); ob_sub(&m1, &t1, &t2);
    ob_mul(&t1, &Q->c[ia], &R->w); ob_mul(&t2, &Q->w, &R->c[ia]); ob_sub(&m2, &t1, &t2);
    ob_mul(&t1, &Q->c[ia], &R->c[ib]); ob_mul(&t2, &Q->c[ib], &R->c[ia]); ob_sub(&m3, &t1, &t2);
    ob_mul(&acc, &P->c[ia], &m1);
    ob_mul(&t1, &P->c[ib], &m2);
    ob_sub(&acc, &acc, &t1);
    ob_mul(&t1, &P->w, &m3);
    ob_add(&acc, &acc, &t1);
    return ob_sgn(&acc);
}

/* -------------------------------------------------------------- output */

typedef struct {
    int n;
    ohp p[ORACLE_MAX_SEG][2];
} segset;

static void seg_push(segset* s, const ohp* a, const ohp* b) {
    if (s->n >= ORACLE_MAX_SEG) abort();
    s->p[s->n][0] = *a;
    s->p[s->n][1] = *b;
    s->n++;
}

/* ------------------------------------------------------ polygon clipping */

#define POLY_CAP 16

/* Clip a convex polygon (homogeneous points lying in one plane) by the closed
 * half-plane sigma * orient2d(u, v, X) >= 0 of the (ia, ib) projection
 * (Sutherland-Hodgman). */
static int clip_halfplane(ohp* out, const ohp* in, int n, const ov3* u, const ov3* v, int sigma,
                          int ia, int ib) {
    if (n == 0) return 0;
    ob_t f[POLY_CAP];
    ob_t da, db;
    ob_sub(&da, &v->c[ia], &u->c[ia]);
    ob_sub(&db, &v->c[ib], &u->c[ib]);
    for (int i = 0; i < n; ++i) {
        ob_t xa, xb, t1, t2;
        ob_mul(&t1, &u->c[ia], &in[i].w);
        ob_sub(&xa, &in[i].c[ia], &t1);
        ob_mul(&t1, &u->c[ib], &in[i].w);
        ob_sub(&xb, &in[i].c[ib], &t1);
        ob_mul(&t1, &da, &xb);
        ob_mul(&t2, &db, &xa);
        ob_sub(&f[i], &t1, &t2);
        if (sigma < 0) ob_neg(&f[i], &f[i]);
    }
    int m = 0;
    for (int i = 0; i < n; ++i) {
        int s = (i + n - 1) % n;
        int ein = ob_sgn(&f[i]) >= 0, sin_ = ob_sgn(&f[s]) >= 0;
        if (ein) {
            if (!sin_) { if (m >= POLY_CAP) abort(); hp_crossing(&out[m++], &in[s], &in[i], &f[s], &f[i]); }
            if (m >= POLY_CAP) abort();
            out[m++] = in[i];
        } else if (sin_) {
            if (m >= POLY_CAP) abort();
            hp_crossing(&out[m++], &in[s], &in[i], &f[s], &f[i]);
        }
    }
    return m;
}

/* Reduce a clipped convex point cycle to canonical boundary segments. */
static void polygon_to_segments(const ohp* in, int n, int ia, int ib, segset* out) {
    ohp p[POLY_CAP];
    int m = 0;
    for (int i = 0; i < n; ++i) {
        int dup = 0;
        for (int j = 0; j < m; ++j)
            if (hp_eq(&p[j], &in[i])) { dup = 1; break; }
        if (!dup) p[m++] = in[i];
    }
    if (m == 0) return;
    if (m == 1) { seg_push(out, &p[0], &p[0]); return; }
    /* all collinear? */
    int collinear = 1;
    for (int i = 2; i < m && collinear; ++i)
        if (hp_orient2d(&p[0], &p[1], &p[i], ia, ib) != 0) collinear = 0;
    if (collinear) {
        /* extremes along an axis on which the points differ */
        int k = 0;
        for (int t = 0; t < 3; ++t)
            if (hp_cmp_axis(&p[0], &p[1], t) != 0) { k = t; break; }
        int lo = 0, hi = 0;
        for (int i = 1; i < m; ++i) {
            if (hp_cmp_axis(&p[i], &p[lo], k) < 0) lo = i;
            if (hp_cmp_axis(&p[i], &p[hi], k) > 0) hi = i;
        }
        seg_push(out, &p[lo], &p[hi]);
        return;
    }
    /* The cycle order of the clipper is kept after dropping duplicates (a
     * duplicate is always adjacent to its twin in a convex clip); remove
     * vertices in the interior of an edge. */
    int changed = 1;
    while (changed && m > 3) {
        changed = 0;
        for (int i = 0; i < m; ++i) {
            const ohp* a = &p[(i + m - 1) % m];
            const ohp* c = &p[(i + 1) % m];
            if (hp_orient2d(a, &p[i], c, ia, ib) == 0) {
                for (int j = i; j + 1 < m; ++j) p[j] = p[j + 1];
                m--;
                changed = 1;
                break;
            }
        }
    }
    for (int i = 0; i < m; ++i) seg_push(out, &p[i], &p[(i + 1) % m]);
}

static int clip_by_triangle(ohp* poly, int n, const ov3 T[3], int ia, int ib) {
    int sigma = orient2d_sgn(&T[0], &T[1], &T[2], ia, ib);
    ohp tmp[POLY_CAP];
    for (int e = 0; e < 3 && n > 0; ++e) {
        n = clip_halfplane(tmp, poly, n, &T[e], &T[(e + 1) % 3], sigma, ia, ib);
        memcpy(poly, tmp, sizeof(ohp) * (size_t)n);
    }
    return n;
}

/* ---------------------------------------------- degenerate configurations */

/* Point-set reduction of a triangle: 2 (proper triangle), 1 (segment with
 * endpoints T[*e0], T[*e1]) or 0 (a point). */
static int simplex_dim(const ov3 T[3], const ov3* n, int* e0, int* e1) {
    if (!v_is_zero(n)) return 2;
    if (v_eq(&T[0], &T[1]) && v_eq(&T[1], &T[2])) { *e0 = *e1 = 0; return 0; }
    for (int k = 0; k < 3; ++k) {
        int lo = 0, hi = 0;
        for (int i = 1; i < 3; ++i) {
            if (ob_cmp(&T[i].c[k], &T[lo].c[k]) < 0) lo = i;
            if (ob_cmp(&T[i].c[k], &T[hi].c[k]) > 0) hi = i;
        }
        if (ob_cmp(&T[lo].c[k], &T[hi].c[k]) != 0) { *e0 = lo; *e1 = hi; return 1; }
    }
    abort();
}

/* closed segment [p,q] (p may equal q) against a proper triangle T */
static int seg_tri(const ov3* p, const ov3* q, const ov3 T[3], const ov3* nT, segset* out) {
    ob_t op, oq;
    orient3d_val(&op, &T[0], &T[1], &T[2], p);
    orient3d_val(&oq, &T[0], &T[1], &T[2], q);
    int sp = ob_sgn(&op), sq = ob_sgn(&oq);
    if (sp * sq > 0) return 0;
    if (sp == 0 && sq == 0) {
        int ia, ib;
        kept_axes(dominant_axis(nT), &ia, &ib);
        ohp poly[POLY_CAP];
        hp_from_int(&poly[0], p);
        hp_from_int(&poly[1], q);
        int n = clip_by_triangle(poly, 2, T, ia, ib);
        if (n == 0) return 0;
        segset tmp = {0};
        polygon_to_segments(poly, n, ia, ib, &tmp);
        if (out) for (int i = 0; i < tmp.n; ++i) seg_push(out, &tmp.p[i][0], &tmp.p[i][1]);
        return tmp.n > 0;
    }
    /* transversal: the line pq meets the plane once; inside test by the
     * Plucker signs of pq against the triangle edges */
    int e0 = orient3d_sgn(p, q, &T[0], &T[1]);
    int e1 = orient3d_sgn(p, q, &T[1], &T[2]);
    int e2 = orient3d_sgn(p, q, &T[2], &T[0]);
    if (!((e0 >= 0 && e1 >= 0 && e2 >= 0) || (e0 <= 0 && e1 <= 0 && e2 <= 0))) return 0;
    if (out) {
        ohp X;
        if (sp == 0) hp_from_int(&X, p);
        else if (sq == 0) hp_from_int(&X, q);
        else hp_crossing_int(&X, p, q, &op, &oq);
        seg_push(out, &X, &X);
    }
    return 1;
}

/* closed segments [p,q] and [r,s], each possibly a single point */
static int seg_seg(const ov3* p, const ov3* q, const ov3* r, const ov3* s, segset* out) {
    int pt1 = v_eq(p, q), pt2 = v_eq(r, s);
    if (pt1 && pt2) {
        if (!v_eq(p, r)) return 0;
        if (out) { ohp X; hp_from_int(&X, p); seg_push(out, &X, &X); }
        return 1;
    }
    if (pt1 || pt2) {
        const ov3* X = pt1 ? p : r;
        const ov3* a = pt1 ? r : p;
        const ov3* b = pt1 ? s : q;
        ov3 ab, ax, bx, c;
        v_sub(&ab, b, a);
        v_sub(&ax, X, a);
        v_sub(&bx, X, b);
        v_cross(&c, &ab, &ax);
        if (!v_is_zero(&c)) return 0;
        ob_t d1, d2;
        v_dot(&d1, &ax, &ab);
        v_dot(&d2, &bx, &ab);
        if (ob_sgn(&d1) < 0 || ob_sgn(&d2) > 0) return 0;
        if (out) { ohp H; hp_from_int(&H, X); seg_push(out, &H, &H); }
        return 1;
    }
    if (orient3d_sgn(p, q, r, s) != 0) return 0;
    ov3 pq, rs, pr, c;
    v_sub(&pq, q, p);
    v_sub(&rs, s, r);
    v_cross(&c, &pq, &rs);
    if (v_is_zero(&c)) {
        ov3 c2;
        v_sub(&pr, r, p);
        v_cross(&c2, &pq, &pr);
        if (!v_is_zero(&c2)) return 0; /* parallel, distinct lines */
        int k = 0;
        for (int t = 0; t < 3; ++t)
            if (!ob_is_zero(&pq.c[t])) { k = t; break; }
        const ov3* lo1 = ob_cmp(&p->c[k], &q->c[k]) <= 0 ? p : q;
        const ov3* hi1 = lo1 == p ? q : p;
        const ov3* lo2 = ob_cmp(&r->c[k], &s->c[k]) <= 0 ? r : s;
        const ov3* hi2 = lo2 == r ? s : r;
        const ov3* lo = ob_cmp(&lo1->c[k], &lo2->c[k]) >= 0 ? lo1 : lo2;
        const ov3* hi = ob_cmp(&hi1->c[k], &hi2->c[k]) <= 0 ? hi1 : hi2;
        if (ob_cmp(&lo->c[k], &hi->c[k]) > 0) return 0;
        if (out) { ohp A, B; hp_from_int(&A, lo); hp_from_int(&B, hi); seg_push(out, &A, &B); }
        return 1;
    }
    int ia, ib;
    kept_axes(dominant_axis(&c), &ia, &ib);
    ob_t o1, o2;
    orient2d_val(&o1, p, q, r, ia, ib);
    orient2d_val(&o2, p, q, s, ia, ib);
    int s3 = orient2d_sgn(r, s, p, ia, ib);
    int s4 = orient2d_sgn(r, s, q, ia, ib);
    if (ob_sgn(&o1) * ob_sgn(&o2) > 0 || s3 * s4 > 0) return 0;
    if (out) { ohp X; hp_crossing_int(&X, r, s, &o1, &o2); seg_push(out, &X, &X); }
    return 1;
}

/* ---------------------------------------------------------- main decision */

/* Decision of classifyPairBigIntV2 (ExactBigIntPredicate.cu:381-489) for two
 * proper triangles, plus the exact intersection set when they meet. */
static int exact_pair(const ov3 A[3], const ov3 B[3], int* coplanar, int* degenerate, segset* segs) {
    ov3 d1, d2, nA, nB;
    v_sub(&d1, &A[1], &A[0]);
    v_sub(&d2, &A[2], &A[0]);
    v_cross(&nA, &d1, &d2);
    v_sub(&d1, &B[1], &B[0]);
    v_sub(&d2, &B[2], &B[0]);
    v_cross(&nB, &d1, &d2);
    *coplanar = 0;
    *degenerate = 0;

    int a0, a1, b0, b1;
    int dimA = simplex_dim(A, &nA, &a0, &a1);
    int dimB = simplex_dim(B, &nB, &b0, &b1);
    if (dimA < 2 || dimB < 2) {
        *degenerate = 1;
        if (dimA == 2) return seg_tri(&B[b0], &B[b1], A, &nA, segs);
        if (dimB == 2) return seg_tri(&A[a0], &A[a1], B, &nB, segs);
        return seg_seg(&A[a0], &A[a1], &B[b0], &B[b1], segs);
    }

    ob_t ob[3], oa[3];
    int sb[3], sa[3];
    for (int i = 0; i < 3; ++i) { orient3d_val(&ob[i], &A[0], &A[1], &A[2], &B[i]); sb[i] = ob_sgn(&ob[i]); }
    if (sb[0] == 0 && sb[1] == 0 && sb[2] == 0) {
        *coplanar = 1;
        int r = coplanar_tritri(A, B);
        if (r == PAIR_GREEN && segs) {
            int ia, ib;
            kept_axes(dominant_axis(&nA), &ia, &ib);
            ohp poly[POLY_CAP];
            for (int i = 0; i < 3; ++i) hp_from_int(&poly[i], &A[i]);
            int n = clip_by_triangle(poly, 3, B, ia, ib);
            polygon_to_segments(poly, n, ia, ib, segs);
            if (segs->n == 0) abort(); /* decision and construction disagree */
        }
        return r == PAIR_GREEN;
    }
    if ((sb[0] > 0 && sb[1] > 0 && sb[2] > 0) || (sb[0] < 0 && sb[1] < 0 && sb[2] < 0)) return 0;
    for (int i = 0; i < 3; ++i) { orient3d_val(&oa[i], &B[0], &B[1], &B[2], &A[i]); sa[i] = ob_sgn(&oa[i]); }
    if ((sa[0] > 0 && sa[1] > 0 && sa[2] > 0) || (sa[0] < 0 && sa[1] < 0 && sa[2] < 0)) return 0;

    int hit = 0;
    for (int i = 0; i < 3 && !hit; ++i) hit = edgetri(&A[i], &A[(i + 1) % 3], B) == PAIR_GREEN;
    for (int i = 0; i < 3 && !hit; ++i) hit = edgetri(&B[i], &B[(i + 1) % 3], A) == PAIR_GREEN;
    if (!hit || !segs) return hit;

    /* Non-coplanar: T_A meets plane B in S1, T_B meets plane A in S2, both on
     * the line L = plane A ^ plane B; the intersection is S1 ^ S2 ordered
     * along an axis on which L's direction is non-zero. */
    ov3 dir;
    v_cross(&dir, &nA, &nB);
    int k = 0;
    for (int t = 1; t < 3; ++t)
        if (ob_cmpabs(&dir.c[t], &dir.c[k]) > 0) k = t;
    ohp lo[2], hi[2];
    for (int side = 0; side < 2; ++side) {
        const ov3* T = side == 0 ? A : B;
        const ob_t* o = side == 0 ? oa : ob;
        const int* s = side == 0 ? sa : sb;
        ohp pts[4];
        int m = 0;
        for (int i = 0; i < 3; ++i)
            if (s[i] == 0) hp_from_int(&pts[m++], &T[i]);
        for (int i = 0; i < 3; ++i) {
            int j = (i + 1) % 3;
            if (s[i] * s[j] < 0) hp_crossing_int(&pts[m++], &T[i], &T[j], &o[i], &o[j]);
        }
        if (m == 0) abort();
        int il = 0, ih = 0;
        for (int i = 1; i < m; ++i) {
            if (hp_cmp_axis(&pts[i], &pts[il], k) < 0) il = i;
            if (hp_cmp_axis(&pts[i], &pts[ih], k) > 0) ih = i;
        }
        lo[side] = pts[il];
        hi[side] = pts[ih];
    }
    const ohp* L = hp_cmp_axis(&lo[0], &lo[1], k) >= 0 ? &lo[0] : &lo[1];
    const ohp* H = hp_cmp_axis(&hi[0], &hi[1], k) <= 0 ? &hi[0] : &hi[1];
    if (hp_cmp_axis(L, H, k) > 0) abort(); /* decision and construction disagree */
    seg_push(segs, L, H);
    return 1;
}

/* ------------------------------------------------------- double -> integer */

/* Common scale 2^e (e <= 0) that turns all 18 coordinates into integers. */
static void pair_to_ints(const double a[9], const double b[9], ov3 A[3], ov3 B[3], int* exp2) {
    int emin = 0;
    const double* src[2] = {a, b};
    for (int s = 0; s < 2; ++s)
        for (int i = 0; i < 9; ++i) {
            double x = src[s][i];
            if (x == 0.0) continue;
            int e;
            double m = frexp(x, &e);
            int64_t M = (int64_t)ldexp(m, 53);
            uint64_t um = M < 0 ? (uint64_t)(-M) : (uint64_t)M;
            int elow = e - 53 + __builtin_ctzll(um);
            if (elow < emin) emin = elow;
        }
    *exp2 = emin;
    for (int s = 0; s < 2; ++s) {
        ov3* T = s == 0 ? A : B;
        for (int i = 0; i < 9; ++i) {
            double x = src[s][i];
            ob_t* dst = &T[i / 3].c[i % 3];
            if (x == 0.0) { ob_zero(dst); continue; }
            int e;
            double m = frexp(x, &e);
            int64_t M = (int64_t)ldexp(m, 53);
            uint64_t um = M < 0 ? (uint64_t)(-M) : (uint64_t)M;
            int t = __builtin_ctzll(um);
            int64_t odd = M / ((int64_t)1 << t);
            ob_from_i64_shl(dst, odd, (e - 53 + t) - emin);
        }
    }
}

/* ------------------------------------------------------------- FP filter */

/* Shewchuk's orient3d with its static error bound errboundA = (7 + 56 eps)
 * eps * permanent, eps = 2^-53 (J. R. Shewchuk, "Adaptive Precision
 * Floating-Point Arithmetic and Fast Robust Geometric Predicates", 1997).
 * Returns +1/-1 if the sign of the exact determinant is certified, else 0. */
int oracle_orient3d_filter(const double a[3], const double b[3], const double c[3], const double d[3]) {
    const double eps = 1.1102230246251565e-16; /* 2^-53 */
    const double bound = (7.0 + 56.0 * eps) * eps;
    double adx = a[0] - d[0], ady = a[1] - d[1], adz = a[2] - d[2];
    double bdx = b[0] - d[0], bdy = b[1] - d[1], bdz = b[2] - d[2];
    double cdx = c[0] - d[0], cdy = c[1] - d[1], cdz = c[2] - d[2];
    double bdxcdy = bdx * cdy, cdxbdy = cdx * bdy;
    double cdxady = cdx * ady, adxcdy = adx * cdy;
    double adxbdy = adx * bdy, bdxady = bdx * ady;
    double det = adz * (bdxcdy - cdxbdy) + bdz * (cdxady - adxcdy) + cdz * (adxbdy - bdxady);
    double perm = (fabs(bdxcdy) + fabs(cdxbdy)) * fabs(adz) + (fabs(cdxady) + fabs(adxcdy)) * fabs(bdz) +
                  (fabs(adxbdy) + fabs(bdxady)) * fabs(cdz);
    /* below ~2^-900 products may have underflowed and the bound is void */
    if (!(perm >= 1e-270) || !isfinite(perm)) return 0;
    double err = bound * perm;
    if (det > err) return 1;
    if (-det > err) return -1;
    return 0;
}

/* edgeTriWithSigns, OriginShiftedFloatPredicateCore.cuh:70-85 */
static int filt_edgetri(const double* a, const double* b, int sa, int sb, const double* p, const double* q,
                        const double* r) {
    if (sa == 0 || sb == 0) return PAIR_YELLOW;
    if (sa == sb) return PAIR_NO;
    int e0 = oracle_orient3d_filter(a, b, p, q);
    int e1 = oracle_orient3d_filter(a, b, q, r);
    int e2 = oracle_orient3d_filter(a, b, r, p);
    if (e0 == 0 || e1 == 0 || e2 == 0) return PAIR_YELLOW;
    if ((e0 > 0 && e1 > 0 && e2 > 0) || (e0 < 0 && e1 < 0 && e2 < 0)) return PAIR_GREEN;
    return PAIR_NO;
}

/* classifyLocal, OriginShiftedFloatPredicateCore.cuh:91-124 */
static int filt_classify(const double a[9], const double b[9]) {
    const double* A[3] = {a, a + 3, a + 6};
    const double* B[3] = {b, b + 3, b + 6};
    int sB[3], sA[3];
    for (int i = 0; i < 3; ++i) sB[i] = oracle_orient3d_filter(A[0], A[1], A[2], B[i]);
    if (sB[0] != 0 && sB[0] == sB[1] && sB[1] == sB[2]) return PAIR_NO;
    for (int i = 0; i < 3; ++i) sA[i] = oracle_orient3d_filter(B[0], B[1], B[2], A[i]);
    if (sA[0] != 0 && sA[0] == sA[1] && sA[1] == sA[2]) return PAIR_NO;
    for (int i = 0; i < 3; ++i) {
        int j = (i + 1) % 3;
        int r = filt_edgetri(A[i], A[j], sA[i], sA[j], B[0], B[1], B[2]);
        if (r != PAIR_NO) return r;
    }
    for (int i = 0; i < 3; ++i) {
        int j = (i + 1) % 3;
        int r = filt_edgetri(B[i], B[j], sB[i], sB[j], A[0], A[1], A[2]);
        if (r != PAIR_NO) return r;
    }
    return PAIR_NO;
}

/* ------------------------------------------------------------- entry points */

int oracle_orient3d_sign(const double a[3], const double b[3], const double c[3], const double d[3]) {
    double pa[9] = {a[0], a[1], a[2], b[0], b[1], b[2], c[0], c[1], c[2]};
    double pb[9] = {d[0], d[1], d[2], d[0], d[1], d[2], d[0], d[1], d[2]};
    ov3 A[3], B[3];
    int e;
    pair_to_ints(pa, pb, A, B, &e);
    return orient3d_sgn(&A[0], &A[1], &A[2], &B[0]);
}

int oracle_tritri(const double a[9], const double b[9], int use_filter, oracle_tt_result* out) {
    memset(&out->info, 0, sizeof(out->info));
    out->info.filter_status = -1;
    for (int i = 0; i < 9; ++i)
        if (!isfinite(a[i]) || !isfinite(b[i])) return -1;
    if (use_filter) {
        int f = filt_classify(a, b);
        out->info.filter_status = f;
        if (f == PAIR_NO) return 0;
    }
    ov3 A[3], B[3];
    int e2;
    pair_to_ints(a, b, A, B, &e2);
    segset s;
    s.n = 0;
    int cop, deg;
    int hit = exact_pair(A, B, &cop, &deg, &s);
    if (use_filter && out->info.filter_status == PAIR_GREEN && !hit) abort(); /* unsound filter */
    out->info.intersect = hit;
    out->info.coplanar = hit ? cop : 0;
    out->info.degenerate = hit ? deg : 0;
    out->info.exp2 = e2;
    if (!hit) return 0;
    int limbs = 1;
    for (int i = 0; i < s.n; ++i)
        for (int e = 0; e < 2; ++e) {
            const ohp* P = &s.p[i][e];
            for (int k = 0; k < 3; ++k) {
                int l = ob_export_limbs(&P->c[k]);
                if (l > limbs) limbs = l;
            }
            int l = ob_export_limbs(&P->w);
            if (l > limbs) limbs = l;
        }
    if (limbs > ORACLE_EXPORT_LIMBS) return -2;
    out->info.limbs = limbs;
    out->info.nseg = s.n;
    for (int i = 0; i < s.n; ++i)
        for (int e = 0; e < 2; ++e) {
            const ohp* P = &s.p[i][e];
            for (int k = 0; k < 3; ++k) ob_export(&P->c[k], out->seg[i][4 * e + k], limbs);
            ob_export(&P->w, out->seg[i][4 * e + 3], limbs);
        }
    return 0;
}
