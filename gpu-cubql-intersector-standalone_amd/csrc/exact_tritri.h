// exact_tritri.h -- exact intersection of two closed triangles.
//
// Replaces the reference's exact stage (classifyPairBigIntV2,
// src/GPU_exact_predicates/ExactBigIntPredicate.cu:381-489) and its CGAL
// fallback/segment extraction (src/CPU/YellowFilter.h:80-137,
// src/CPU/MeshIntersectionCurves.cpp:27-92) with one integer algorithm that
// decides AND constructs:
//
//   T_A ^ T_B is non-empty iff some edge of one triangle meets the other
//   triangle, and the intersection set is the convex hull of those
//   edge/triangle pieces.
//
// * Non-coplanar pairs: every piece lies on the line L = plane_A ^ plane_B,
//   so the result is the one segment (or point) spanned by the extreme piece
//   endpoints along an axis on which L's direction is non-zero.
// * Coplanar pairs: the pieces of the edges of A clipped by T_B, and of the
//   edges of B clipped by T_A (skipping B edges collinear with an A edge,
//   whose piece A already contributes), are exactly the boundary edges of the
//   overlap polygon; a single-point overlap is reported as (P, P).
// * Zero-area triangles are the segment or point they span.
//
// The algorithm is written once over a number policy P: the GPU runs it on
// int64/int128/I256 for grid coordinates (|x| <= 2^26) and on
// int128/W256/M512 for general doubles scaled to integers below 2^61; the
// host residual path runs it on arbitrary-precision integers for the pairs
// beyond that.
//
// Every result also counts the orient3d and orient2d signs it evaluated
// (npred): the cascade's predicates per stage.
#pragma once

#include "gi_int.h"

namespace gi {

struct GridPolicy {
    typedef int64_t C;  // coordinates, differences, cross products, 2D orientations
    typedef i128 B;     // orient3d values, homogeneous numerators/denominators
    typedef I256 H;     // products of two B values (comparisons only)
    static GI_HD B toB(C x) { return (B)x; }
    static GI_HD B mulCC(C a, C b) { return (B)a * (B)b; }
    static GI_HD B mulBC(B a, C b) { return a * (B)b; }
    static GI_HD H mulBB(B a, B b) { return mul128(a, b); }
    static GI_HD int cmpH(const H& a, const H& b) { return cmp256(a, b); }
    static GI_HD int sgnC(C x) { return sgn64(x); }
    static GI_HD int sgnB(B x) { return sgn128(x); }
    static GI_HD B pick3(int k, const B& a, const B& b, const B& c) { return k == 0 ? a : (k == 1 ? b : c); }
    static GI_HD int cmpAbsC(C a, C b) {
        uint64_t ua = a < 0 ? (uint64_t)0 - (uint64_t)a : (uint64_t)a;
        uint64_t ub = b < 0 ? (uint64_t)0 - (uint64_t)b : (uint64_t)b;
        return (ua > ub) - (ua < ub);
    }
    static GI_HD int cmpAbsB(B a, B b) {
        u128 ua = uabs128(a), ub = uabs128(b);
        return (ua > ub) - (ua < ub);
    }
};

// General doubles on the GPU: a pair's coordinates as integers with
// |x| < 2^61 (see gi_int.h for the widths this bound gives).
struct WidePolicy {
    typedef i128 C;
    typedef W256 B;
    typedef M512 H;
    static GI_HD B toB(C x) { return w256_from(x); }
    static GI_HD B mulCC(C a, C b) { return w256_mul(a, b); }
    static GI_HD B mulBC(const B& a, C b) { return w256_mul(a, b); }
    static GI_HD H mulBB(const B& a, const B& b) { return m512_mul(a, b); }
    static GI_HD int cmpH(const H& a, const H& b) { return m512_cmp(a, b); }
    static GI_HD int sgnC(C x) { return sgn128(x); }
    static GI_HD int sgnB(const B& x) { return w256_sgn(x); }
    // one of three values by a run-time index, limb by limb through masks: a
    // select between whole structs is folded into an indexed access of a
    // per-lane array in scratch memory
    static GI_HD B pick3(int k, const B& a, const B& b, const B& c) {
        const uint64_t ma = k == 0 ? ~0ull : 0ull, mb = k == 1 ? ~0ull : 0ull, mc = k == 2 ? ~0ull : 0ull;
        B r;
#pragma unroll
        for (int i = 0; i < 4; ++i) r.w[i] = (a.w[i] & ma) | (b.w[i] & mb) | (c.w[i] & mc);
        return r;
    }
    static GI_HD int cmpAbsC(C a, C b) {
        u128 ua = uabs128(a), ub = uabs128(b);
        return (ua > ub) - (ua < ub);
    }
    static GI_HD int cmpAbsB(const B& a, const B& b) { return w256_cmp_abs(a, b); }
};

enum { GI_SEG_MAX = 6 };

template <class P>
struct ExactTT {
    typedef typename P::C C;
    typedef typename P::B B;
    typedef typename P::H H;

    struct V3 { C v[3]; };
    struct HP { B h[3]; B w; };  // point (h/w), w > 0
    struct Rat { C n, d; };      // n/d, d > 0
    // Piece of an edge clipped by a triangle: points e0 + t (e1 - e0), t in [t0, t1]
    struct Piece { int edge; Rat t0, t1; };

    // Result of a pair without its coplanar pieces: everything a
    // non-coplanar or degenerate pair needs. With `deferred` set (coplanar
    // pair met by intersect_fast) nothing else is valid.
    struct Fast {
        int hit, coplanar, degenerate, nseg;
        int npred;              // orient3d and orient2d signs evaluated
        int point_from_piece;   // coplanar single-point overlap: index into pc
        int deferred;
        HP p, q;                // the single segment of non-coplanar/degenerate pairs
    };
    struct Result : Fast {
        Piece pc[GI_SEG_MAX];   // coplanar boundary edges (edge 0..2: A, 3..5: B)
    };

    // ------------------------------------------------------------ vectors
    // Component k of a (run-time k) by selects: an array indexed at run time
    // would be kept in per-lane scratch memory on the GPU.
    static GI_HD C comp(const V3& a, int k) { return k == 0 ? a.v[0] : (k == 1 ? a.v[1] : a.v[2]); }
    static GI_HD V3 sub(const V3& a, const V3& b) {
        V3 r;
        for (int k = 0; k < 3; ++k) r.v[k] = a.v[k] - b.v[k];
        return r;
    }
    static GI_HD V3 cross(const V3& a, const V3& b) {
        V3 r;
        r.v[0] = a.v[1] * b.v[2] - a.v[2] * b.v[1];
        r.v[1] = a.v[2] * b.v[0] - a.v[0] * b.v[2];
        r.v[2] = a.v[0] * b.v[1] - a.v[1] * b.v[0];
        return r;
    }
    static GI_HD B dotB(const V3& a, const V3& b) {
        return P::mulCC(a.v[0], b.v[0]) + P::mulCC(a.v[1], b.v[1]) + P::mulCC(a.v[2], b.v[2]);
    }
    static GI_HD bool isZero(const V3& a) {
        return P::sgnC(a.v[0]) == 0 && P::sgnC(a.v[1]) == 0 && P::sgnC(a.v[2]) == 0;
    }
    static GI_HD bool eq(const V3& a, const V3& b) {
        return P::sgnC(a.v[0] - b.v[0]) == 0 && P::sgnC(a.v[1] - b.v[1]) == 0 && P::sgnC(a.v[2] - b.v[2]) == 0;
    }
    // (a - d) . ((b - d) x (c - d))
    static GI_HD B orient3d(const V3& a, const V3& b, const V3& c, const V3& d) {
        return dotB(sub(a, d), cross(sub(b, d), sub(c, d)));
    }
    // axis to drop when projecting onto the plane of normal n
    static GI_HD int dropAxis(const V3& n) {
        int k = P::cmpAbsC(n.v[1], n.v[0]) > 0 ? 1 : 0;
        if (P::cmpAbsC(n.v[2], comp(n, k)) > 0) k = 2;
        return k;
    }
    static GI_HD C orient2d(const V3& p, const V3& q, const V3& r, int ia, int ib) {
        const C pa = comp(p, ia), pb = comp(p, ib);
        return (comp(q, ia) - pa) * (comp(r, ib) - pb) - (comp(q, ib) - pb) * (comp(r, ia) - pa);
    }
    static GI_HD int ratCmp(const Rat& a, const Rat& b) {
        B l = P::mulCC(a.n, b.d), r = P::mulCC(b.n, a.d);
        return P::sgnB(l - r);
    }

    // ------------------------------------------------------------ points
    static GI_HD HP hpInt(const V3& p) {
        HP r;
        for (int k = 0; k < 3; ++k) r.h[k] = P::toB(p.v[k]);
        r.w = P::toB(C(1));
        return r;
    }
    static GI_HD HP hpOnEdge(const V3& e0, const V3& e1, const Rat& t) {
        HP r;
        for (int k = 0; k < 3; ++k) r.h[k] = P::mulCC(e0.v[k], t.d) + P::mulCC(t.n, e1.v[k] - e0.v[k]);
        r.w = P::toB(t.d);
        return r;
    }
    // zero of an affine function with values fu at u and fv at v (opposite signs)
    static GI_HD HP hpCrossing(const V3& u, const V3& v, const B& fu, const B& fv) {
        HP r;
        for (int k = 0; k < 3; ++k) r.h[k] = P::mulBC(fu, v.v[k]) - P::mulBC(fv, u.v[k]);
        r.w = fu - fv;
        if (P::sgnB(r.w) < 0) {
            for (int k = 0; k < 3; ++k) r.h[k] = -r.h[k];
            r.w = -r.w;
        }
        return r;
    }
    static GI_HD B hcomp(const HP& a, int k) { return P::pick3(k, a.h[0], a.h[1], a.h[2]); }
    static GI_HD int hpCmpAxis(const HP& a, const HP& b, int k) {
        return P::cmpH(P::mulBB(hcomp(a, k), b.w), P::mulBB(hcomp(b, k), a.w));
    }

    // Running extremes of points along axis k.
    struct Span {
        int n;
        int npred;  // signs evaluated while collecting the pieces
        HP lo, hi;
    };
    static GI_HD void spanAdd(Span& s, const HP& x, int k) {
        if (s.n == 0) { s.lo = x; s.hi = x; s.n = 1; return; }
        if (hpCmpAxis(x, s.lo, k) < 0) s.lo = x;
        if (hpCmpAxis(x, s.hi, k) > 0) s.hi = x;
        s.n++;
    }

    // ----------------------------------------------------- 2D clipping
    // Clip the segment p + t (q - p), t in [0, 1], by the closed triangle T in
    // the (ia, ib) projection. Returns 0 (empty), 1 (one point t0), 2 (t0 < t1).
    static GI_HD int clip2d(const V3& p, const V3& q, const V3* T, int ia, int ib, Rat& t0, Rat& t1, int& np) {
        int sigma = P::sgnC(orient2d(T[0], T[1], T[2], ia, ib));
        np += 1;
        t0.n = C(0); t0.d = C(1);
        t1.n = C(1); t1.d = C(1);
#pragma unroll
        for (int e = 0; e < 3; ++e) {
            const V3& u = T[e];
            const V3& v = T[e == 2 ? 0 : e + 1];
            C g0 = orient2d(u, v, p, ia, ib);
            C g1 = orient2d(u, v, q, ia, ib);
            np += 2;
            if (sigma < 0) { g0 = -g0; g1 = -g1; }
            int s0 = P::sgnC(g0), s1 = P::sgnC(g1);
            if (s0 < 0 && s1 < 0) return 0;
            if (s0 >= 0 && s1 >= 0) continue;
            Rat c;
            if (s0 >= 0) {  // leaves the half-plane: upper bound g0 / (g0 - g1)
                c.n = g0; c.d = g0 - g1;
                if (ratCmp(c, t1) < 0) t1 = c;
            } else {        // enters: lower bound (-g0) / (g1 - g0)
                c.n = -g0; c.d = g1 - g0;
                if (ratCmp(c, t0) > 0) t0 = c;
            }
        }
        int c = ratCmp(t0, t1);
        if (c > 0) return 0;
        return c == 0 ? 1 : 2;
    }

    // ------------------------------------------------ degenerate shapes
    // T[i] for a run-time i by selects (constant indices keep T in registers)
    static GI_HD V3 pick(const V3* T, int i) { return i == 0 ? T[0] : (i == 1 ? T[1] : T[2]); }

    // 2 for a proper triangle; 1 for a segment T[e0]T[e1]; 0 for a point.
    static GI_HD int simplexDim(const V3* T, const V3& n, int& e0, int& e1) {
        e0 = 0; e1 = 0;
        if (!isZero(n)) return 2;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            int lo = 0, hi = 0;
            C vlo = T[0].v[k], vhi = T[0].v[k];
#pragma unroll
            for (int i = 1; i < 3; ++i) {
                if (P::sgnC(T[i].v[k] - vlo) < 0) { lo = i; vlo = T[i].v[k]; }
                if (P::sgnC(T[i].v[k] - vhi) > 0) { hi = i; vhi = T[i].v[k]; }
            }
            if (P::sgnC(vhi - vlo) != 0) { e0 = lo; e1 = hi; return 1; }
        }
        return 0;
    }

    // segment [p, q] (possibly p == q) against a proper triangle T with normal nT
    template <class R>
    static GI_HD int segTri(const V3& p, const V3& q, const V3* T, const V3& nT, R& r) {
        B op = dotB(nT, sub(p, T[0]));
        B oq = dotB(nT, sub(q, T[0]));
        int sp = P::sgnB(op), sq = P::sgnB(oq);
        r.npred += 2;
        if (sp * sq > 0) return 0;
        if (sp == 0 && sq == 0) {
            int ia, ib, k = dropAxis(nT);
            ia = k == 0 ? 1 : 0;
            ib = k == 2 ? 1 : 2;
            Rat t0, t1;
            int kind = clip2d(p, q, T, ia, ib, t0, t1, r.npred);
            if (kind == 0) return 0;
            if (eq(p, q)) { r.p = hpInt(p); r.q = r.p; }
            else { r.p = hpOnEdge(p, q, t0); r.q = kind == 1 ? r.p : hpOnEdge(p, q, t1); }
            r.nseg = 1;
            return 1;
        }
        int e0 = P::sgnB(orient3d(p, q, T[0], T[1]));
        int e1 = P::sgnB(orient3d(p, q, T[1], T[2]));
        int e2 = P::sgnB(orient3d(p, q, T[2], T[0]));
        r.npred += 3;
        if (!((e0 >= 0 && e1 >= 0 && e2 >= 0) || (e0 <= 0 && e1 <= 0 && e2 <= 0))) return 0;
        r.p = sp == 0 ? hpInt(p) : (sq == 0 ? hpInt(q) : hpCrossing(p, q, op, oq));
        r.q = r.p;
        r.nseg = 1;
        return 1;
    }

    // closed segments [p, q] and [s, t] (each possibly a point)
    template <class R>
    static GI_HD int segSeg(const V3& p, const V3& q, const V3& s, const V3& t, R& r) {
        bool pt1 = eq(p, q), pt2 = eq(s, t);
        if (pt1 && pt2) {
            if (!eq(p, s)) return 0;
            r.p = hpInt(p); r.q = r.p; r.nseg = 1;
            return 1;
        }
        if (pt1 || pt2) {
            const V3& X = pt1 ? p : s;
            const V3& a = pt1 ? s : p;
            const V3& b = pt1 ? t : q;
            V3 ab = sub(b, a), ax = sub(X, a), bx = sub(X, b);
            r.npred += 1;
            if (!isZero(cross(ab, ax))) return 0;
            if (P::sgnB(dotB(ax, ab)) < 0 || P::sgnB(dotB(bx, ab)) > 0) return 0;
            r.p = hpInt(X); r.q = r.p; r.nseg = 1;
            return 1;
        }
        r.npred += 1;
        if (P::sgnB(orient3d(p, q, s, t)) != 0) return 0;
        V3 pq = sub(q, p), st = sub(t, s);
        V3 c = cross(pq, st);
        if (isZero(c)) {
            if (!isZero(cross(pq, sub(s, p)))) return 0;  // parallel, distinct lines
            const int k = P::sgnC(pq.v[0]) != 0 ? 0 : (P::sgnC(pq.v[1]) != 0 ? 1 : 2);
            const V3& lo1 = P::sgnC(comp(p, k) - comp(q, k)) <= 0 ? p : q;
            const V3& hi1 = P::sgnC(comp(p, k) - comp(q, k)) <= 0 ? q : p;
            const V3& lo2 = P::sgnC(comp(s, k) - comp(t, k)) <= 0 ? s : t;
            const V3& hi2 = P::sgnC(comp(s, k) - comp(t, k)) <= 0 ? t : s;
            const V3& lo = P::sgnC(comp(lo1, k) - comp(lo2, k)) >= 0 ? lo1 : lo2;
            const V3& hi = P::sgnC(comp(hi1, k) - comp(hi2, k)) <= 0 ? hi1 : hi2;
            if (P::sgnC(comp(lo, k) - comp(hi, k)) > 0) return 0;
            r.p = hpInt(lo); r.q = hpInt(hi); r.nseg = 1;
            return 1;
        }
        int k = dropAxis(c);
        int ia = k == 0 ? 1 : 0, ib = k == 2 ? 1 : 2;
        C o1 = orient2d(p, q, s, ia, ib), o2 = orient2d(p, q, t, ia, ib);
        C o3 = orient2d(s, t, p, ia, ib), o4 = orient2d(s, t, q, ia, ib);
        r.npred += 4;
        if (P::sgnC(o1) * P::sgnC(o2) > 0 || P::sgnC(o3) * P::sgnC(o4) > 0) return 0;
        r.p = hpCrossing(s, t, P::toB(o1), P::toB(o2));
        r.q = r.p; r.nseg = 1;
        return 1;
    }

    // The piece of edge (Ti, Tj) of a triangle met by the (non-coplanar)
    // triangle O with normal nO: 0, 1 or 2 points on the line L, in p0 (and
    // p1). oi/oj and si/sj: orientations of Ti and Tj against O's plane and
    // their signs. np counts the signs evaluated.
    static GI_HD int edgePiece(const V3& Ti, const V3& Tj, const V3* O, const V3& nO, const B& oi, const B& oj, int si,
                               int sj, HP& p0, HP& p1, int& np) {
        if (si * sj > 0) return 0;
        if (si == 0 && sj == 0) {
            int kk = dropAxis(nO);
            int ia = kk == 0 ? 1 : 0, ib = kk == 2 ? 1 : 2;
            Rat t0, t1;
            int kind = clip2d(Ti, Tj, O, ia, ib, t0, t1, np);
            if (kind == 0) return 0;
            p0 = hpOnEdge(Ti, Tj, t0);
            if (kind == 2) p1 = hpOnEdge(Ti, Tj, t1);
            return kind;
        }
        int e0 = P::sgnB(orient3d(Ti, Tj, O[0], O[1]));
        int e1 = P::sgnB(orient3d(Ti, Tj, O[1], O[2]));
        int e2 = P::sgnB(orient3d(Ti, Tj, O[2], O[0]));
        np += 3;
        if (!((e0 >= 0 && e1 >= 0 && e2 >= 0) || (e0 <= 0 && e1 <= 0 && e2 <= 0))) return 0;
        p0 = si == 0 ? hpInt(Ti) : (sj == 0 ? hpInt(Tj) : hpCrossing(Ti, Tj, oi, oj));
        return 1;
    }

    // The pieces of the edges of T met by the (non-coplanar) triangle O, added
    // to the span along axis k; o/s: orientations of T's vertices against O's
    // plane. Unrolled, every index is a compile-time constant.
    static GI_HD void edgesAgainst(const V3* T, const V3* O, const V3& nO, const B* o, const int* s, int k,
                                   Span& span) {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const int j = i == 2 ? 0 : i + 1;
            HP p0, p1;
            const int m = edgePiece(T[i], T[j], O, nO, o[i], o[j], s[i], s[j], p0, p1, span.npred);
            if (m >= 1) spanAdd(span, p0, k);
            if (m == 2) spanAdd(span, p1, k);
        }
    }

    // ------------------------------------------------------------- main
    // Full decision and construction.
    static GI_HD void intersect(const V3* A, const V3* Bt, Result& r) { intersect_t<Result, true>(A, Bt, r); }
    // The same without the coplanar case, which is only flagged (deferred):
    // no polygon pieces, so nothing indexed at run time in per-lane storage.
    static GI_HD void intersect_fast(const V3* A, const V3* Bt, Fast& r) { intersect_t<Fast, false>(A, Bt, r); }

    template <class R, bool kCoplanar>
    static GI_HD void intersect_t(const V3* A, const V3* Bt, R& r) {
        r.hit = 0; r.coplanar = 0; r.degenerate = 0; r.nseg = 0; r.point_from_piece = -1; r.deferred = 0;
        r.npred = 0;
        V3 nA = cross(sub(A[1], A[0]), sub(A[2], A[0]));
        V3 nB = cross(sub(Bt[1], Bt[0]), sub(Bt[2], Bt[0]));
        int a0, a1, b0, b1;
        int dA = simplexDim(A, nA, a0, a1);
        int dB = simplexDim(Bt, nB, b0, b1);
        if (dA < 2 || dB < 2) {
            r.degenerate = 1;
            const V3 pa0 = pick(A, a0), pa1 = pick(A, a1), pb0 = pick(Bt, b0), pb1 = pick(Bt, b1);
            if (dA == 2) r.hit = segTri(pb0, pb1, A, nA, r);
            else if (dB == 2) r.hit = segTri(pa0, pa1, Bt, nB, r);
            else r.hit = segSeg(pa0, pa1, pb0, pb1, r);
            return;
        }
        B oB[3], oA[3];
        int sB[3], sA[3];
        for (int i = 0; i < 3; ++i) { oB[i] = dotB(nA, sub(Bt[i], A[0])); sB[i] = P::sgnB(oB[i]); }
        r.npred += 3;
        if (sB[0] == 0 && sB[1] == 0 && sB[2] == 0) {
            if constexpr (kCoplanar) coplanar(A, Bt, nA, r);
            else r.deferred = 1;
            return;
        }
        if (sB[0] * sB[1] > 0 && sB[1] * sB[2] > 0) return;
        for (int i = 0; i < 3; ++i) { oA[i] = dotB(nB, sub(A[i], Bt[0])); sA[i] = P::sgnB(oA[i]); }
        r.npred += 3;
        if (sA[0] * sA[1] > 0 && sA[1] * sA[2] > 0) return;

        // axis of the largest |component| of L's direction nA x nB
        B dir[3];
        dir[0] = P::mulCC(nA.v[1], nB.v[2]) - P::mulCC(nA.v[2], nB.v[1]);
        dir[1] = P::mulCC(nA.v[2], nB.v[0]) - P::mulCC(nA.v[0], nB.v[2]);
        dir[2] = P::mulCC(nA.v[0], nB.v[1]) - P::mulCC(nA.v[1], nB.v[0]);
        int k = P::cmpAbsB(dir[1], dir[0]) > 0 ? 1 : 0;
        if (P::cmpAbsB(dir[2], k ? dir[1] : dir[0]) > 0) k = 2;
        Span span;
        span.n = 0;
        span.npred = 0;
        edgesAgainst(A, Bt, nB, oA, sA, k, span);  // edges of A against triangle B
        edgesAgainst(Bt, A, nA, oB, sB, k, span);  // edges of B against triangle A
        r.npred += span.npred;
        if (span.n == 0) return;
        r.hit = 1;
        r.nseg = 1;
        r.p = span.lo;
        r.q = span.hi;
    }

    static GI_HD void coplanar(const V3* A, const V3* Bt, const V3& nA, Result& r) {
        r.coplanar = 1;
        int kk = dropAxis(nA);
        int ia = kk == 0 ? 1 : 0, ib = kk == 2 ? 1 : 2;
        int any = 0;
        for (int e = 0; e < 6; ++e) {
            const V3* T = e < 3 ? A : Bt;
            const V3* O = e < 3 ? Bt : A;
            int i = e % 3, j = i == 2 ? 0 : i + 1;
            if (e >= 3) {
                // skip a B edge that lies on the line of an A edge
                bool collinear = false;
                for (int f = 0; f < 3 && !collinear; ++f) {
                    int g = f == 2 ? 0 : f + 1;
                    collinear = P::sgnC(orient2d(A[f], A[g], T[i], ia, ib)) == 0 &&
                                P::sgnC(orient2d(A[f], A[g], T[j], ia, ib)) == 0;
                    r.npred += 2;
                }
                if (collinear) {
                    // its overlap with T_A still decides intersection
                    Rat t0, t1;
                    if (!any && clip2d(T[i], T[j], O, ia, ib, t0, t1, r.npred)) {
                        any = 1;
                        r.pc[GI_SEG_MAX - 1].edge = e; r.pc[GI_SEG_MAX - 1].t0 = t0; r.pc[GI_SEG_MAX - 1].t1 = t0;
                        r.point_from_piece = GI_SEG_MAX - 1;
                    }
                    continue;
                }
            }
            Rat t0, t1;
            int kind = clip2d(T[i], T[j], O, ia, ib, t0, t1, r.npred);
            if (kind == 0) continue;
            if (kind == 1) {
                if (!any) {
                    any = 1;
                    r.pc[GI_SEG_MAX - 1].edge = e; r.pc[GI_SEG_MAX - 1].t0 = t0; r.pc[GI_SEG_MAX - 1].t1 = t0;
                    r.point_from_piece = GI_SEG_MAX - 1;
                }
                continue;
            }
            any = 1;
            Piece& pc = r.pc[r.nseg++];
            pc.edge = e; pc.t0 = t0; pc.t1 = t1;
        }
        r.hit = any;
        if (any && r.nseg == 0) {
            // single-point overlap
            r.pc[0] = r.pc[r.point_from_piece];
            r.nseg = 1;
            r.point_from_piece = 0;
        } else {
            r.point_from_piece = -1;
        }
    }

    // endpoints of segment i of a result
    static GI_HD void segment(const V3* A, const V3* Bt, const Result& r, int i, HP& p, HP& q) {
        if (!r.coplanar) { p = r.p; q = r.q; return; }
        const Piece& pc = r.pc[i];
        const V3* T = pc.edge < 3 ? A : Bt;
        int a = pc.edge % 3, b = a == 2 ? 0 : a + 1;
        p = hpOnEdge(T[a], T[b], pc.t0);
        q = r.point_from_piece >= 0 ? p : hpOnEdge(T[a], T[b], pc.t1);
    }
};

}  // namespace gi
