"""Independent exact pins for triangle/triangle intersection (pure Python).

Nothing here shares an algorithm with the oracle or the HIP path:

* decision and intersection set come from the linear program
      lambda, mu >= 0, sum(lambda) = 1, sum(mu) = 1, sum lambda_i A_i = sum mu_j B_j
  solved by enumerating its basic feasible solutions in exact rationals
  (the feasible set is a bounded polytope, so it is non-empty iff it has a
  vertex, and the images of its vertices span the intersection set);
* orient3d signs come from the plain 3x3 integer/rational determinant.

The closed triangles are point sets (a zero-area triangle is the segment or
point it spans), which is the semantics the oracle and the HIP path use.
"""
from __future__ import annotations

from fractions import Fraction as F
from itertools import combinations


def det3(r0, r1, r2):
    return (r0[0] * (r1[1] * r2[2] - r1[2] * r2[1])
            - r0[1] * (r1[0] * r2[2] - r1[2] * r2[0])
            + r0[2] * (r1[0] * r2[1] - r1[1] * r2[0]))


def orient3d(a, b, c, d):
    """Exact (a-d).((b-d)x(c-d)) on Fractions/ints."""
    a, b, c, d = [tuple(F(x) for x in p) for p in (a, b, c, d)]
    return det3([a[i] - d[i] for i in range(3)], [b[i] - d[i] for i in range(3)],
                [c[i] - d[i] for i in range(3)])


def _solve(M, r, cols):
    """Unique solution of M[:, cols] x = r, or None (inconsistent or rank
    deficient)."""
    rows = [[M[i][c] for c in cols] + [r[i]] for i in range(len(M))]
    n = len(cols)
    piv_row = 0
    where = [-1] * n
    for col in range(n):
        sel = None
        for i in range(piv_row, len(rows)):
            if rows[i][col] != 0:
                sel = i
                break
        if sel is None:
            return None  # dependent columns
        rows[piv_row], rows[sel] = rows[sel], rows[piv_row]
        pv = rows[piv_row][col]
        rows[piv_row] = [x / pv for x in rows[piv_row]]
        for i in range(len(rows)):
            if i != piv_row and rows[i][col] != 0:
                f = rows[i][col]
                rows[i] = [x - f * y for x, y in zip(rows[i], rows[piv_row])]
        where[col] = piv_row
        piv_row += 1
    for i in range(piv_row, len(rows)):
        if rows[i][n] != 0:
            return None
    return [rows[where[c]][n] for c in range(n)]


def intersection_points(A, B):
    """Images of all basic feasible solutions: a finite point set whose convex
    hull is A ^ B (empty iff the triangles do not meet)."""
    A = [tuple(F(x) for x in p) for p in A]
    B = [tuple(F(x) for x in p) for p in B]
    M = [[F(1), F(1), F(1), F(0), F(0), F(0)], [F(0), F(0), F(0), F(1), F(1), F(1)]]
    for k in range(3):
        M.append([A[0][k], A[1][k], A[2][k], -B[0][k], -B[1][k], -B[2][k]])
    r = [F(1), F(1), F(0), F(0), F(0)]
    pts = set()
    for size in range(1, 6):
        for cols in combinations(range(6), size):
            x = _solve(M, r, cols)
            if x is None or any(v < 0 for v in x):
                continue
            lam = [F(0)] * 3
            for c, v in zip(cols, x):
                if c < 3:
                    lam[c] = v
            pts.add(tuple(sum(lam[i] * A[i][k] for i in range(3)) for k in range(3)))
    return pts


def _sub(p, q):
    return tuple(p[i] - q[i] for i in range(3))


def _cross(u, v):
    return (u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0])


def canonical_segments(pts):
    """Boundary of conv(pts) as sorted canonical segments: [] for an empty set,
    [(P, P)] for a point, one segment for a collinear set, the maximal edges of
    the convex polygon otherwise."""
    pts = sorted(set(pts))
    if not pts:
        return []
    if len(pts) == 1:
        return [(pts[0], pts[0])]
    p0 = pts[0]
    d = _sub(pts[1], p0)
    normal = None
    for p in pts[2:]:
        c = _cross(d, _sub(p, p0))
        if any(x != 0 for x in c):
            normal = c
            break
    if normal is None:
        return [(pts[0], pts[-1])]  # lexicographic order is monotone on a line
    ax = max(range(3), key=lambda k: abs(normal[k]))
    ia, ib = [k for k in range(3) if k != ax]
    proj = sorted(pts, key=lambda p: (p[ia], p[ib]))

    def cross2(o, a, b):
        return (a[ia] - o[ia]) * (b[ib] - o[ib]) - (a[ib] - o[ib]) * (b[ia] - o[ia])

    lower, upper = [], []
    for p in proj:
        while len(lower) >= 2 and cross2(lower[-2], lower[-1], p) <= 0:
            lower.pop()
        lower.append(p)
    for p in reversed(proj):
        while len(upper) >= 2 and cross2(upper[-2], upper[-1], p) <= 0:
            upper.pop()
        upper.append(p)
    hull = lower[:-1] + upper[:-1]
    segs = []
    for i in range(len(hull)):
        p, q = hull[i], hull[(i + 1) % len(hull)]
        segs.append((p, q) if p <= q else (q, p))
    return sorted(segs)


def tritri(A, B):
    """(intersect, canonical segments) of two closed triangles."""
    segs = canonical_segments(intersection_points(A, B))
    return bool(segs), segs


def coplanar(A, B):
    A = [tuple(F(x) for x in p) for p in A]
    B = [tuple(F(x) for x in p) for p in B]
    n = _cross(_sub(A[1], A[0]), _sub(A[2], A[0]))
    if all(x == 0 for x in n):
        return False
    return all(orient3d(A[0], A[1], A[2], b) == 0 for b in B) and any(
        x != 0 for x in _cross(_sub(B[1], B[0]), _sub(B[2], B[0])))
