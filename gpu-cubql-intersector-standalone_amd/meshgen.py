"""Deterministic synthetic meshes for the benchmark and parity workloads.

Everything is generated with integer hashing (SplitMix64) and IEEE basic
operations only (no libm transcendentals except where a mesh is snapped to an
integer grid afterwards), so the same seed gives the same bytes on any host.

Workloads:
  icosphere pair   subdivision level 5 (20,480 triangles), radial jitter
                   ~N(0, 1e-3) (Irwin-Hall of 12 uniforms), float64, the
                   second mesh translated by (0.6, 0.1, 0.05)
  sphere x torus   1024x512 UV sphere displaced by 0.05 * value noise and a
                   1024x512 torus (R=0.8, r=0.3) under seeded Euler angles in
                   [0, pi), both 2^20 triangles, snapped by round((x+2) * 2^22)
  height field     512x512 cells, vertices (i*2048, j*2048, h), h in {0..8};
                   copies shifted by (1024, 1024, 0) and rotated 90 degrees
                   about z: (x, y) -> (2^20 - y, x)
"""
from __future__ import annotations

import numpy as np

GRID_SCALE = float(1 << 22)

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed: int, n: int, stream: int = 0) -> np.ndarray:
    """n successive SplitMix64 outputs for (seed, stream)."""
    with np.errstate(over="ignore"):
        base = np.uint64((seed * 0x9E3779B97F4A7C15 + stream * 0xD1B54A32D192ED03) & 0xFFFFFFFFFFFFFFFF)
        x = base + np.arange(1, n + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
        z = x
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform01(seed: int, n: int, stream: int = 0) -> np.ndarray:
    return (splitmix64(seed, n, stream) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal_ih(seed: int, n: int) -> np.ndarray:
    """Approximate standard normals: sum of 12 uniforms minus 6."""
    u = uniform01(seed, 12 * n).reshape(n, 12)
    return u.sum(axis=1) - 6.0


def hash3(seed: int, i, j, k) -> np.ndarray:
    """Uniform [-1, 1) value per integer lattice point."""
    with np.errstate(over="ignore"):
        h = (np.asarray(i, dtype=np.int64).astype(np.uint64) * np.uint64(0x8DA6B343)
             ^ np.asarray(j, dtype=np.int64).astype(np.uint64) * np.uint64(0xD8163841)
             ^ np.asarray(k, dtype=np.int64).astype(np.uint64) * np.uint64(0xCB1AB31F)
             ^ np.uint64((seed * 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF))
        z = h
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (2.0 / 9007199254740992.0) - 1.0


def value_noise3(p: np.ndarray, seed: int, freq: float = 4.0) -> np.ndarray:
    """Trilinear value noise with smoothstep weights, range about [-1, 1]."""
    q = p * freq
    f = np.floor(q)
    t = q - f
    w = t * t * (3.0 - 2.0 * t)
    i = f.astype(np.int64)
    acc = np.zeros(p.shape[0])
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                wx = w[:, 0] if dx else 1.0 - w[:, 0]
                wy = w[:, 1] if dy else 1.0 - w[:, 1]
                wz = w[:, 2] if dz else 1.0 - w[:, 2]
                acc += wx * wy * wz * hash3(seed, i[:, 0] + dx, i[:, 1] + dy, i[:, 2] + dz)
    return acc


# ------------------------------------------------------------------ meshes


def icosphere(level: int):
    """Unit icosphere by midpoint subdivision and normalisation (sqrt only)."""
    t = (1.0 + np.sqrt(5.0)) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, dtype=np.float64) / np.sqrt(np.dot(p, p)) for p in v]
    faces = list(f)
    for _ in range(level):
        cache = {}
        nf = []

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = (verts[a] + verts[b]) * 0.5
                verts.append(m / np.sqrt(np.dot(m, m)))
                cache[key] = len(verts) - 1
            return cache[key]

        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    return np.array(verts), np.array(faces, dtype=np.uint32)


def jittered_icosphere(level: int, seed: int, sigma: float = 1e-3, offset=(0.0, 0.0, 0.0)):
    v, f = icosphere(level)
    r = 1.0 + sigma * normal_ih(seed, v.shape[0])
    return v * r[:, None] + np.asarray(offset, dtype=np.float64)[None, :], f


def uv_sphere(nu: int, nv: int, seed: int, amp: float = 0.05):
    """UV sphere with nu longitudes and nv latitude bands (2*nu*nv triangles,
    single pole vertices, so each pole band has nu zero-area triangles),
    radius 1 + amp * value_noise."""
    th = np.arange(1, nv) * (np.pi / nv)            # interior latitude rings
    ph = np.arange(nu) * (2.0 * np.pi / nu)
    st, ct = np.sin(th), np.cos(th)
    ring = np.stack([np.outer(st, np.cos(ph)), np.outer(st, np.sin(ph)), np.outer(ct, np.ones(nu))], axis=-1)
    pts = np.concatenate([[[0.0, 0.0, 1.0]], ring.reshape(-1, 3), [[0.0, 0.0, -1.0]]], axis=0)
    if amp != 0.0:
        pts = pts * (1.0 + amp * value_noise3(pts, seed))[:, None]
    top, bot = 0, pts.shape[0] - 1

    def vid(band, k):  # vertex on latitude line `band` (0 = north pole, nv = south pole)
        k = k % nu
        if band == 0:
            return np.full_like(k, top)
        if band == nv:
            return np.full_like(k, bot)
        return 1 + (band - 1) * nu + k

    tris = []
    k = np.arange(nu)
    for band in range(nv):
        a, b = vid(band, k), vid(band, k + 1)
        c, d = vid(band + 1, k), vid(band + 1, k + 1)
        tris.append(np.stack([a, c, d], axis=1))
        tris.append(np.stack([a, d, b], axis=1))
    return pts, np.concatenate(tris).astype(np.uint32)


def torus(nu: int, nv: int, R: float = 0.8, r: float = 0.3):
    u = np.arange(nu) * (2.0 * np.pi / nu)
    v = np.arange(nv) * (2.0 * np.pi / nv)
    cu, su = np.cos(u)[:, None], np.sin(u)[:, None]
    cv, sv = np.cos(v)[None, :], np.sin(v)[None, :]
    x = (R + r * cv) * cu
    y = (R + r * cv) * su
    z = np.broadcast_to(r * sv, x.shape)
    pts = np.stack([x, y, z], axis=-1).reshape(-1, 3)
    i = np.arange(nu)[:, None]
    j = np.arange(nv)[None, :]
    a = (i * nv + j)
    b = (((i + 1) % nu) * nv + j)
    c = (i * nv + (j + 1) % nv)
    d = (((i + 1) % nu) * nv + (j + 1) % nv)
    t1 = np.stack([a, b, d], axis=-1).reshape(-1, 3)
    t2 = np.stack([a, d, c], axis=-1).reshape(-1, 3)
    return pts, np.concatenate([t1, t2]).astype(np.uint32)


def euler_matrix(ax: float, ay: float, az: float) -> np.ndarray:
    cx, sx, cy, sy, cz, sz = np.cos(ax), np.sin(ax), np.cos(ay), np.sin(ay), np.cos(az), np.sin(az)
    return np.array([[cy * cz, sx * sy * cz - cx * sz, cx * sy * cz + sx * sz],
                     [cy * sz, sx * sy * sz + cx * cz, cx * sy * sz - sx * cz],
                     [-sy, sx * cy, cx * cy]])


def snap(pts: np.ndarray, scale: float = GRID_SCALE, offset: float = 2.0) -> np.ndarray:
    """Integer grid coordinates round((x + offset) * scale), as float64."""
    return np.rint((pts + offset) * scale)


# -------------------------------------------------------------- workloads


def icosphere_pair(level: int = 5):
    """Plain method, small: two jittered icospheres, float64 (not on a grid)."""
    a = jittered_icosphere(level, seed=1)
    b = jittered_icosphere(level, seed=2, offset=(0.6, 0.1, 0.05))
    return a, b


def sphere_torus(nu: int = 1024, nv: int = 512, seed_noise: int = 3, seed_rot: int = 4):
    """Plain method at scale: displaced sphere x rotated torus on the 2^24 grid."""
    sp, st = uv_sphere(nu, nv, seed_noise)
    tp, tt = torus(nu, nv)
    ang = uniform01(seed_rot, 3) * np.pi
    tp = tp @ euler_matrix(*ang).T
    return (snap(sp), st), (snap(tp), tt)


def height_field(cells: int = 512, seed: int = 5, spacing: int = 2048):
    n = cells + 1
    h = (splitmix64(seed, n * n) % np.uint64(9)).astype(np.float64).reshape(n, n)
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    pts = np.stack([i * float(spacing), j * float(spacing), h], axis=-1).reshape(-1, 3)
    a = (i[:-1, :-1] * n + j[:-1, :-1]).ravel()
    b = (i[1:, :-1] * n + j[1:, :-1]).ravel()
    c = (i[:-1, 1:] * n + j[:-1, 1:]).ravel()
    d = (i[1:, 1:] * n + j[1:, 1:]).ravel()
    tris = np.concatenate([np.stack([a, b, d], 1), np.stack([a, d, c], 1)]).astype(np.uint32)
    return pts, tris


def height_field_pairs(cells: int = 512, seed: int = 5, spacing: int = 2048):
    """Degeneracy-heavy stress: base vs half-cell shift, base vs 90-degree turn."""
    p, t = height_field(cells, seed, spacing)
    shifted = p + np.array([spacing / 2, spacing / 2, 0.0])
    L = float(cells * spacing)
    rot = np.stack([L - p[:, 1], p[:, 0], p[:, 2]], axis=1)
    return (p, t), (shifted, t), (rot, t)


# ----------------------------------------------------------------- serving


def gaussian(seed: int, n: int, stream: int = 0) -> np.ndarray:
    """Box-Muller normals from two SplitMix64 uniform streams."""
    u1 = uniform01(seed, n, 2 * stream)
    u2 = uniform01(seed, n, 2 * stream + 1)
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)


def quat_matrix(q) -> np.ndarray:
    """Rotation matrix of a quaternion (w, x, y, z); normalised here."""
    w, x, y, z = np.asarray(q, dtype=np.float64) / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def affine_about(R: np.ndarray, center, trans) -> np.ndarray:
    """Row-major 3x4 map p -> R (p - c) + c + t."""
    c = np.asarray(center, dtype=np.float64)
    m = np.zeros((3, 4))
    m[:, :3] = R
    m[:, 3] = c + np.asarray(trans, dtype=np.float64) - R @ c
    return m


def apply_affine(m: np.ndarray, x: np.ndarray, snap_to_grid: bool = True) -> np.ndarray:
    """The library's vertex map: ((m0 x + m1 y) + m2 z) + m3 per row, each
    operation rounded separately, then rint when snapping."""
    out = np.empty_like(x)
    for r in range(3):
        out[:, r] = ((m[r, 0] * x[:, 0] + m[r, 1] * x[:, 1]) + m[r, 2] * x[:, 2]) + m[r, 3]
    return np.rint(out) if snap_to_grid else out


def serving_workload(n_requests: int = 400, qps: float = 200.0, seed: int = 7, static_res=(1024, 1024),
                     torus_res=(512, 256), probe_res=(64, 32), torus_share: float = 0.8):
    """Static displaced sphere (2 * nu * nv triangles, 2,097,152 at 1024x1024)
    and two moving meshes registered around the grid centre 2^23: a torus
    (262,144 triangles at 512x256, R=0.8, r=0.3) and a probe sphere of radius
    0.1 (4,096 triangles at 64x32). Requests: Poisson arrivals at `qps`;
    `torus_share` of them move the torus by a rotation uniform on SO(3)
    (quaternion of 4 normals) about the centre plus a translation
    U[-1/2, 1/2]^3; the rest put the probe at a uniform direction on the unit
    sphere. All in grid units (2^22 per unit), snapped.
    Returns dict(static=(xyz, tri), meshes=[torus, probe], requests=[(t_s, mesh_id, m)])."""
    sp, st = uv_sphere(static_res[0], static_res[1], 3)
    tp, tt = torus(*torus_res)
    pp, pt = uv_sphere(probe_res[0], probe_res[1], 11, amp=0.0)
    c = np.full(3, 2.0 * GRID_SCALE)
    arrivals = np.cumsum(-np.log1p(-uniform01(seed, n_requests, 0)) / qps)
    kind = uniform01(seed, n_requests, 1) >= torus_share           # True -> probe
    q = np.stack([gaussian(seed, n_requests, 1 + k) for k in range(4)], axis=1)
    t = np.stack([uniform01(seed, n_requests, 20 + k) for k in range(3)], axis=1) - 0.5
    d = np.stack([gaussian(seed, n_requests, 5 + k) for k in range(3)], axis=1)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    reqs = []
    for i in range(n_requests):
        R = quat_matrix(q[i])
        if kind[i]:
            reqs.append((float(arrivals[i]), 1, affine_about(R, c, d[i] * GRID_SCALE)))
        else:
            reqs.append((float(arrivals[i]), 0, affine_about(R, c, t[i] * GRID_SCALE)))
    return dict(static=(snap(sp), st), meshes=[(snap(tp), tt), (snap(0.1 * pp), pt)], requests=reqs)


# ---------------------------------------------------------------- batches


def geodesic_icosphere(freq: int):
    """Icosahedron with every face cut into freq^2 triangles (20 * freq^2),
    points projected to the unit sphere. A point (u A + v B + w C) / freq,
    u + v + w = freq, is evaluated left to right; on a shared edge the zero
    weight's term is exactly 0, so both faces produce the same bits and the
    vertices are merged exactly."""
    t = (1.0 + np.sqrt(5.0)) / 2.0
    V = np.array([(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
                  (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)], dtype=np.float64)
    F = np.array([(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
                  (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
                  (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)])
    f = int(freq)
    ii, jj = np.meshgrid(np.arange(f + 1), np.arange(f + 1), indexing="ij")
    keep = ii + jj <= f
    v_w = jj[keep].astype(np.float64)             # weight of C
    u_w = (f - ii[keep] - jj[keep]).astype(np.float64)  # weight of A
    b_w = ii[keep].astype(np.float64)             # weight of B
    A, B, Cc = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    P = ((u_w[None, :, None] * A[:, None, :] + b_w[None, :, None] * B[:, None, :]) + v_w[None, :, None] * Cc[:, None, :])
    P = P / float(f)
    P = P / np.sqrt((P * P).sum(axis=2, keepdims=True))
    # local index of (i, j) within a face's point list
    lid = -np.ones((f + 1, f + 1), dtype=np.int64)
    lid[keep] = np.arange(keep.sum())
    tris = []
    for i in range(f):
        for j in range(f - i):
            tris.append((lid[i, j], lid[i + 1, j], lid[i, j + 1]))
            if i + j + 1 < f:
                tris.append((lid[i + 1, j], lid[i + 1, j + 1], lid[i, j + 1]))
    tl = np.array(tris, dtype=np.int64)
    npf = int(keep.sum())
    allp = P.reshape(-1, 3)
    allt = (tl[None, :, :] + (np.arange(20) * npf)[:, None, None]).reshape(-1, 3)
    # merge bit-identical points: unique over a 64-bit mix of the coordinate
    # bits, then check that every merged point really is identical
    bits = np.ascontiguousarray(allp).view(np.uint64).reshape(-1, 3)
    h = bits[:, 0] * np.uint64(0x9E3779B97F4A7C15)
    h = (h ^ (h >> np.uint64(29))) + bits[:, 1] * np.uint64(0xBF58476D1CE4E5B9)
    h = (h ^ (h >> np.uint64(31))) + bits[:, 2] * np.uint64(0x94D049BB133111EB)
    _, first, inv = np.unique(h, return_index=True, return_inverse=True)
    uniq = allp[first]
    assert np.array_equal(uniq[inv.reshape(-1)], allp), "hash collision between distinct points"
    return uniq, inv.reshape(-1)[allt].astype(np.uint32)


def _batch_mesh(kind: int, n: int, seed: int):
    """kind 0 icosphere (20 f^2), 1 displaced UV sphere (2 nu nv), 2 torus
    (2 nu nv, R = 0.8, r = 0.3); size nearest to n triangles."""
    if kind == 0:
        return geodesic_icosphere(max(1, int(round(np.sqrt(n / 20.0)))))
    if kind == 1:
        nv = max(2, int(round(np.sqrt(n / 4.0))))
        nu = max(3, int(round(n / (2.0 * nv))))
        p, t = uv_sphere(nu, nv, seed, amp=0.0)
        ph = seed * 0.618
        r = 1.0 + 0.05 * np.sin(4.0 * p[:, 0] + ph) * np.sin(4.0 * p[:, 1] + 2 * ph) * np.sin(4.0 * p[:, 2] + 3 * ph)
        return p * r[:, None], t
    nv = max(3, int(round(np.sqrt(n / (2.0 * 0.8 / 0.3)))))
    nu = max(3, int(round(n / (2.0 * nv))))
    return torus(nu, nv)


def batch_workload(n_pairs: int = 4096, seed: int = 6, nmin: int = 500, nmax: int = 200000):
    """Independent mesh pairs: each side an icosphere, displaced sphere or
    torus (uniform), N log-uniform in [nmin, nmax], rotated uniformly on SO(3)
    (quaternion of 4 normals); B's centre at distance U[0.3, 1.5] (radius 1)
    from A's in a uniform direction; both snapped to the 2^22-per-unit grid.
    Returns a list of ((xyzA, triA), (xyzB, triB))."""
    kinds = (uniform01(seed, 2 * n_pairs, 0) * 3).astype(np.int64).reshape(n_pairs, 2)
    sizes = np.exp(np.log(nmin) + uniform01(seed, 2 * n_pairs, 1) * np.log(nmax / nmin)).reshape(n_pairs, 2)
    q = np.stack([gaussian(seed, 2 * n_pairs, 2 + k) for k in range(4)], axis=1).reshape(n_pairs, 2, 4)
    d = np.stack([gaussian(seed, n_pairs, 6 + k) for k in range(3)], axis=1)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    dist = 0.3 + 1.2 * uniform01(seed, n_pairs, 30)
    cache = {}
    out = []
    for i in range(n_pairs):
        sides = []
        for s in range(2):
            key = (int(kinds[i, s]), int(round(sizes[i, s])))
            if key not in cache:
                cache[key] = _batch_mesh(key[0], key[1], 17 + key[1] % 97)
            p, t = cache[key]
            p = p @ quat_matrix(q[i, s]).T
            if s == 1:
                p = p + d[i] * dist[i]
            sides.append((snap(p), t))
        out.append((sides[0], sides[1]))
    return out
