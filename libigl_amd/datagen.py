"""Seeded synthetic inputs for the winding-number workloads.

All generators are deterministic functions of their seed (numpy PCG64), so the
GPU run, the CPU baseline and the parity tests see identical inputs.

Workloads (sizes as named in the benchmark configuration):
  * exact:   level-5 icosphere (20,480 tris), radial N(0, 0.01) noise,
             100k query points uniform in [-1.5, 1.5]^3.
  * soup:    level-6 icosphere (81,920 tris) displaced by 16 Gaussian bumps;
             1M queries: half uniform in the bbox, half within 1e-3 of the
             surface along the face normal.
  * cloud:   2M oriented points area-uniform on the level-5 noisy sphere,
             exact face normals, A = total area / n; queries on a 160^3 grid
             over the bbox padded by 10%.
  * mixed:   level-8 icosphere (1,310,720 tris), 64 bumps; 10M queries, 80%
             uniform and 20% within 1e-4 of the surface.
  * serving: 64 displaced spheres with triangle counts log-uniform in
             [1k, 500k]; query batches log-uniform in [1k, 1M].
"""
from __future__ import annotations

import numpy as np

__all__ = [
    "icosphere", "noisy_sphere", "bumpy_sphere", "uv_sphere", "face_normals",
    "sample_surface", "near_surface_queries", "uniform_queries", "grid_queries",
    "exact_workload", "soup_workload", "cloud_workload", "mixed_workload",
    "serving_meshes",
]


def icosphere(level: int):
    """Unit icosphere by midpoint subdivision: 20*4^level faces,
    10*4^level+2 vertices, outward (counter-clockwise) orientation."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    V = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
                  [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    F = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    for _ in range(level):
        nv = V.shape[0]
        e = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]], axis=0)
        es = np.sort(e, axis=1)
        key = es[:, 0] * (nv + 1) + es[:, 1]
        uk, inv = np.unique(key, return_inverse=True)
        a = uk // (nv + 1)
        b = uk % (nv + 1)
        mid = V[a] + V[b]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        m = nv + inv.reshape(3, -1).T  # midpoints of edges (01, 12, 20)
        V = np.concatenate([V, mid], axis=0)
        f0, f1, f2 = F[:, 0], F[:, 1], F[:, 2]
        m01, m12, m20 = m[:, 0], m[:, 1], m[:, 2]
        F = np.concatenate([
            np.stack([f0, m01, m20], 1), np.stack([f1, m12, m01], 1),
            np.stack([f2, m20, m12], 1), np.stack([m01, m12, m20], 1)], axis=0)
    return V, F.astype(np.int32)


def noisy_sphere(level: int, sigma: float = 0.01, seed: int = 1):
    """Icosphere with every vertex scaled radially by (1 + eps), eps ~ N(0, sigma)."""
    V, F = icosphere(level)
    rng = np.random.default_rng(seed)
    V = V * (1.0 + rng.normal(0.0, sigma, size=(V.shape[0], 1)))
    return V, F


def _random_unit(rng, n):
    x = rng.normal(size=(n, 3))
    return x / np.linalg.norm(x, axis=1, keepdims=True)


def _bump(V, nbumps, rng):
    c = _random_unit(rng, nbumps)
    a = rng.uniform(-0.3, 0.3, size=nbumps)
    w = rng.uniform(0.1, 0.4, size=nbumps)
    r = np.ones(V.shape[0])
    for j in range(nbumps):
        d2 = np.sum((V - c[j]) ** 2, axis=1)
        r += a[j] * np.exp(-d2 / (w[j] * w[j]))
    return V * r[:, None]


def bumpy_sphere(level: int, nbumps: int, seed: int = 2):
    """Icosphere displaced radially: r = 1 + sum_j a_j exp(-|x-c_j|^2 / w_j^2),
    a_j ~ U(-0.3, 0.3), w_j ~ U(0.1, 0.4), c_j uniform on the sphere."""
    V, F = icosphere(level)
    rng = np.random.default_rng(seed)
    return _bump(V, nbumps, rng), F


def uv_sphere(ntris_target: int, nbumps: int = 8, seed: int = 3):
    """Latitude/longitude sphere with about ntris_target triangles, displaced by
    Gaussian bumps (used where triangle counts are not powers of four)."""
    nlat = max(3, int(round((ntris_target / 4.0) ** 0.5)))
    nlon = 2 * nlat
    th = np.pi * (np.arange(1, nlat) / nlat)
    ph = 2 * np.pi * (np.arange(nlon) / nlon)
    TH, PH = np.meshgrid(th, ph, indexing="ij")
    ring = np.stack([np.sin(TH) * np.cos(PH), np.sin(TH) * np.sin(PH), np.cos(TH)], -1).reshape(-1, 3)
    V = np.concatenate([[[0, 0, 1.0]], ring, [[0, 0, -1.0]]], axis=0)
    top, bot = 0, V.shape[0] - 1
    j = np.arange(nlon)
    jn = (j + 1) % nlon
    caps_top = np.stack([np.full(nlon, top), 1 + j, 1 + jn], 1)
    i = np.arange(nlat - 2)[:, None]
    a = 1 + i * nlon + j[None, :]
    b = 1 + (i + 1) * nlon + j[None, :]
    c = 1 + (i + 1) * nlon + jn[None, :]
    d = 1 + i * nlon + jn[None, :]
    quads = np.concatenate([np.stack([a, b, c], -1).reshape(-1, 1, 3), np.stack([a, c, d], -1).reshape(-1, 1, 3)],
                           axis=1).reshape(-1, 3)
    last = 1 + (nlat - 2) * nlon
    caps_bot = np.stack([np.full(nlon, bot), last + jn, last + j], 1)
    F = np.concatenate([caps_top, quads, caps_bot], 0)
    F = np.asarray(F, dtype=np.int32)
    rng = np.random.default_rng(seed)
    return _bump(V, nbumps, rng), F


def face_normals(V, F):
    a, b, c = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    n = np.cross(b - a, c - a)
    dbl = np.linalg.norm(n, axis=1)
    return n / np.maximum(dbl, 1e-300)[:, None], 0.5 * dbl


def sample_surface(V, F, n, rng):
    """n points area-uniform on the mesh: (points, unit face normals, face ids)."""
    N, area = face_normals(V, F)
    p = area / area.sum()
    fid = rng.choice(F.shape[0], size=n, p=p)
    r1 = np.sqrt(rng.random(n))
    r2 = rng.random(n)
    a, b, c = V[F[fid, 0]], V[F[fid, 1]], V[F[fid, 2]]
    pts = (1 - r1)[:, None] * a + (r1 * (1 - r2))[:, None] * b + (r1 * r2)[:, None] * c
    return pts, N[fid], fid


def uniform_queries(lo, hi, n, rng):
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    return lo + (hi - lo) * rng.random((n, 3))


def near_surface_queries(V, F, n, dist, rng):
    """Surface points pushed by +-dist along the face normal."""
    pts, N, _ = sample_surface(V, F, n, rng)
    s = np.where(rng.random(n) < 0.5, -1.0, 1.0)
    return pts + (s * dist)[:, None] * N


def grid_queries(lo, hi, res, pad=0.1):
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    ext = hi - lo
    lo = lo - pad * ext
    hi = hi + pad * ext
    axes = [np.linspace(lo[d], hi[d], res) for d in range(3)]
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)


def exact_workload(nq=100_000, level=5, seed=11):
    V, F = noisy_sphere(level, 0.01, seed)
    rng = np.random.default_rng(seed + 1000)
    Q = uniform_queries([-1.5] * 3, [1.5] * 3, nq, rng)
    return V, F, Q


def soup_workload(nq=1_000_000, level=6, nbumps=16, seed=22, near_dist=1e-3):
    V, F = bumpy_sphere(level, nbumps, seed)
    rng = np.random.default_rng(seed + 1000)
    nu = nq // 2
    Qu = uniform_queries(V.min(0), V.max(0), nu, rng)
    Qn = near_surface_queries(V, F, nq - nu, near_dist, rng)
    return V, F, np.concatenate([Qu, Qn], 0)


def cloud_workload(npts=2_000_000, res=160, level=5, seed=33):
    V, F = noisy_sphere(level, 0.01, 11)
    rng = np.random.default_rng(seed)
    P, N, _ = sample_surface(V, F, npts, rng)
    _, area = face_normals(V, F)
    A = np.full(npts, area.sum() / npts)
    Q = grid_queries(P.min(0), P.max(0), res, 0.1)
    return P, N, A, Q


def mixed_workload(nq=10_000_000, level=8, nbumps=64, seed=44, near_dist=1e-4):
    V, F = bumpy_sphere(level, nbumps, seed)
    rng = np.random.default_rng(seed + 1000)
    nu = (nq * 8) // 10
    Qu = uniform_queries(V.min(0), V.max(0), nu, rng)
    Qn = near_surface_queries(V, F, nq - nu, near_dist, rng)
    return V, F, np.concatenate([Qu, Qn], 0)


def serving_meshes(nmeshes=64, lo=1_000, hi=500_000, seed=55):
    """Triangle counts log-uniform in [lo, hi]; one displaced sphere each."""
    rng = np.random.default_rng(seed)
    counts = np.exp(rng.uniform(np.log(lo), np.log(hi), size=nmeshes)).astype(np.int64)
    meshes = []
    for i, c in enumerate(counts):
        V, F = uv_sphere(int(c), nbumps=8, seed=seed * 1000 + i)
        meshes.append((V, F))
    return meshes


def bumpy_curve(nseg: int, nbumps: int = 16, seed: int = 66):
    """A closed counter-clockwise curve of nseg segments: the unit circle with
    its radius displaced by nbumps Gaussian bumps, r(t) = 1 + sum a_j exp(-d^2/w_j^2)
    (d the angular distance to the bump's centre), a_j ~ U(-0.3, 0.3),
    w_j ~ U(0.1, 0.4)."""
    rng = np.random.default_rng(seed)
    t = np.linspace(0.0, 2.0 * np.pi, nseg, endpoint=False)
    r = np.ones(nseg)
    for _ in range(nbumps):
        c, a, w = rng.uniform(0, 2 * np.pi), rng.uniform(-0.3, 0.3), rng.uniform(0.1, 0.4)
        d = np.angle(np.exp(1j * (t - c)))
        r += a * np.exp(-(d * d) / (w * w))
    P = np.stack([r * np.cos(t), r * np.sin(t)], 1)
    E = np.stack([np.arange(nseg), (np.arange(nseg) + 1) % nseg], 1).astype(np.int32)
    return P, E


def curve_workload(nseg=262_144, nq=1_000_000, seed=66, near_dist=1e-3):
    """2D fast winding number workload: bumpy_curve(nseg), half the queries
    uniform in its box, half within near_dist of a segment (a random point of
    a random segment moved along the segment's normal)."""
    P, E = bumpy_curve(nseg, 16, seed)
    rng = np.random.default_rng(seed + 1000)
    nu = nq // 2
    Qu = rng.uniform(P.min(0), P.max(0), (nu, 2))
    k = rng.integers(0, nseg, nq - nu)
    a, b = P[E[k, 0]], P[E[k, 1]]
    s = rng.random((nq - nu, 1))
    d = b - a
    n = np.stack([d[:, 1], -d[:, 0]], 1) / np.maximum(np.linalg.norm(d, axis=1, keepdims=True), 1e-300)
    Qn = a + s * d + n * rng.uniform(-near_dist, near_dist, (nq - nu, 1))
    return P, E, np.concatenate([Qu, Qn], 0)
