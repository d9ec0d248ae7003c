"""Seeded synthetic meshes and query sets for the benchmark workloads.

Every random number comes from a counter-based splitmix64 stream, so a
(seed, count) pair always yields the same values on any machine and numpy
version.  Seeds follow the convention seed = 1000 * workload + stream, with
stream 1 = mesh, 2 = queries, 3 = rays, 4 = perturbation, 5 = sampling.

Workloads (sizes as named in the benchmark definition):
  1. icosphere subdivided 5x (20,480 tris), radial noise sigma 0.01, 1M queries
     uniform in the AABB scaled 1.5x, squared max radius inf;
  2. torus R=1, r=0.3, 1024 x 512 grid (1,048,576 tris), jitter U(-1e-3, 1e-3);
     16M queries (half uniform in the 1.2x AABB, half within 0.05 of the
     surface) and 16M rays (origins in the AABB, directions uniform on S^2);
  3. star polygon with 262,144 segments, 8M queries in [-1.5, 1.5]^2 and rays;
  4. class-I geodesic sphere of frequency 100 (200,000 tris), radial noise 0.01,
     4M queries within 0.1 of the surface;
  5. 256 meshes with log-uniform triangle (1k-500k) and query (1k-1M) counts.
"""
import math

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _finalize(z):
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


class Rng:
    """splitmix64 stream: value i = finalize(seed_state + (i + 1) * golden)."""

    def __init__(self, seed):
        with np.errstate(over="ignore"):
            self.state = _finalize(np.uint64(seed) + _GOLDEN)
        self.counter = 0

    def bits(self, n):
        with np.errstate(over="ignore"):
            i = np.arange(self.counter + 1, self.counter + 1 + n, dtype=np.uint64)
            self.counter += n
            return _finalize(self.state + i * _GOLDEN)

    def uniform(self, n, lo=0.0, hi=1.0):
        """float64 in [lo, hi) from the top 53 bits."""
        u = (self.bits(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u

    def normal(self, n):
        """Box-Muller; u1 is moved to (0, 1] so the log is finite."""
        m = (n + 1) // 2
        u1 = 1.0 - self.uniform(m)
        u2 = self.uniform(m)
        r = np.sqrt(-2.0 * np.log(u1))
        z = np.concatenate([r * np.cos(2.0 * np.pi * u2), r * np.sin(2.0 * np.pi * u2)])
        return z[:n]

    def sphere(self, n, dim=3):
        """uniform on S^(dim-1) by normalized Gaussians"""
        g = self.normal(n * dim).reshape(n, dim)
        nrm = np.linalg.norm(g, axis=1, keepdims=True)
        nrm[nrm == 0] = 1.0
        return g / nrm

    def integers(self, n, high):
        return (self.uniform(n) * high).astype(np.int64).clip(0, high - 1)


def seed(workload, stream):
    return 1000 * workload + stream


# ----------------------------------------------------------------------------
# meshes
# ----------------------------------------------------------------------------
def icosahedron():
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
                  [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    return v / np.linalg.norm(v, axis=1, keepdims=True), f


def icosphere(subdivisions):
    """midpoint subdivision with projection to the unit sphere after each level"""
    v, f = icosahedron()
    for _ in range(subdivisions):
        edges = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        key = np.sort(edges, axis=1)
        uniq, inv = np.unique(key, axis=0, return_inverse=True)
        inv = inv.ravel()
        mid = 0.5 * (v[uniq[:, 0]] + v[uniq[:, 1]])
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = len(v)
        v = np.concatenate([v, mid])
        nf = len(f)
        m01, m12, m20 = (inv[:nf] + base, inv[nf:2 * nf] + base, inv[2 * nf:] + base)
        a, b, c = f[:, 0], f[:, 1], f[:, 2]
        f = np.concatenate([np.stack([a, m01, m20], 1), np.stack([b, m12, m01], 1),
                            np.stack([c, m20, m12], 1), np.stack([m01, m12, m20], 1)])
    return v, f


def geodesic_sphere(freq):
    """class-I geodesic sphere: each icosahedron face split into freq^2
    triangles on a barycentric grid, vertices projected to the unit sphere;
    20 freq^2 triangles and 10 freq^2 + 2 vertices"""
    v0, f0 = icosahedron()
    n = freq
    ij = [(i, j) for i in range(n + 1) for j in range(n + 1 - i)]
    loc = {p: k for k, p in enumerate(ij)}
    ij = np.array(ij, dtype=np.float64)
    tris = []
    for i in range(n):
        for j in range(n - i):
            tris.append((loc[(i, j)], loc[(i + 1, j)], loc[(i, j + 1)]))
            if i + j < n - 1:
                tris.append((loc[(i + 1, j)], loc[(i + 1, j + 1)], loc[(i, j + 1)]))
    tris = np.array(tris, dtype=np.int64)
    npf = len(ij)
    A, B, C = v0[f0[:, 0]], v0[f0[:, 1]], v0[f0[:, 2]]
    pts = (A[:, None, :] + (ij[None, :, 0:1] / n) * (B - A)[:, None, :]
           + (ij[None, :, 1:2] / n) * (C - A)[:, None, :]).reshape(-1, 3)
    pts /= np.linalg.norm(pts, axis=1, keepdims=True)
    faces = (tris[None, :, :] + (np.arange(20) * npf)[:, None, None]).reshape(-1, 3)
    key = np.round(pts * 1e9).astype(np.int64)
    uniq, first, inv = np.unique(key, axis=0, return_index=True, return_inverse=True)
    inv = inv.ravel()
    order = np.argsort(first)          # keep first-appearance vertex order
    rank = np.empty_like(order)
    rank[order] = np.arange(len(order))
    verts = pts[first[order]]
    faces = rank[inv[faces]]
    return verts, faces


def torus(U, V, R=1.0, r=0.3):
    i = np.arange(U)
    j = np.arange(V)
    th = 2.0 * np.pi * i / U
    ph = 2.0 * np.pi * j / V
    TH, PH = np.meshgrid(th, ph, indexing="ij")
    x = (R + r * np.cos(PH)) * np.cos(TH)
    y = (R + r * np.cos(PH)) * np.sin(TH)
    z = r * np.sin(PH)
    verts = np.stack([x, y, z], -1).reshape(-1, 3)
    I, J = np.meshgrid(i, j, indexing="ij")
    v00 = I * V + J
    v10 = ((I + 1) % U) * V + J
    v01 = I * V + (J + 1) % V
    v11 = ((I + 1) % U) * V + (J + 1) % V
    f = np.concatenate([np.stack([v00, v10, v11], -1).reshape(-1, 3),
                        np.stack([v00, v11, v01], -1).reshape(-1, 3)])
    return verts, f


def star_polygon(n_segments, rng, modes=64, amplitude=0.2):
    N = n_segments
    a = rng.uniform(modes) / np.arange(1, modes + 1)
    phi = rng.uniform(modes, 0.0, 2.0 * np.pi)
    th = 2.0 * np.pi * np.arange(N) / N
    s = np.zeros(N)
    for k in range(modes):
        s += a[k] * np.sin((k + 1) * th + phi[k])
    rad = 1.0 + amplitude * s / np.max(np.abs(s))
    verts = np.stack([rad * np.cos(th), rad * np.sin(th)], -1)
    idx = np.stack([np.arange(N), (np.arange(N) + 1) % N], -1)
    return verts, idx


def radial_noise(v, sigma, rng):
    r = 1.0 + sigma * rng.normal(len(v))
    return v * r[:, None]


# ----------------------------------------------------------------------------
# queries
# ----------------------------------------------------------------------------
def aabb(v):
    return v.min(0), v.max(0)


def uniform_in_box(n, lo, hi, scale, rng):
    c = 0.5 * (lo + hi)
    h = 0.5 * (hi - lo) * scale
    dim = len(lo)
    u = rng.uniform(n * dim).reshape(n, dim)
    return c + (2.0 * u - 1.0) * h


def near_surface(n, v, f, radius, rng):
    """area-weighted triangle, uniform barycentric point, then + u * radius * w
    with u ~ U(0, 1) and w uniform on S^2"""
    a, b, c = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    area = 0.5 * np.linalg.norm(np.cross(b - a, c - a), axis=1)
    cdf = np.cumsum(area)
    t = rng.uniform(n) * cdf[-1]
    tri = np.searchsorted(cdf, t, side="right").clip(0, len(f) - 1)
    r1 = np.sqrt(rng.uniform(n))
    r2 = rng.uniform(n)
    wa, wb = 1.0 - r1, r1 * (1.0 - r2)
    wc = 1.0 - wa - wb
    p = wa[:, None] * a[tri] + wb[:, None] * b[tri] + wc[:, None] * c[tri]
    u = rng.uniform(n)
    w = rng.sphere(n, 3)
    return p + (u * radius)[:, None] * w


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


# ----------------------------------------------------------------------------
# workloads
# ----------------------------------------------------------------------------
def icosphere_20k_workload(n_queries=1 << 20):
    v, f = icosphere(5)
    v = radial_noise(v, 0.01, Rng(seed(1, 1)))
    lo, hi = aabb(v)
    q = uniform_in_box(n_queries, lo, hi, 1.5, Rng(seed(1, 2)))
    return dict(positions=f32(v), indices=f.astype(np.int32), queries=f32(q),
                r2=np.full(n_queries, np.inf, np.float32))


def torus_mesh(U=1024, V=512, jitter=1e-3, stream_seed=None):
    v, f = torus(U, V)
    rng = Rng(seed(2, 1) if stream_seed is None else stream_seed)
    v = v + rng.uniform(v.size, -jitter, jitter).reshape(v.shape)
    return v, f


def torus_1m_workload(n_queries=1 << 24, n_rays=1 << 24, U=1024, V=512):
    v, f = torus_mesh(U, V)
    lo, hi = aabb(v)
    rq = Rng(seed(2, 2))
    n_u = n_queries // 2
    q = np.concatenate([uniform_in_box(n_u, lo, hi, 1.2, rq), near_surface(n_queries - n_u, v, f, 0.05, rq)])
    rr = Rng(seed(2, 3))
    o = uniform_in_box(n_rays, lo, hi, 1.0, rr)
    d = rr.sphere(n_rays, 3)
    return dict(positions=f32(v), indices=f.astype(np.int32), queries=f32(q),
                r2=np.full(n_queries, np.inf, np.float32),
                ray_o=f32(o), ray_d=f32(d), ray_tmax=np.full(n_rays, np.inf, np.float32))


def star_refit_workload(n_segments=1 << 18, n_queries=1 << 23):
    v, idx = star_polygon(n_segments, Rng(seed(3, 1)))
    rq = Rng(seed(3, 2))
    q = rq.uniform(n_queries * 2, -1.5, 1.5).reshape(-1, 2)
    rr = Rng(seed(3, 3))
    ang = rr.uniform(n_queries, 0.0, 2.0 * np.pi)
    d = np.stack([np.cos(ang), np.sin(ang)], -1)
    return dict(positions=f32(v), indices=idx.astype(np.int32), queries=f32(q),
                r2=np.full(n_queries, np.inf, np.float32), ray_o=f32(q), ray_d=f32(d),
                ray_tmax=np.full(n_queries, np.inf, np.float32))


def perturb_star(v, step, fraction=0.1, amount=0.01):
    """one refit step: a random 10% of the vertices move radially by U(-0.01, 0.01)"""
    rng = Rng(seed(3, 4) + 10000 * step)
    n = len(v)
    k = max(1, int(round(fraction * n)))
    sel = np.argsort(rng.uniform(n))[:k]
    dr = rng.uniform(k, -amount, amount)
    out = v.astype(np.float64).copy()
    rad = np.linalg.norm(out[sel], axis=1, keepdims=True)
    out[sel] = out[sel] * (1.0 + dr[:, None] / rad)
    return f32(out)


def geodesic_workload(n_queries=1 << 22, freq=100):
    v, f = geodesic_sphere(freq)
    v = radial_noise(v, 0.01, Rng(seed(4, 1)))
    q = near_surface(n_queries, v, f, 0.1, Rng(seed(4, 2)))
    return dict(positions=f32(v), indices=f.astype(np.int32), queries=f32(q),
                r2=np.full(n_queries, np.inf, np.float32))


def batch_workload(n_meshes=256, tri_range=(1e3, 5e5), query_range=(1e3, 1e6), max_meshes=None):
    rng = Rng(seed(5, 1))
    T = np.exp(rng.uniform(n_meshes, math.log(tri_range[0]), math.log(tri_range[1])))
    Q = np.exp(rng.uniform(n_meshes, math.log(query_range[0]), math.log(query_range[1]))).astype(np.int64)
    meshes = []
    count = n_meshes if max_meshes is None else min(n_meshes, max_meshes)
    for m in range(count):
        if m % 2 == 0:
            Vm = max(3, int(round(math.sqrt(T[m] / 4.0))))
            v, f = torus_mesh(2 * Vm, Vm, stream_seed=seed(5, 1) + 10000 * (m + 1))
        else:
            n = max(1, int(round(math.sqrt(T[m] / 20.0))))
            v, f = geodesic_sphere(n)
            v = radial_noise(v, 0.01, Rng(seed(5, 1) + 10000 * (m + 1)))
        lo, hi = aabb(v)
        rq = Rng(seed(5, 2) + 10000 * (m + 1))
        nq = int(Q[m])
        n_u = nq // 2
        q = np.concatenate([uniform_in_box(n_u, lo, hi, 1.2, rq), near_surface(nq - n_u, v, f, 0.05, rq)])
        meshes.append(dict(positions=f32(v), indices=f.astype(np.int32), queries=f32(q)))
    return meshes
