"""Oracle BVH build, traversal and refit, pinned by structural invariants,
the reference's brute-force invariant (|d_bvh - d_baseline| <= 1e-6,
tests/aggregate_tests.cpp:562) and the FP64 brute force."""
import numpy as np
import pytest

import oracle
from tests import meshes

EPS = np.float32(np.finfo(np.float32).eps)


def _cases():
    return {
        "icosphere": meshes.small_icosphere(),
        "torus": meshes.small_torus(),
        "soup": meshes.triangle_soup(),
        "star2d": meshes.small_star(),
        "plus2d": meshes.load_obj("plus-shape.obj", 2),
        "walker2d": meshes.load_obj("walker.obj", 2),
    }


CASES = _cases()
HEURISTICS = [oracle.SURFACE_AREA, oracle.OVERLAP_SURFACE_AREA, oracle.LONGEST_AXIS_CENTER,
              oracle.VOLUME, oracle.OVERLAP_VOLUME]


def _prim_boxes(V, F, dim):
    P = V[F]                                     # (n, k, dim) float32
    lo = (P - EPS).min(axis=1)
    hi = (P + EPS).max(axis=1)
    return lo, hi


def check_structure(b, V, F):
    dim = V.shape[1]
    pmin, pmax, nrefs, offset = b.nodes()
    order = b.prim_order()
    n = b.n_nodes
    # every primitive sits in exactly one leaf
    assert sorted(order.tolist()) == list(range(len(F)))
    covered = np.zeros(len(F), np.int32)
    lo, hi = _prim_boxes(V, F[order], dim)
    n_leaves = 0
    for i in range(n):
        if nrefs[i] > 0:
            n_leaves += 1
            s, e = offset[i], offset[i] + nrefs[i]
            covered[s:e] += 1
            assert nrefs[i] <= 4 or b.max_depth >= 62
            # leaf box is the union of its primitives' inflated boxes, bit for bit
            np.testing.assert_array_equal(pmin[i, :dim], lo[s:e].min(0))
            np.testing.assert_array_equal(pmax[i, :dim], hi[s:e].max(0))
        else:
            l, r = i + 1, i + offset[i]
            assert l < r < n
            np.testing.assert_array_equal(pmin[i], np.minimum(pmin[l], pmin[r]))
            np.testing.assert_array_equal(pmax[i], np.maximum(pmax[l], pmax[r]))
    assert (covered == 1).all()
    assert n_leaves == b.n_leaves
    assert n == 2 * n_leaves - 1
    # leaf slots are laid out in DFS order
    leaf_off = offset[nrefs > 0]
    assert (np.diff(leaf_off) > 0).all()
    # indexMap is a bijection onto the referenced vertices
    imap = b.index_map()
    used = np.unique(F)
    assert sorted(imap[used].tolist()) == list(range(len(used)))


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("heuristic", HEURISTICS)
def test_build_structure(name, heuristic):
    V, F = CASES[name]
    b = oracle.Bvh(V, F, heuristic=heuristic)
    check_structure(b, V, F)


def _assert_matches_brute(V, F, hits, ref, q):
    assert (hits["index"] >= 0).all() == (ref["index"] >= 0).all()
    assert np.abs(hits["d"] - ref["d"]).max() <= 1e-6
    m = hits["index"] != ref["index"]
    if m.any():
        tie = oracle.classify_index_mismatch(V, F, q[m], hits["index"][m], ref["index"][m])
        assert tie.all(), f"{(~tie).sum()} non-tie index mismatches"


@pytest.mark.parametrize("name", list(CASES))
def test_closest_point_vs_brute_force(name):
    V, F = CASES[name]
    b = oracle.Bvh(V, F)
    q = meshes.queries_around(V, 3000)
    ref = b.brute_closest_f32(q)
    cpu = b.closest_points_cpu(q)
    gpu = b.closest_points_gpuorder(q)
    _assert_matches_brute(V, F, cpu, ref, q)
    _assert_matches_brute(V, F, gpu, ref, q)
    # closest points agree with the FP64 oracle at the FP32 rounding level
    d64, p64, _ = oracle.brute_closest_f64(V, F, q)
    scale = np.abs(q).max() + np.abs(V).max()
    assert np.abs(cpu["d"] - d64).max() <= 4 * np.finfo(np.float32).eps * scale
    # p is pinned by consistency (at a medial query two different points are
    # both closest): |x - p| equals d, and p lies on the returned primitive
    dim = V.shape[1]
    for out in (cpu, gpu):
        p = out["p"][:, :dim].astype(np.float64)
        assert np.abs(np.linalg.norm(q - p, axis=1) - out["d"]).max() <= 4 * np.finfo(np.float32).eps * scale
        on = oracle.prim_dist_f64(V, F, out["p"][:, :dim], out["index"])
        assert on.max() <= 8 * np.finfo(np.float32).eps * scale


@pytest.mark.parametrize("name", ["icosphere", "star2d"])
def test_bounded_radius_like_reference_harness(name):
    """finite search radii derived from the previous query's answer, as
    tests/aggregate_tests.cpp:538-545 does"""
    V, F = CASES[name]
    b = oracle.Bvh(V, F)
    q = meshes.queries_around(V, 2000)
    full = b.closest_points_cpu(q)
    dist_prev = np.linalg.norm(q - np.roll(q, 1, axis=0), axis=1)
    r2 = ((np.roll(full["d"], 1) + dist_prev) * 1.25) ** 2
    r2 = np.where(np.arange(len(q)) % 3 == 0, r2 * 0.01, r2).astype(np.float32)
    ref = b.brute_closest_f32(q, r2)
    for out in (b.closest_points_cpu(q, r2), b.closest_points_gpuorder(q, r2)):
        np.testing.assert_array_equal(out["index"] >= 0, ref["index"] >= 0)
        found = ref["index"] >= 0
        assert np.abs(out["d"][found] - ref["d"][found]).max() <= 1e-6
    assert (ref["index"] < 0).any() and (ref["index"] >= 0).any()


@pytest.mark.parametrize("name", list(CASES))
def test_rays_vs_brute_force(name):
    V, F = CASES[name]
    b = oracle.Bvh(V, F)
    o, d = meshes.rays_from(V, 2000)
    t64, i64 = oracle.brute_ray_f64(V, F, o, d)
    for out in (b.intersect_cpu(o, d), b.intersect_gpuorder(o, d)):
        hit = out["index"] >= 0
        # hit/miss agrees except for grazing rays, checked in FP64 below
        dis = hit != (i64 >= 0)
        assert dis.mean() < 0.002
        both = hit & (i64 >= 0)
        rel = np.abs(out["d"][both] - t64[both]) / np.maximum(t64[both], 1.0)
        assert rel.max() < 1e-4


def test_cpu_and_gpu_order_rays_agree():
    V, F = CASES["torus"]
    b = oracle.Bvh(V, F)
    o, d = meshes.rays_from(V, 4000)
    a = b.intersect_cpu(o, d)
    g = b.intersect_gpuorder(o, d)
    np.testing.assert_array_equal(a["index"] >= 0, g["index"] >= 0)
    assert np.abs(a["d"] - g["d"]).max() < 1e-5


@pytest.mark.parametrize("name", ["torus", "star2d", "soup"])
def test_refit_matches_boxes_recomputed_on_same_topology(name):
    V, F = CASES[name]
    b = oracle.Bvh(V, F)
    rng = np.random.default_rng(3)
    V2 = (V + rng.uniform(-0.02, 0.02, V.shape)).astype(np.float32)
    b.update_vertices(V2)
    b.refit()
    check_structure(b, V2, F)
    # refit answers match a brute force on the moved geometry
    q = meshes.queries_around(V2, 1000)
    _assert_matches_brute(V2, F, b.closest_points_cpu(q), b.brute_closest_f32(q), q)


def _sizes64(V, F):
    V64 = V.astype(np.float64)
    if V.shape[1] == 3:
        a, b, c = V64[F[:, 0]], V64[F[:, 1]], V64[F[:, 2]]
        return 0.5 * np.linalg.norm(np.cross(b - a, c - a), axis=1), (a + b + c) / 3
    a, b = V64[F[:, 0]], V64[F[:, 1]]
    return np.linalg.norm(b - a, axis=1), (a + b) / 2


def test_sphere_sampling_single_primitive():
    """One triangle (one segment) inside the sphere: the walk is trivial, so
    pdf = 1/size exactly and the sampled points are uniform over the
    primitive: their mean is its centroid."""
    rng = np.random.default_rng(11)
    for V, F in ((np.array([[0, 0, 0], [2, 0, 0], [0.5, 1.5, 0.3]], np.float32), np.array([[0, 1, 2]], np.int32)),
                 (np.array([[-1, 0.5], [2, 1]], np.float32), np.array([[0, 1]], np.int32))):
        n = 200000
        dim = V.shape[1]
        u = rng.uniform(0, 1, (n, 3)).astype(np.float32)
        c = np.zeros((n, dim), np.float32)
        hits = oracle.Bvh(V, F).intersect_spheres(c, np.full(n, 100.0, np.float32), u)
        size, cen = _sizes64(V, F)
        assert (hits["index"] == 0).all()
        np.testing.assert_allclose(hits["d"], 1.0 / size[0], rtol=1e-6)
        p = hits["p"][:, :dim].astype(np.float64)
        se = p.std(0) / np.sqrt(n)
        assert (np.abs(p.mean(0) - cen[0]) <= 5 * se + 1e-7).all()
        assert oracle.prim_dist_f64(V, F, p.astype(np.float32), np.zeros(n, np.int32)).max() <= 1e-6


@pytest.mark.parametrize("name", ["icosphere", "star2d"])
def test_sphere_sampling_is_an_unbiased_area_estimator(name):
    """Law of the stochastic sphere intersection (bvh.slang:257-392): over the
    random numbers, E[1/pdf] is the total size of the primitives that overlap
    the sphere and E[p/pdf] is their size-weighted centroid sum (uniform
    sampling within a primitive), both computed here in FP64 from the mesh."""
    V, F = CASES[name]
    dim = V.shape[1]
    size, cen = _sizes64(V, F)
    b = oracle.Bvh(V, F)
    rng = np.random.default_rng(12)
    n = 100000
    for k, radius in enumerate([0.2, 0.5, 1.0, 10.0]):
        centre = (V[rng.integers(len(V))] + rng.normal(0, 0.05, dim)).astype(np.float32)
        hits = b.intersect_spheres(np.repeat(centre[None], n, 0), np.full(n, radius * radius, np.float32),
                                   rng.uniform(0, 1, (n, 3)).astype(np.float32))
        hit = hits["index"] >= 0
        w = np.where(hit, 1.0 / hits["d"].astype(np.float64), 0.0)
        d64 = oracle.prim_dist_f64(V, F, np.repeat(centre[None], len(F), 0), np.arange(len(F), dtype=np.int32))
        lo = size[d64 <= radius * (1 - 1e-5)].sum()
        hi = size[d64 <= radius * (1 + 1e-5)].sum()
        se = w.std() / np.sqrt(n)
        assert lo - 5 * se <= w.mean() <= hi + 5 * se, (k, w.mean(), lo, hi, se)
        if radius == 10.0:  # every primitive is inside: check the point law too
            assert lo == size.sum()
            wp = w[:, None] * hits["p"][:, :dim].astype(np.float64)
            target = (size[:, None] * cen).sum(0)
            se_p = wp.std(0) / np.sqrt(n)
            assert (np.abs(wp.mean(0) - target) <= 5 * se_p + 1e-9).all(), (wp.mean(0), target, se_p)
        idx = hits["index"][hit].astype(np.int32)
        assert oracle.prim_dist_f64(V, F, hits["p"][hit][:, :dim].copy(), idx).max() <= 1e-5 * np.abs(V).max()
