"""Known-answer tests pinning the oracle's primitive routines.

Each case has a closed-form answer (Voronoi region of a unit right triangle or
segment, axis-aligned rays); the oracle is queried through a one-primitive
scene, whose root is a leaf, in all four of its paths (CPU order, GPU order,
FP32 brute force, FP64 brute force)."""
import math

import numpy as np
import pytest

import oracle

TRI_V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
TRI_F = np.array([[0, 1, 2]], np.int32)

# (query, closest point, uv = (bary_a, bary_b), distance) per Voronoi region of
# findClosestPointTriangle (triangles.inl:374-458)
TRI_CASES = [
    ((-1.0, -1.0, 0.5), (0, 0, 0), (1, 0), 1.5),            # vertex a
    ((2.0, -1.0, 0.0), (1, 0, 0), (0, 1), math.sqrt(2)),    # vertex b
    ((-1.0, 2.0, 0.0), (0, 1, 0), (0, 0), math.sqrt(2)),    # vertex c
    ((0.25, -1.0, 0.0), (0.25, 0, 0), (0.75, 0.25), 1.0),   # edge ab
    ((-1.0, 0.5, 0.0), (0, 0.5, 0), (0.5, 0.0), 1.0),       # edge ac
    ((1.0, 1.0, 0.0), (0.5, 0.5, 0), (0.0, 0.5), math.sqrt(0.5)),  # edge bc
    ((0.25, 0.25, 2.0), (0.25, 0.25, 0), (0.5, 0.25), 2.0),  # face, above
    ((0.25, 0.5, -3.0), (0.25, 0.5, 0), (0.25, 0.25), 3.0),  # face, below
]

SEG_V = np.array([[0, 0], [2, 0]], np.float32)
SEG_F = np.array([[0, 1]], np.int32)
SEG_CASES = [
    ((-1.0, 1.0), (0, 0), 0.0, math.sqrt(2)),
    ((3.0, -1.0), (2, 0), 1.0, math.sqrt(2)),
    ((0.5, 2.0), (0.5, 0), 0.25, 2.0),
    ((1.5, -0.5), (1.5, 0), 0.75, 0.5),
]


@pytest.fixture(scope="module")
def tri():
    return oracle.Bvh(TRI_V, TRI_F)


@pytest.fixture(scope="module")
def seg():
    return oracle.Bvh(SEG_V, SEG_F)


def _paths(b, q):
    return {"cpu": b.closest_points_cpu(q, record_normal=True),
            "gpu": b.closest_points_gpuorder(q, record_normal=True),
            "bf32": b.brute_closest_f32(q)}


def test_triangle_regions(tri):
    q = np.array([c[0] for c in TRI_CASES], np.float32)
    for name, out in _paths(tri, q).items():
        for k, (_, p, uv, d) in enumerate(TRI_CASES):
            np.testing.assert_allclose(out["p"][k], p, atol=1e-7, err_msg=f"{name} case {k}")
            np.testing.assert_allclose(out["uv"][k], uv, atol=1e-7, err_msg=f"{name} case {k}")
            assert abs(out["d"][k] - d) <= 1e-6 * d, (name, k)
            assert out["index"][k] == 0
    d64, p64, i64 = oracle.brute_closest_f64(TRI_V, TRI_F, q)
    np.testing.assert_allclose(d64, [c[3] for c in TRI_CASES], rtol=1e-12)
    np.testing.assert_allclose(p64, [c[1] for c in TRI_CASES], atol=1e-12)


def test_triangle_normals(tri):
    q = np.array([c[0] for c in TRI_CASES], np.float32)
    for name in ("cpu", "gpu"):
        out = _paths(tri, q)[name]
        np.testing.assert_array_equal(out["n"], np.tile([0, 0, 1], (len(q), 1)).astype(np.float32))


def test_segment_regions(seg):
    q = np.array([c[0] for c in SEG_CASES], np.float32)
    for name, out in _paths(seg, q).items():
        for k, (_, p, t, d) in enumerate(SEG_CASES):
            np.testing.assert_allclose(out["p"][k][:2], p, atol=1e-7)
            assert out["p"][k][2] == 0.0
            assert out["uv"][k][0] == pytest.approx(t, abs=1e-7)
            assert out["uv"][k][1] == 0.0
            assert abs(out["d"][k] - d) <= 1e-6 * d, (name, k)
    d64, _, _ = oracle.brute_closest_f64(SEG_V, SEG_F, q)
    np.testing.assert_allclose(d64, [c[3] for c in SEG_CASES], rtol=1e-12)


def test_segment_normal(seg):
    q = np.array([[0.5, 2.0]], np.float32)
    for out in (seg.closest_points_cpu(q, record_normal=True), seg.closest_points_gpuorder(q, record_normal=True)):
        # LineSegment::normal = (s.y, -s.x) normalized, s = pb - pa = (2, 0)
        np.testing.assert_array_equal(out["n"][0], [0.0, -1.0, 0.0])


def test_radius_bound_excludes(tri):
    q = np.array([[0.25, 0.25, 2.0]], np.float32)
    for r2, found in ((3.9, False), (4.0, True), (4.1, True)):
        r = np.array([r2], np.float32)
        for out in (tri.closest_points_cpu(q, r), tri.closest_points_gpuorder(q, r), tri.brute_closest_f32(q, r)):
            assert (out["index"][0] == 0) == found
            if not found:
                assert out["d"][0] == np.finfo(np.float32).max


def test_ray_triangle_axis_aligned(tri):
    o = np.array([[0.25, 0.25, 5.0], [0.25, 0.25, -2.0], [2.0, 2.0, 5.0], [0.1, 0.1, 1.0]], np.float32)
    d = np.array([[0, 0, -1], [0, 0, 1], [0, 0, -1], [0, 0, 1]], np.float32)
    for name, out in (("cpu", tri.intersect_cpu(o, d)), ("gpu", tri.intersect_gpuorder(o, d))):
        assert out["index"][0] == 0 and out["d"][0] == 5.0, name
        np.testing.assert_allclose(out["p"][0], [0.25, 0.25, 0.0], atol=1e-7)
        np.testing.assert_allclose(out["uv"][0], [0.5, 0.25], atol=1e-7)
        assert abs(out["n"][0][2]) == 1.0
        assert out["index"][1] == 0 and out["d"][1] == 2.0, name
        assert out["index"][2] == -1 and out["d"][2] == np.finfo(np.float32).max, name
        assert out["index"][3] == -1, name  # pointing away
    t64, i64 = oracle.brute_ray_f64(TRI_V, TRI_F, o, d)
    np.testing.assert_array_equal(i64, [0, 0, -1, -1])
    np.testing.assert_allclose(t64[:2], [5.0, 2.0])


def test_ray_tmax_and_occlusion(tri):
    o = np.array([[0.25, 0.25, 5.0]], np.float32)
    d = np.array([[0, 0, -1]], np.float32)
    for f in (tri.intersect_cpu, tri.intersect_gpuorder):
        assert f(o, d, np.array([4.5], np.float32))["index"][0] == -1
        assert f(o, d, np.array([5.0], np.float32))["index"][0] == 0
    # occlusion: CPU leaves the record at defaults (triangles.inl:247-248),
    # GPU writes the occluder's index only (bvh.slang:183-187)
    cpu = tri.intersect_cpu(o, d, occlusion=True)
    gpu = tri.intersect_gpuorder(o, d, occlusion=True)
    assert cpu["index"][0] == -1
    assert gpu["index"][0] == 0 and gpu["d"][0] == np.finfo(np.float32).max
    np.testing.assert_array_equal(gpu["p"][0], [0, 0, 0])


def test_ray_segment(seg):
    o = np.array([[1.0, -1.0], [1.0, 1.0], [3.0, -1.0], [-1.0, 0.5]], np.float32)
    d = np.array([[0, 1], [0, -1], [0, 1], [1, 0]], np.float32)
    for name, out in (("cpu", seg.intersect_cpu(o, d)), ("gpu", seg.intersect_gpuorder(o, d))):
        np.testing.assert_array_equal(out["index"], [0, 0, -1, -1], err_msg=name)
        np.testing.assert_array_equal(out["d"][:2], [1.0, 1.0])
        np.testing.assert_allclose(out["uv"][:2, 0], [0.5, 0.5])
        np.testing.assert_allclose(out["p"][:2], [[1, 0, 0], [1, 0, 0]])
    t64, i64 = oracle.brute_ray_f64(SEG_V, SEG_F, o, d)
    np.testing.assert_array_equal(i64, [0, 0, -1, -1])


def test_parallel_ray_misses(tri):
    # det == 0: the ray runs inside the triangle's plane (triangles.inl:234-235)
    o = np.array([[-1.0, 0.25, 0.0]], np.float32)
    d = np.array([[1.0, 0.0, 0.0]], np.float32)
    assert tri.intersect_cpu(o, d)["index"][0] == -1
    assert tri.intersect_gpuorder(o, d)["index"][0] == -1


def test_f64_brute_force_against_dense_sampling():
    """the FP64 oracle itself, on random triangles, against a dense barycentric
    sampling (an upper bound that must come within the grid resolution)"""
    rng = np.random.default_rng(5)
    for _ in range(20):
        v = rng.normal(size=(3, 3)).astype(np.float32)
        x = rng.normal(size=(1, 3)).astype(np.float32) * 2
        d64, p64, _ = oracle.brute_closest_f64(v, np.array([[0, 1, 2]], np.int32), x)
        n = 400
        i, j = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="ij")
        m = i + j <= n
        a, b = i[m] / n, j[m] / n
        pts = (1 - a - b)[:, None] * v[0] + a[:, None] * v[1] + b[:, None] * v[2]
        dmin = np.linalg.norm(pts - x.astype(np.float64), axis=1).min()
        edge = max(np.linalg.norm(v[1] - v[0]), np.linalg.norm(v[2] - v[0]), np.linalg.norm(v[2] - v[1]))
        assert d64[0] <= dmin + 1e-12
        assert dmin - d64[0] <= edge / n
