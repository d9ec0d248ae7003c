"""Degenerate inputs through the HIP kernels, bit-exact against the oracle's
reference-order traversal: zero-area and collinear triangles, zero-length
segments, duplicate primitives, NaN query coordinates (never found, as in
the reference, where every comparison with NaN is false), zero search
radius, queries on the surface, and a scene far from the origin."""
import numpy as np
import pytest

import oracle
from tests import meshes

pytestmark = pytest.mark.gpu
fcpw = pytest.importorskip("fcpw_amd")


def make_scene(V, F):
    dim = V.shape[1]
    scene = fcpw.Scene3D() if dim == 3 else fcpw.Scene2D()
    scene.set_object_count(1)
    scene.set_object_vertices(V, 0)
    (scene.set_object_triangles if dim == 3 else scene.set_object_line_segments)(F, 0)
    scene.build(fcpw.AggregateType.Bvh_SurfaceArea, False)
    g = fcpw.GPUScene3D("") if dim == 3 else fcpw.GPUScene2D("")
    g.transfer_to_gpu(scene)
    return scene, g


def records(hits):
    return np.frombuffer(hits._as_bytes(), dtype=fcpw.INTERACTION_DTYPE)


def assert_same(rec, ref):
    np.testing.assert_array_equal(rec["index"].view(np.int32), ref["index"])
    for f in ("d", "p", "uv", "n"):
        a, b = rec[f], ref[f]
        assert ((a == b) | (np.isnan(a) & np.isnan(b))).all(), f


def test_degenerate_triangles_and_duplicates(gpu):
    V, F = meshes.degenerate_soup()
    scene, g = make_scene(V, F)
    q = meshes.queries_around(V, 20000, seed=4)
    q[:64] = V[:64]  # exactly on vertices: d = 0
    hits = fcpw.GPUInteractionList()
    g.find_closest_points(q, np.full(len(q), np.inf, np.float32), hits, record_normals=True)
    ref = oracle.Bvh(V, F).closest_points_gpuorder(q, record_normal=True)
    assert_same(records(hits), ref)
    o, d = meshes.rays_from(V, 10000, seed=5)
    g.intersect(o, d, np.full(len(o), np.inf, np.float32), hits)
    assert_same(records(hits), oracle.Bvh(V, F).intersect_gpuorder(o, d))


def test_zero_length_segments(gpu):
    V, F = meshes.small_star(512)
    F = np.vstack([F, [[3, 3], [10, 10]], F[:5]]).astype(np.int32)
    scene, g = make_scene(V, F)
    q = meshes.queries_around(V, 20000, seed=6)
    hits = fcpw.GPUInteractionList()
    g.find_closest_points(q, np.full(len(q), np.inf, np.float32), hits)
    assert_same(records(hits), oracle.Bvh(V, F).closest_points_gpuorder(q))


def test_nan_queries_are_misses_and_zero_radius(gpu):
    V, F = meshes.small_torus()
    scene, g = make_scene(V, F)
    q = meshes.queries_around(V, 10000, seed=7)
    q[::97, 1] = np.nan
    r2 = np.full(len(q), np.inf, np.float32)
    r2[1::5] = 0.0
    q[1:200:5] = V[F[:40, 0]]  # on a vertex with r2 = 0: found at d = 0
    hits = fcpw.GPUInteractionList()
    g.find_closest_points(q, r2, hits)
    rec = records(hits)
    ref = oracle.Bvh(V, F).closest_points_gpuorder(q, r2)
    assert_same(rec, ref)
    nanq = np.isnan(q).any(1)
    assert (rec["index"][nanq] == fcpw.UINT_MAX).all()
    assert (rec["d"][1:200:5] == 0).all() and (rec["index"][1:200:5] != fcpw.UINT_MAX).all()


def test_scene_far_from_the_origin(gpu):
    V, F = meshes.small_icosphere(3)
    V = (V.astype(np.float64) + np.array([1e4, -2e4, 3e3])).astype(np.float32)
    scene, g = make_scene(V, F)
    q = meshes.queries_around(V, 20000, seed=8)
    hits = fcpw.GPUInteractionList()
    g.find_closest_points(q, np.full(len(q), np.inf, np.float32), hits)
    assert_same(records(hits), oracle.Bvh(V, F).closest_points_gpuorder(q))
