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


def test_device_entry_points_check_their_tensors_before_launch(gpu):
    """shape, dtype, contiguity and device of every tensor argument are checked
    on the host (the C ABI can only check alignment)"""
    import torch
    V, F = meshes.small_icosphere(1)
    s = fcpw.Scene3D()
    s.set_object_count(1)
    s.set_object_vertices(V, 0)
    s.set_object_triangles(F, 0)
    s.build(fcpw.AggregateType.Bvh_SurfaceArea, False)
    g = fcpw.GPUScene3D("")
    g.transfer_to_gpu(s)
    n = 64
    sph = torch.zeros((n, 4), device="cuda")
    out = torch.zeros((n, 10), device="cuda")
    rays = torch.zeros((n, 10), device="cuda")
    rnd = torch.zeros((n, 3), device="cuda")
    bad = [
        lambda: g.find_closest_points_device(sph, torch.zeros((n - 1, 10), device="cuda")),
        lambda: g.find_closest_points_device(sph, torch.zeros((n, 9), device="cuda")),
        lambda: g.find_closest_points_device(sph, torch.zeros((n, 10), dtype=torch.float64, device="cuda")),
        lambda: g.find_closest_points_device(sph, torch.zeros((10, n), device="cuda").t()),
        lambda: g.find_closest_points_device(sph.cpu(), out),
        lambda: g.find_closest_points_device(sph[:, :3], out),
        lambda: g.find_closest_points_device(sph, out, object_index=torch.zeros(n, dtype=torch.int64, device="cuda")),
        lambda: g.find_closest_points_device(sph, out, object_index=torch.zeros(n - 1, dtype=torch.int32, device="cuda")),
        lambda: g.intersect_device(rays[:, :9], out),
        lambda: g.intersect_device(rays, out[: n // 2]),
        lambda: g.intersect_spheres_device(sph, rnd[:, :2], out),
        lambda: g.intersect_spheres_device(sph, rnd, out.cpu()),
        lambda: g.find_closest_points_device(np.zeros((n, 4), np.float32), out),
    ]
    for call in bad:
        with pytest.raises((TypeError, ValueError)):
            call()
    g.find_closest_points_device(sph, out)
    g.intersect_spheres_device(sph, rnd, out)
    torch.cuda.synchronize()


def test_backward_sees_the_geometry_of_its_own_forward(gpu):
    """a second forward with other V before the first backward: the first
    backward restores its V (and refits), so both gradients equal those of a
    forward/backward pair run alone; without V to restore it raises"""
    import torch
    from fcpw_amd.torch_ops import closest_point_distance
    V, F = meshes.small_icosphere(2)
    s = fcpw.Scene3D()
    s.set_object_count(1)
    s.set_object_vertices(V, 0)
    s.set_object_triangles(F, 0)
    s.build(fcpw.AggregateType.Bvh_SurfaceArea, False)
    g = fcpw.GPUScene3D("")
    g.transfer_to_gpu(s)
    x = torch.from_numpy(meshes.queries_around(V, 256, seed=3)).cuda()
    V1 = torch.from_numpy(V).cuda()
    V2 = (V1 * 1.01).contiguous()

    def alone(Vv):
        a = Vv.clone().requires_grad_(True)
        d, _, _ = closest_point_distance(g, x, a)
        d.sum().backward()
        return a.grad.clone()
    ref1, ref2 = alone(V1), alone(V2)
    a1, a2 = V1.clone().requires_grad_(True), V2.clone().requires_grad_(True)
    d1, _, _ = closest_point_distance(g, x, a1)
    d2, _, _ = closest_point_distance(g, x, a2)
    d1.sum().backward()
    d2.sum().backward()
    # FP32 atomics sum in no fixed order: equal within rounding
    torch.testing.assert_close(a1.grad, ref1, rtol=1e-5, atol=1e-6)
    torch.testing.assert_close(a2.grad, ref2, rtol=1e-5, atol=1e-6)
    assert (ref1 - ref2).abs().max() > 1e-3
    xg = x.clone().requires_grad_(True)
    d3, _, _ = closest_point_distance(g, xg)  # no V: the scene's current geometry
    g._update_positions_device(V1.data_ptr(), V1.shape[0], torch.cuda.current_stream().cuda_stream)
    with pytest.raises(RuntimeError, match="no V to restore"):
        d3.sum().backward()
    torch.cuda.synchronize()


def test_triangle_regions_and_their_boundaries(gpu):
    """The triangle routine's seven Voronoi regions and the planes between them:
    the closed-form answers of the unit right triangle (vertex, edge and face
    regions) within FP32 rounding, and a grid of queries with coordinates in
    {-1, -0.5, 0, 0.25, 0.5, 1, 1.5} (many exactly on a region boundary, where
    the reference's <= and >= decide), plus queries at the vertices and edge
    midpoints of a soup nudged by +-1 ulp, bit-exact against the oracle (d, p,
    uv, n, index)"""
    from tests.test_oracle_primitives import TRI_CASES, TRI_F, TRI_V
    scene, g = make_scene(TRI_V, TRI_F)
    q = np.array([c[0] for c in TRI_CASES], np.float32)
    hits = fcpw.GPUInteractionList()
    g.find_closest_points(q, np.full(len(q), np.inf, np.float32), hits, record_normals=True)
    rec = records(hits)
    for k, (_, p, uv, d) in enumerate(TRI_CASES):
        np.testing.assert_allclose(rec["p"][k], p, atol=1e-7)
        np.testing.assert_allclose(rec["uv"][k], uv, atol=1e-7)
        assert abs(rec["d"][k] - d) <= 1e-6 * d
    vals = np.array([-1.0, -0.5, 0.0, 0.25, 0.5, 1.0, 1.5], np.float32)
    grid = np.stack(np.meshgrid(vals, vals, np.array([-1.0, 0.0, 1.0], np.float32), indexing="ij"), -1)
    q = grid.reshape(-1, 3).astype(np.float32)
    g.find_closest_points(q, np.full(len(q), np.inf, np.float32), hits, record_normals=True)
    assert_same(records(hits), oracle.Bvh(TRI_V, TRI_F).closest_points_gpuorder(q, record_normal=True))
    V, F = meshes.triangle_soup(400)
    scene, g = make_scene(V, F)
    rng = np.random.default_rng(52)
    tri = V[F]
    pts = np.concatenate([tri.reshape(-1, 3), 0.5 * (tri + np.roll(tri, 1, axis=1)).reshape(-1, 3)])
    q = np.concatenate([pts, np.nextafter(pts, pts + 1), np.nextafter(pts, pts - 1),
                        pts + rng.normal(scale=1e-3, size=pts.shape)]).astype(np.float32)
    g.find_closest_points(q, np.full(len(q), np.inf, np.float32), hits, record_normals=True)
    assert_same(records(hits), oracle.Bvh(V, F).closest_points_gpuorder(q, record_normal=True))
