"""BVH built on the GPU (fcpw_hip_build_bvh, GPUScene.build_on_gpu): the tree
is the reference's (bvh.inl:3-234) -- checked node for node (boxes, counts,
child offsets), primitive order per leaf included, against the oracle's
restatement of the host build for all five heuristics, 2D and 3D, on meshes
with degenerate and duplicate primitives -- and queries on it equal the
oracle's reference-order traversal bit for bit.  At 1,048,576 triangles the
device tree equals the product's host build."""
import time

import numpy as np
import pytest

import oracle
from tests import meshes
from tests.test_host_scene import HEUR, assert_same_tree

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
fcpw = pytest.importorskip("fcpw_amd")

CASES = {
    "icosphere": meshes.small_icosphere,
    "torus": meshes.small_torus,
    "soup": meshes.triangle_soup,
    "bunny": lambda: meshes.load_obj("bunny.obj", 3),
    "degenerate": meshes.degenerate_soup,
    "star2d": meshes.small_star,
    "walker2d": lambda: meshes.load_obj("walker.obj", 2),
    "spiral2d": lambda: meshes.load_obj("spiral.obj", 2),
    "tiny": lambda: (np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), np.array([[0, 1, 2]], np.int32)),
}


def scene_of(V, F):
    s = fcpw.Scene3D() if V.shape[1] == 3 else fcpw.Scene2D()
    s.set_object_count(1)
    s.set_object_vertices(V, 0)
    (s.set_object_triangles if V.shape[1] == 3 else s.set_object_line_segments)(F, 0)
    return s


def records(hits):
    return np.frombuffer(hits._as_bytes(), dtype=fcpw.INTERACTION_DTYPE)


@pytest.mark.parametrize("agg", list(HEUR))
@pytest.mark.parametrize("name", list(CASES))
def test_gpu_build_is_the_reference_tree(gpu, name, agg):
    V, F = CASES[name]()
    dim = V.shape[1]
    s = scene_of(V, F)
    g = fcpw.GPUScene3D("") if dim == 3 else fcpw.GPUScene2D("")
    g.build_on_gpu(s, agg, export_tree=True)
    ob = oracle.Bvh(V, F, heuristic=HEUR[agg])
    assert_same_tree(s, ob)
    q = meshes.queries_around(V, 5000, seed=21)
    hits = fcpw.GPUInteractionList()
    g.find_closest_points(q, np.full(len(q), np.inf, np.float32), hits)
    ref = ob.closest_points_gpuorder(q)
    rec = records(hits)
    np.testing.assert_array_equal(rec["index"].view(np.int32), ref["index"])
    np.testing.assert_array_equal(rec["d"], ref["d"])
    np.testing.assert_array_equal(rec["p"], ref["p"])


def test_gpu_build_then_refit(gpu):
    V, F = meshes.small_torus()
    s = scene_of(V, F)
    g = fcpw.GPUScene3D("")
    g.build_on_gpu(s)
    ob = oracle.Bvh(V, F)
    rng = np.random.default_rng(22)
    V2 = (V + rng.normal(0, 0.01, V.shape)).astype(np.float32)
    s.update_object_vertices(V2, 0)
    g.refit(s)
    ob.update_vertices(V2)
    ob.refit()
    q = meshes.queries_around(V2, 5000, seed=23)
    hits = fcpw.GPUInteractionList()
    g.find_closest_points(q, np.full(len(q), np.inf, np.float32), hits)
    ref = ob.closest_points_gpuorder(q)
    np.testing.assert_array_equal(records(hits)["d"], ref["d"])
    np.testing.assert_array_equal(records(hits)["index"].view(np.int32), ref["index"])


def test_gpu_build_at_scale_equals_host_build(gpu):
    from fcpw_amd import synth
    v, f = synth.torus_mesh(1024, 512)
    V, F = synth.f32(v), f.astype(np.int32)
    host = scene_of(V, F)
    t0 = time.perf_counter()
    host.build(fcpw.AggregateType.Bvh_SurfaceArea, False)
    t_host = time.perf_counter() - t0
    dev = scene_of(V, F)
    g = fcpw.GPUScene3D("")
    g.build_on_gpu(dev, fcpw.AggregateType.Bvh_SurfaceArea, export_tree=True)  # warm-up (module load)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g.build_on_gpu(dev, fcpw.AggregateType.Bvh_SurfaceArea, export_tree=True)
    t_dev = time.perf_counter() - t0
    a, b = host._bvh_nodes(), dev._bvh_nodes()
    for x, y in zip(a[:5], b[:5]):
        np.testing.assert_array_equal(x, y)
    assert a[5:] == b[5:]
    print(f"\nbuild of {len(F)} triangles: host {t_host * 1e3:.1f} ms, device (with export) {t_dev * 1e3:.1f} ms")


@pytest.mark.parametrize("name", ["torus", "walker2d"])
def test_gpu_build_with_silhouettes(gpu, name):
    """a scene marked for silhouettes: the device tree is exported, the silhouette
    data built on the host from it, and the silhouette query matches the oracle"""
    V, F = CASES[name]()
    dim = V.shape[1]
    s = scene_of(V, F)
    s.compute_silhouettes()
    g = fcpw.GPUScene3D("") if dim == 3 else fcpw.GPUScene2D("")
    g.build_on_gpu(s)
    ob = oracle.Bvh(V, F)
    ob.build_silhouettes()
    q = meshes.queries_around(V, 5000, seed=24)
    hits = fcpw.GPUInteractionList()
    r2 = np.full(len(q), np.finfo(np.float32).max, np.float32)
    g.find_closest_silhouette_points(q, r2, np.zeros(len(q), np.int32), hits)
    np.testing.assert_array_equal(records(hits)["d"], ob.closest_silhouettes(q, r2, brute=True)["d"])


def test_gpu_build_of_the_star_polygon_equals_host_build(gpu):
    from fcpw_amd import synth
    w = synth.star_refit_workload(n_queries=1 << 10)
    V, F = w["positions"], w["indices"]
    host = scene_of(V, F)
    host.build(fcpw.AggregateType.Bvh_SurfaceArea, False)
    dev = scene_of(V, F)
    fcpw.GPUScene2D("").build_on_gpu(dev, fcpw.AggregateType.Bvh_SurfaceArea, export_tree=True)
    a, b = host._bvh_nodes(), dev._bvh_nodes()
    for x, y in zip(a[:5], b[:5]):
        np.testing.assert_array_equal(x, y)
