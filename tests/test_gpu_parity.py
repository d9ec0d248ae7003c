"""GPU parity: the HIP kernels through the drop-in API against the oracle.

Bit-exact bar: the kernels follow the reference GPU traversal order and its
arithmetic (no fused multiply-add), so d, p, uv, n and index must equal the
oracle's restatement of include/fcpw/gpu/bvh.slang bit for bit.
Reference CPU path bar (include/fcpw/aggregates/bvh.inl): indices equal except
at FP64-classified ties, distances within 1e-5 relative (the stated tolerance).
"""
import numpy as np
import pytest

import oracle
from tests import meshes

pytestmark = pytest.mark.gpu

fcpw = pytest.importorskip("fcpw_amd")
REL_TOL = 1e-5


def make_scene(V, F, heuristic=None):
    dim = V.shape[1]
    scene = fcpw.Scene3D() if dim == 3 else fcpw.Scene2D()
    scene.set_object_count(1)
    scene.set_object_vertices(V, 0)
    if dim == 3:
        scene.set_object_triangles(F, 0)
    else:
        scene.set_object_line_segments(F, 0)
    scene.build(heuristic or fcpw.AggregateType.Bvh_SurfaceArea, False)
    gpu = fcpw.GPUScene3D(fcpw.slang_directory_path) if dim == 3 else fcpw.GPUScene2D(fcpw.slang_directory_path)
    gpu.transfer_to_gpu(scene)
    return scene, gpu


def as_records(hits):
    return np.frombuffer(hits._as_bytes(), dtype=fcpw.INTERACTION_DTYPE)


def assert_bit_equal(rec, ref, fields=("p", "n", "uv", "d")):
    idx = rec["index"].view(np.int32)
    bad = np.nonzero(idx != ref["index"])[0]
    assert len(bad) == 0, f"{len(bad)} index mismatches, first at {bad[:5]}"
    for f in fields:
        a, b = rec[f], ref[f]
        eq = (a == b) | (np.isnan(a) & np.isnan(b))
        assert eq.all(), f"field {f}: {np.count_nonzero(~eq)} mismatches"


def abs_floor(V, q):
    """FP32 floor of any distance computation: x - p cancels, leaving a few ulp
    of the coordinate magnitude regardless of d (4 ulp of the largest coordinate)"""
    scale = max(float(np.abs(V).max()), float(np.abs(q).max()))
    return 4.0 * float(np.finfo(np.float32).eps) * scale


def check_vs_cpu_path(V, F, q, rec, cpu):
    """reference CPU order: same hit set, indices equal except at ties,
    |d_gpu - d_cpu| <= 1e-5 d + floor (the two paths sum their FP32 dot products
    in different orders, so near d -> 0 only the absolute floor is attainable)"""
    idx = rec["index"].view(np.int32)
    np.testing.assert_array_equal(idx >= 0, cpu["index"] >= 0)
    found = idx >= 0
    err = np.abs(rec["d"][found].astype(np.float64) - cpu["d"][found])
    tol = REL_TOL * cpu["d"][found].astype(np.float64) + abs_floor(V, q)
    assert (err <= tol).all(), f"max excess {np.max(err - tol)}"
    m = idx != cpu["index"]
    if m.any():
        tie = oracle.classify_index_mismatch(V, F, q[m], idx[m], cpu["index"][m])
        assert tie.all(), f"{np.count_nonzero(~tie)} non-tie index mismatches"


MESHES = {
    "icosphere": meshes.small_icosphere,
    "torus": meshes.small_torus,
    "soup": meshes.triangle_soup,
    "bunny": lambda: meshes.load_obj("bunny.obj", 3),
    "star2d": meshes.small_star,
    "plus2d": lambda: meshes.load_obj("plus-shape.obj", 2),
    "walker2d": lambda: meshes.load_obj("walker.obj", 2),
    "spiral2d": lambda: meshes.load_obj("spiral.obj", 2),
}


@pytest.fixture(scope="module")
def cache():
    return {}


def get(cache, name):
    if name not in cache:
        V, F = MESHES[name]()
        scene, gpu = make_scene(V, F)
        cache[name] = (V, F, scene, gpu, oracle.Bvh(V, F))
    return cache[name]


@pytest.mark.parametrize("name", list(MESHES))
@pytest.mark.parametrize("normals", [False, True])
def test_closest_points_bit_exact(gpu, cache, name, normals):
    V, F, scene, g, ob = get(cache, name)
    q = meshes.queries_around(V, 20000)
    r2 = np.full(len(q), np.inf, np.float32)
    hits = fcpw.GPUInteractionList()
    g.find_closest_points(q, r2, hits, record_normals=normals)
    rec = as_records(hits)
    ref = ob.closest_points_gpuorder(q, r2, record_normal=normals)
    assert_bit_equal(rec, ref)
    check_vs_cpu_path(V, F, q, rec, ob.closest_points_cpu(q, r2))


@pytest.mark.parametrize("name", ["icosphere", "star2d", "bunny"])
def test_closest_points_finite_radius(gpu, cache, name):
    V, F, scene, g, ob = get(cache, name)
    q = meshes.queries_around(V, 8000)
    rng = np.random.default_rng(1)
    r2 = (rng.uniform(0.0, 0.3, len(q)) ** 2).astype(np.float32)
    r2[::7] = 0.0
    spheres = fcpw.GPUBoundingSphereList()
    for x, r in zip(q[:2000], r2[:2000]):
        c = fcpw.GPUFloat3(float(x[0]), float(x[1]), float(x[2]) if V.shape[1] == 3 else 0.0)
        spheres.append(fcpw.GPUBoundingSphere(c, float(r)))
    hits = fcpw.GPUInteractionList()
    g.find_closest_points(q, r2, hits)
    rec = as_records(hits)
    ref = ob.closest_points_gpuorder(q, r2)
    assert_bit_equal(rec, ref)
    assert (rec["index"] == fcpw.UINT_MAX).any() and (rec["index"] != fcpw.UINT_MAX).any()
    miss = rec["index"] == fcpw.UINT_MAX
    assert (rec["d"][miss] == np.finfo(np.float32).max).all() and (rec["p"][miss] == 0).all()
    # the GPUBoundingSphereList overload gives the same records
    hits2 = fcpw.GPUInteractionList()
    g.find_closest_points(spheres, hits2)
    np.testing.assert_array_equal(as_records(hits2), rec[:2000])


@pytest.mark.parametrize("name", list(MESHES))
@pytest.mark.parametrize("occlusion", [False, True])
def test_rays_bit_exact(gpu, cache, name, occlusion):
    V, F, scene, g, ob = get(cache, name)
    o, d = meshes.rays_from(V, 20000)
    tmax = np.full(len(o), np.inf, np.float32)
    tmax[::5] = 0.5
    hits = fcpw.GPUInteractionList()
    g.intersect(o, d, tmax, hits, check_for_occlusion=occlusion)
    rec = as_records(hits)
    ref = ob.intersect_gpuorder(o, d, tmax, occlusion=occlusion)
    assert_bit_equal(rec, ref)
    if not occlusion:
        cpu = ob.intersect_cpu(o, d, tmax)
        hit = rec["index"] != fcpw.UINT_MAX
        assert np.mean(hit != (cpu["index"] >= 0)) < 1e-3


@pytest.mark.parametrize("name", ["torus", "star2d", "bunny"])
def test_refit_after_vertex_update(gpu, name):
    V, F = MESHES[name]()
    scene, g = make_scene(V, F)
    ob = oracle.Bvh(V, F)
    rng = np.random.default_rng(7)
    q = meshes.queries_around(V, 10000)
    for step in range(3):
        V = (V + rng.uniform(-0.01, 0.01, V.shape)).astype(np.float32)
        scene.update_object_vertices(V, 0)
        g.refit(scene)
        ob.update_vertices(V)
        ob.refit()
        hits = fcpw.GPUInteractionList()
        g.find_closest_points(q, np.full(len(q), np.inf, np.float32), hits)
        assert_bit_equal(as_records(hits), ob.closest_points_gpuorder(q))
        o, d = meshes.rays_from(V, 5000, seed=step)
        hits = fcpw.GPUInteractionList()
        g.intersect(o, d, np.full(len(o), np.inf, np.float32), hits)
        assert_bit_equal(as_records(hits), ob.intersect_gpuorder(o, d))


def test_single_primitive_root_leaf(gpu):
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    F = np.array([[0, 1, 2]], np.int32)
    scene, g = make_scene(V, F)
    q = np.array([[0.25, 0.25, 2.0], [-1, -1, 0.5], [1, 1, 0]], np.float32)
    hits = fcpw.GPUInteractionList()
    g.find_closest_points(q, np.full(3, np.inf, np.float32), hits)
    np.testing.assert_allclose(hits.get_distances(), [2.0, 1.5, np.sqrt(0.5)], rtol=1e-6)
    np.testing.assert_array_equal(hits.get_indices(), [0, 0, 0])
    # refit of a tree whose root is a leaf
    V2 = V + np.float32(1.0)
    scene.update_object_vertices(V2, 0)
    g.refit(scene)
    g.find_closest_points(q + 1.0, np.full(3, np.inf, np.float32), hits)
    np.testing.assert_allclose(hits.get_distances(), [2.0, 1.5, np.sqrt(0.5)], rtol=1e-6)


def test_empty_query_batch(gpu, cache):
    V, F, scene, g, ob = get(cache, "torus")
    hits = fcpw.GPUInteractionList()
    g.find_closest_points(np.zeros((0, 3), np.float32), np.zeros(0, np.float32), hits)
    assert len(hits) == 0
    g.intersect(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32), np.zeros(0, np.float32), hits)
    assert len(hits) == 0


def test_options_do_not_change_results(gpu, cache):
    """every tuning choice gives the same bits (the traversal order never changes)"""
    V, F, scene, g, ob = get(cache, "bunny")
    q = meshes.queries_around(V, 20000, seed=5)
    r2 = np.full(len(q), np.inf, np.float32)
    configs = [dict(coop=0, block=b, stack=k, sort=srt) for b in (64, 128, 256) for k in (8, 16, 24) for srt in (0, 1)]
    configs += [dict(coop=1, stack=k, hold_min=h, sort=srt, xcds=x) for k in (0, 4, 6, 8, 12, 16, 24)
                for h in (1, 2, 8, 64) for srt in (0, 1) for x in (1, 8)]
    configs += [dict(coop=1, stack=8, hold_min=1, sort=1, xcds=8, morton_bits=b) for b in (18, 21, 24)]
    # chunked tile mappings, both kernels
    configs += [dict(coop=cp, xcds=x, xcd_chunk=c) for cp in (0, 1) for x in (3, 4, 8) for c in (1, 5, 12)]
    base = None
    for cfg in configs:
        for key, val in cfg.items():
            g.set_option(key, val)
        hits = fcpw.GPUInteractionList()
        g.find_closest_points(q, r2, hits)
        rec = as_records(hits).copy()
        if base is None:
            base = rec
            assert_bit_equal(rec, ob.closest_points_gpuorder(q, r2))
        else:
            np.testing.assert_array_equal(rec, base, err_msg=str(cfg))
    for key, val in dict(coop=1, stack=6, hold_min=1, sort=1, xcds=4, xcd_chunk=0, block=128, morton_bits=24).items():
        g.set_option(key, val)


@pytest.mark.parametrize("name", ["bunny", "star2d"])
@pytest.mark.parametrize("occlusion", [False, True])
def test_ray_options_do_not_change_results(gpu, cache, name, occlusion):
    """the cooperative ray kernel (evaluate t with no upper limit, replay the
    acceptance in leaf order) gives the bits of the one-ray-per-lane kernel
    and of the oracle, for every knob it reads"""
    V, F, scene, g, ob = get(cache, name)
    o, d = meshes.rays_from(V, 20000, seed=31)
    tmax = np.full(len(o), np.inf, np.float32)
    tmax[::7] = 0.3
    ref = ob.intersect_gpuorder(o, d, tmax, occlusion=occlusion)
    configs = [dict(ray_coop=0, block=b, stack=k, sort=srt) for b in (64, 256) for k in (8, 16) for srt in (0, 1)]
    configs += [dict(ray_coop=1, ray_stack=k, ray_hold_min=h, sort=srt, xcds=x) for k in (4, 8, 16)
                for h in (1, 8, 64) for srt in (0, 1) for x in (1, 8)]
    configs += [dict(ray_coop=rc, ray_sort_dir=sd, morton_bits=b) for rc in (0, 1) for sd in (0, 1)
                for b in (18, 21, 24)]
    for cfg in configs:
        for key, val in cfg.items():
            g.set_option(key, val)
        hits = fcpw.GPUInteractionList()
        g.intersect(o, d, tmax, hits, check_for_occlusion=occlusion)
        rec = as_records(hits)
        try:
            assert_bit_equal(rec, ref)
        except AssertionError as e:
            raise AssertionError(f"{cfg}: {e}") from None
    for key, val in dict(ray_coop=1, ray_stack=4, ray_hold_min=8, ray_sort_dir=1, stack=6, sort=1, xcds=4, block=128,
                         morton_bits=24).items():
        g.set_option(key, val)


@pytest.mark.parametrize("name", ["torus", "star2d"])
def test_cooperative_ray_counters_equal_reference_order_counts(gpu, cache, name):
    V, F, scene, g, ob = get(cache, name)
    o, d = meshes.rays_from(V, 20000, seed=32)
    g.set_option("ray_coop", 1)
    g.set_option("stats", 1)
    hits = fcpw.GPUInteractionList()
    g.intersect(o, d, np.full(len(o), np.inf, np.float32), hits)
    got = (g.info("stat_internal_nodes"), g.info("stat_prim_tests"), g.info("stat_queries"))
    ref, ni, npr = ob.intersect_gpuorder(o, d, counters=True)
    assert_bit_equal(as_records(hits), ref)
    assert got == (int(ni.sum()), int(npr.sum()), len(o))
    assert g.info("stat_leaf_passes") > 0
    g.set_option("stats", 0)


@pytest.mark.parametrize("n", [4097, 4096 + 64 * 8 - 1, 50001, 64 * 8 * 13 + 65])
def test_every_query_of_a_ragged_batch_is_written(gpu, cache, n):
    """batch sizes that leave a partial group of workgroups after the XCD
    remap: every output slot is written, by every kernel family"""
    import torch
    V, F, scene, g, ob = get(cache, "icosphere")
    q = meshes.queries_around(V, n, seed=11)
    sph = np.zeros((n, 4), np.float32)
    sph[:, :3] = q
    sph[:, 3] = np.inf
    ref = ob.closest_points_gpuorder(q)
    stream = torch.cuda.current_stream().cuda_stream
    for cfg in (dict(coop=1, sort=1), dict(coop=1, sort=0), dict(coop=0, sort=1, block=128),
                dict(coop=0, sort=0, block=256), dict(coop=1, sort=1, xcd_chunk=5),
                dict(coop=0, sort=1, block=128, xcd_chunk=3, xcds=8)):
        for key, val in cfg.items():
            g.set_option(key, val)
        out = torch.full((n, 10), -7.0, dtype=torch.float32, device="cuda")
        g.find_closest_points_device(torch.from_numpy(sph).cuda(), out)
        d = torch.full((n,), -7.0, device="cuda")
        x = torch.from_numpy(q).cuda()
        g._closest_points_compact_device(x.data_ptr(), 0, n, d.data_ptr(), 0, 0, 0, stream)
        torch.cuda.synchronize()
        rec = out.cpu().numpy().view(fcpw.INTERACTION_DTYPE).ravel()
        assert_bit_equal(rec, ref)
        np.testing.assert_array_equal(d.cpu().numpy(), ref["d"], err_msg=str(cfg))
    for key, val in dict(coop=1, stack=6, hold_min=1, sort=1, xcds=4, xcd_chunk=0, block=128, morton_bits=24).items():
        g.set_option(key, val)
    o, dr = meshes.rays_from(V, n, seed=12)
    ref = ob.intersect_gpuorder(o, dr)
    for rc in (0, 1):
        g.set_option("ray_coop", rc)
        hits = fcpw.GPUInteractionList()
        g.intersect(o, dr, np.full(n, np.inf, np.float32), hits)
        assert_bit_equal(as_records(hits), ref)
    g.set_option("ray_coop", 1)


@pytest.mark.parametrize("name", ["torus", "bunny", "star2d"])
def test_traversal_counters_equal_reference_order_counts(gpu, cache, name):
    """the counting kernels expand exactly the internal nodes and test exactly
    the primitives of the reference GPU traversal order (the byte model's I, P)"""
    V, F, scene, g, ob = get(cache, name)
    q = meshes.queries_around(V, 20000, seed=21)
    r2 = np.full(len(q), np.inf, np.float32)
    g.set_option("stats", 1)
    hits = fcpw.GPUInteractionList()
    g.find_closest_points(q, r2, hits)
    got = (g.info("stat_internal_nodes"), g.info("stat_prim_tests"), g.info("stat_queries"))
    ref, ni, npr, _ = ob.closest_points_gpuorder(q, r2, counters=True)
    assert_bit_equal(as_records(hits), ref)
    assert got == (int(ni.sum()), int(npr.sum()), len(q))
    g.set_option("stats", 1)
    o, d = meshes.rays_from(V, 20000, seed=22)
    g.intersect(o, d, np.full(len(o), np.inf, np.float32), hits)
    got = (g.info("stat_internal_nodes"), g.info("stat_prim_tests"), g.info("stat_queries"))
    ref, ni, npr = ob.intersect_gpuorder(o, d, counters=True)
    assert_bit_equal(as_records(hits), ref)
    assert got == (int(ni.sum()), int(npr.sum()), len(o))
    g.set_option("stats", 0)


def test_timing_events_bracket_each_traversal_launch(gpu, cache):
    import torch
    V, F, scene, g, ob = get(cache, "torus")
    q = meshes.queries_around(V, 50000, seed=23)
    sph = torch.zeros((len(q), 4), dtype=torch.float32, device="cuda")
    sph[:, :3] = torch.from_numpy(q).cuda()
    sph[:, 3] = float("inf")
    out = torch.empty((len(q), 10), dtype=torch.float32, device="cuda")
    g.set_option("timing", 1)
    for _ in range(3):
        g.find_closest_points_device(sph, out)
    assert g.info("kernel_launches") == 3
    ns = g.info("kernel_ns")
    assert 0 < ns < 10 ** 9
    g.set_option("timing", 0)
    g.find_closest_points_device(sph, out)
    assert g.info("kernel_launches") == 3


def test_device_entry_points_match_host(gpu, cache):
    import torch
    V, F, scene, g, ob = get(cache, "icosphere")
    q = meshes.queries_around(V, 30000)
    sph = np.zeros((len(q), 4), np.float32)
    sph[:, :3] = q
    sph[:, 3] = np.inf
    dev_in = torch.from_numpy(sph).cuda()
    dev_out = torch.empty((len(q), 10), dtype=torch.float32, device="cuda")
    g.find_closest_points_device(dev_in, dev_out)
    torch.cuda.synchronize()
    got = dev_out.cpu().numpy().view(fcpw.INTERACTION_DTYPE).ravel()
    hits = fcpw.GPUInteractionList()
    g.find_closest_points(q, sph[:, 3].copy(), hits)
    np.testing.assert_array_equal(got, as_records(hits))


def test_errors_are_reported(gpu):
    g = fcpw.GPUScene3D("")
    with pytest.raises(RuntimeError):
        g.find_closest_points(np.zeros((1, 3), np.float32), np.zeros(1, np.float32), fcpw.GPUInteractionList())
    V, F = meshes.small_icosphere()
    scene, g = make_scene(V, F)
    with pytest.raises(ValueError):
        g.set_option("block", 100)
    for key in ("coop", "ray_coop"):
        with pytest.raises(ValueError):
            g.set_option(key, 2)
        assert g.get_option(key) == 1
    for key, bad in (("ray_stack", 5), ("ray_hold_min", 0), ("sil_hold_min", 65), ("stack", 10),
                     ("xcd_chunk", 17), ("sil_xcds", 0), ("sil_stack", 16), ("inst_stack", 4), ("bwd_order", 2),
                     ("bwd_reduce", -1)):
        with pytest.raises(ValueError):
            g.set_option(key, bad)
    defaults = dict(stack=6, hold_min=1, xcds=4, xcd_chunk=0, ray_stack=4, ray_hold_min=8, sil_hold_min=3,
                    sil_stack=4, inst_stack=12, bwd_order=1, bwd_reduce=2,
                    sil_xcds=8, ray_sort_dir=1)
    assert {k: g.get_option(k) for k in defaults} == defaults
    with pytest.raises(RuntimeError, match="without silhouettes"):  # compute_silhouettes was not called
        g.find_closest_silhouette_points(np.zeros((1, 3), np.float32), np.ones(1, np.float32), np.zeros(1, np.int32),
                                         fcpw.GPUInteractionList())
    with pytest.raises(ValueError):
        g.find_closest_points(np.zeros((4, 2), np.float32), np.zeros(4, np.float32), fcpw.GPUInteractionList())
