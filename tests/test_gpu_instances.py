"""Scenes of several objects and transformed instances on the GPU
(fcpw_hip_upload_instances and the two-level kernels): closest points, rays
and sphere intersection equal the oracle's restatement (oracle.Instances, GPU
order at both levels) bit for bit -- d, p, n, uv, primitive and object index --
on every scene of tests/scenes.py, with bounded and unbounded radii, occlusion,
after a device refit, and through the device-pointer entry points."""
import numpy as np
import pytest

import oracle
from tests import scenes
from tests.test_host_instances import make_scene

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
fcpw = pytest.importorskip("fcpw_amd")


def records(hits):
    return np.frombuffer(hits._as_bytes(), dtype=fcpw.INTERACTION_DTYPE)


def setup(name):
    objs, tr = scenes.scene(name)
    dim = objs[0][0].shape[1]
    s = make_scene(objs, tr)
    g = fcpw.GPUScene3D("") if dim == 3 else fcpw.GPUScene2D("")
    g.transfer_to_gpu(s)
    bv = [oracle.Bvh(V, F) for V, F in objs]
    I = oracle.Instances(bv, oracle.make_instances(dim, len(objs), tr))
    return objs, tr, s, g, bv, I


def assert_hits_equal(rec, obj, ref, ref_obj, fields=("d", "p", "n", "uv")):
    np.testing.assert_array_equal(rec["index"].view(np.int32), ref["index"])
    np.testing.assert_array_equal(obj, ref_obj)
    for f in fields:
        np.testing.assert_array_equal(rec[f], ref[f], err_msg=f)


@pytest.mark.parametrize("normals", [False, True])
@pytest.mark.parametrize("name", scenes.NAMES)
def test_closest_points_equal_oracle(gpu, name, normals):
    objs, tr, s, g, bv, I = setup(name)
    q = scenes.queries(objs, tr, 6000, seed=51)
    hits = fcpw.GPUInteractionList()
    obj = np.zeros(len(q), np.int32)
    g.find_closest_points(q, np.full(len(q), np.inf, np.float32), hits, record_normals=normals, object_indices=obj)
    ref, ref_obj = I.closest_points(q, record_normal=normals)
    assert_hits_equal(records(hits), obj, ref, ref_obj)


@pytest.mark.parametrize("name", scenes.NAMES + ["crowd3d"])
def test_cooperative_and_per_lane_kernels_equal_oracle(gpu, name):
    """the wave-cooperative instance kernel (one stack for the instance tree and
    the object trees, object leaves evaluated by the whole wave) and the
    one-query-per-lane kernel give the oracle's records, for every leaf-pass
    threshold, with bounded and unbounded radii"""
    objs, tr, s, g, bv, I = setup(name)
    q = scenes.queries(objs, tr, 8000, seed=57)
    r2 = np.full(len(q), np.inf, np.float32)
    r2[1::3] = np.float32(0.25)
    ref, ref_obj = I.closest_points(q, r2, record_normal=True)
    for cfg in (dict(coop=1, hold_min=1), dict(coop=1, hold_min=8), dict(coop=1, hold_min=64),
                dict(coop=1, hold_min=1, sort=0, xcds=1), dict(coop=1, inst_stack=8, sort=1, xcds=4),
                dict(coop=0, inst_stack=12)):
        for key, val in cfg.items():
            g.set_option(key, val)
        hits = fcpw.GPUInteractionList()
        obj = np.full(len(q), -7, np.int32)
        g.find_closest_points(q, r2, hits, record_normals=True, object_indices=obj)
        try:
            assert_hits_equal(records(hits), obj, ref, ref_obj)
        except AssertionError as e:
            raise AssertionError(f"{cfg}: {e}") from None
    for key, val in dict(coop=1, hold_min=1, sort=1, xcds=4).items():
        g.set_option(key, val)


@pytest.mark.parametrize("occlusion", [False, True])
@pytest.mark.parametrize("name", scenes.NAMES + ["crowd3d"])
def test_cooperative_and_per_lane_ray_kernels_equal_oracle(gpu, name, occlusion):
    """rays through the wave-cooperative instance kernel (ray leaf pass on the
    mapped rays, winner recomputed from the tMax its instance was entered with)
    and through the one-ray-per-lane kernel, against the oracle"""
    objs, tr, s, g, bv, I = setup(name)
    o, d = scenes.rays(objs, tr, 8000, seed=58)
    tmax = np.full(len(o), np.finfo(np.float32).max, np.float32)
    tmax[::4] = np.float32(2.5)
    ref, ref_obj = I.intersect(o, d, tmax, occlusion=occlusion)
    for cfg in (dict(ray_coop=1, ray_hold_min=1), dict(ray_coop=1, ray_hold_min=8), dict(ray_coop=1, ray_hold_min=64),
                dict(ray_coop=1, ray_hold_min=1, sort=0, xcds=1), dict(ray_coop=1, ray_sort_dir=0, sort=1),
                dict(ray_coop=0)):
        for key, val in cfg.items():
            g.set_option(key, val)
        hits = fcpw.GPUInteractionList()
        obj = np.full(len(o), -7, np.int32)
        g.intersect(o, d, tmax, hits, check_for_occlusion=occlusion, object_indices=obj)
        try:
            assert_hits_equal(records(hits), obj, ref, ref_obj)
        except AssertionError as e:
            raise AssertionError(f"{cfg}: {e}") from None
    for key, val in dict(ray_coop=1, ray_hold_min=8, ray_sort_dir=1, sort=1, xcds=4).items():
        g.set_option(key, val)


@pytest.mark.parametrize("name", scenes.NAMES + ["crowd3d"])
def test_lowered_radii_give_the_oracle_records(gpu, name):
    """inst_seed = B: per block of B ordered queries the 64 in its middle
    first, then the others with their radius lowered to a bound from the
    nearest of those; for every B (every query a representative, blocks that
    are not a multiple of 64, a last partial block), both kernels, radii
    unbounded, bounded and too small to find anything (representatives that
    miss), the records are the oracle's, which runs the reference's walk with
    the caller's radii"""
    objs, tr, s, g, bv, I = setup(name)
    q = scenes.queries(objs, tr, 20000, seed=59)
    rng = np.random.default_rng(60)
    r2 = np.full(len(q), np.inf, np.float32)
    r2[1::3] = rng.uniform(0.0, 0.25, len(r2[1::3])).astype(np.float32)
    r2[2::7] = np.float32(1e-8)
    ref, ref_obj = I.closest_points(q, r2, record_normal=True)
    assert (ref_obj < 0).any() and (ref_obj >= 0).any()
    try:
        for coop in (1, 0):
            for seed in (0, 64, 100, 1024, 1 << 20):
                g.set_option("coop", coop)
                g.set_option("inst_seed", seed)
                hits = fcpw.GPUInteractionList()
                obj = np.full(len(q), -7, np.int32)
                g.find_closest_points(q, r2, hits, record_normals=True, object_indices=obj)
                try:
                    assert_hits_equal(records(hits), obj, ref, ref_obj)
                except AssertionError as e:
                    raise AssertionError(f"coop={coop} inst_seed={seed}: {e}") from None
    finally:
        g.set_option("coop", 1)
        g.set_option("inst_seed", 1024)


@pytest.mark.parametrize("name", ["mixed3d", "many3d", "mixed2d"])
def test_bounded_radii_equal_oracle(gpu, name):
    objs, tr, s, g, bv, I = setup(name)
    q = scenes.queries(objs, tr, 6000, seed=52)
    r2 = np.random.default_rng(53).uniform(0.0, 0.25, len(q)).astype(np.float32)
    hits = fcpw.GPUInteractionList()
    obj = np.zeros(len(q), np.int32)
    g.find_closest_points(q, r2, hits, object_indices=obj)
    ref, ref_obj = I.closest_points(q, r2)
    assert (ref_obj < 0).any() and (ref_obj >= 0).any()
    assert_hits_equal(records(hits), obj, ref, ref_obj)


@pytest.mark.parametrize("occlusion", [False, True])
@pytest.mark.parametrize("name", ["mixed3d", "objects3d", "many3d", "single3d", "mixed2d", "many2d"])
def test_rays_equal_oracle(gpu, name, occlusion):
    objs, tr, s, g, bv, I = setup(name)
    o, d = scenes.rays(objs, tr, 6000, seed=54)
    tmax = np.full(len(o), np.finfo(np.float32).max, np.float32)
    tmax[::3] = 2.0  # some bounded rays
    hits = fcpw.GPUInteractionList()
    obj = np.zeros(len(o), np.int32)
    g.intersect(o, d, tmax, hits, check_for_occlusion=occlusion, object_indices=obj)
    ref, ref_obj = I.intersect(o, d, tmax, occlusion=occlusion)
    assert_hits_equal(records(hits), obj, ref, ref_obj)


@pytest.mark.parametrize("name", ["mixed3d", "many2d"])
def test_refit_on_device_equals_oracle(gpu, name):
    objs, tr, s, g, bv, I = setup(name)
    rng = np.random.default_rng(55)
    for o, (V, F) in enumerate(objs):
        V2 = (V + rng.uniform(-0.03, 0.03, V.shape)).astype(np.float32)
        s.update_object_vertices(V2, o)
        bv[o].update_vertices(V2)
    g.refit(s)
    I.refit()
    q = scenes.queries(objs, tr, 5000, seed=56)
    hits = fcpw.GPUInteractionList()
    obj = np.zeros(len(q), np.int32)
    g.find_closest_points(q, np.full(len(q), np.inf, np.float32), hits, object_indices=obj)
    ref, ref_obj = I.closest_points(q)
    assert_hits_equal(records(hits), obj, ref, ref_obj, fields=("d", "p", "uv"))
    o, d = scenes.rays(objs, tr, 5000, seed=57)
    rh = fcpw.GPUInteractionList()
    g.intersect(o, d, np.full(len(o), np.finfo(np.float32).max, np.float32), rh, object_indices=obj)
    rref, rref_obj = I.intersect(o, d)
    assert_hits_equal(records(rh), obj, rref, rref_obj)


def test_device_entry_points(gpu):
    objs, tr, s, g, bv, I = setup("mixed3d")
    q = scenes.queries(objs, tr, 20000, seed=58)
    dev = torch.device("cuda:0")
    sph = torch.zeros((len(q), 4), dtype=torch.float32, device=dev)
    sph[:, :3] = torch.from_numpy(q).to(dev)
    sph[:, 3] = float("inf")
    out = torch.empty((len(q), 10), dtype=torch.float32, device=dev)
    obj = torch.empty(len(q), dtype=torch.int32, device=dev)
    g.find_closest_points_device(sph, out, record_normals=True, object_index=obj)
    torch.cuda.synchronize()
    rec = out.cpu().numpy().view(fcpw.INTERACTION_DTYPE).ravel()
    ref, ref_obj = I.closest_points(q, record_normal=True)
    assert_hits_equal(rec, obj.cpu().numpy(), ref, ref_obj)
    # without object indices, and the plain entry point
    g.find_closest_points_device(sph, out)
    torch.cuda.synchronize()
    rec = out.cpu().numpy().view(fcpw.INTERACTION_DTYPE).ravel()
    np.testing.assert_array_equal(rec["d"], ref["d"])


def test_single_object_scene_reports_object_zero(gpu):
    objs, tr = scenes.scene("single3d")
    V, F = objs[0]
    s = make_scene([(V, F)], [[]])
    assert not s.is_instanced
    g = fcpw.GPUScene3D("")
    g.transfer_to_gpu(s)
    q = scenes.queries([(V, F)], [[]], 3000, seed=59)
    r2 = np.full(len(q), 0.01, np.float32)
    hits = fcpw.GPUInteractionList()
    obj = np.zeros(len(q), np.int32)
    g.find_closest_points(q, r2, hits, object_indices=obj)
    rec = records(hits)
    hit = rec["index"] != fcpw.UINT_MAX
    assert hit.any() and (~hit).any()
    np.testing.assert_array_equal(obj, np.where(hit, 0, -1))


@pytest.mark.parametrize("name", scenes.NAMES)
def test_sphere_intersection_equals_oracle(gpu, name):
    objs, tr, s, g, bv, I = setup(name)
    c = scenes.queries(objs, tr, 6000, seed=62)
    rng = np.random.default_rng(63)
    r2 = rng.uniform(0.0, 1.0, len(c)).astype(np.float32)
    rn = rng.random((len(c), 3), dtype=np.float32)
    hits = fcpw.GPUInteractionList()
    obj = np.zeros(len(c), np.int32)
    g.intersect(c, r2, rn, hits, object_indices=obj)
    ref, ref_obj = I.intersect_spheres(c, r2, rn)
    assert (ref_obj >= 0).mean() > 0.05 and (ref_obj < 0).any()
    assert_hits_equal(records(hits), obj, ref, ref_obj)
    # device form
    dev = torch.device("cuda:0")
    sph = torch.zeros((len(c), 4), dtype=torch.float32, device=dev)
    sph[:, :c.shape[1]] = torch.from_numpy(c).to(dev)
    sph[:, 3] = torch.from_numpy(r2).to(dev)
    out = torch.empty((len(c), 10), dtype=torch.float32, device=dev)
    objt = torch.empty(len(c), dtype=torch.int32, device=dev)
    g.intersect_spheres_device(sph, torch.from_numpy(rn).to(dev), out, object_index=objt)
    torch.cuda.synchronize()
    assert_hits_equal(out.cpu().numpy().view(fcpw.INTERACTION_DTYPE).ravel(), objt.cpu().numpy(), ref, ref_obj)


def test_silhouette_queries_raise(gpu):
    objs, tr, s, g, bv, I = setup("mixed3d")
    n = 4
    c = np.zeros((n, 3), np.float32)
    hits = fcpw.GPUInteractionList()
    with pytest.raises(RuntimeError, match="several objects|without silhouettes"):
        g.find_closest_silhouette_points(c, np.full(n, 1.0, np.float32), np.zeros(n, np.int32), hits)


def test_reupload_switches_between_instanced_and_single(gpu):
    objs, tr, s, g, bv, I = setup("mixed3d")
    V, F = objs[0]
    single = make_scene([(V, F)], [[]])
    g.transfer_to_gpu(single)
    q = scenes.queries([(V, F)], [[]], 3000, seed=60)
    hits = fcpw.GPUInteractionList()
    g.find_closest_points(q, np.full(len(q), np.inf, np.float32), hits)
    np.testing.assert_array_equal(records(hits)["d"], bv[0].closest_points_gpuorder(q)["d"])
    g.transfer_to_gpu(s)
    q = scenes.queries(objs, tr, 3000, seed=61)
    g.find_closest_points(q, np.full(len(q), np.inf, np.float32), hits)
    np.testing.assert_array_equal(records(hits)["d"], I.closest_points(q)[0]["d"])


def test_crowd_of_instances_equals_oracle(gpu):
    """4,000 instances (a deep instance tree) next to 50 untransformed objects"""
    objs, tr, s, g, bv, I = setup("crowd3d")
    assert g.info("n_instances") == 4050 and g.info("n_objects") == 52
    q = scenes.queries(objs, tr, 20000, seed=82)
    hits = fcpw.GPUInteractionList()
    obj = np.zeros(len(q), np.int32)
    g.find_closest_points(q, np.full(len(q), np.inf, np.float32), hits, record_normals=True, object_indices=obj)
    ref, ref_obj = I.closest_points(q, record_normal=True)
    assert_hits_equal(records(hits), obj, ref, ref_obj)
    o, d = scenes.rays(objs, tr, 20000, seed=83)
    g.intersect(o, d, np.full(len(o), np.finfo(np.float32).max, np.float32), hits, object_indices=obj)
    ref, ref_obj = I.intersect(o, d)
    assert_hits_equal(records(hits), obj, ref, ref_obj)
    r2 = np.full(len(q), 4.0, np.float32)
    rn = np.random.default_rng(84).random((len(q), 3), dtype=np.float32)
    g.intersect(q, r2, rn, hits, object_indices=obj)
    ref, ref_obj = I.intersect_spheres(q, r2, rn)
    assert_hits_equal(records(hits), obj, ref, ref_obj)


@pytest.mark.parametrize("name", ["mixed3d", "mixed2d", "crowd3d"])
def test_build_on_gpu_of_several_objects(gpu, name):
    """every object's tree built on the device: the scene's trees equal the host
    build's, and queries equal the oracle"""
    objs, tr = scenes.scene(name)
    dim = objs[0][0].shape[1]
    host = make_scene(objs, tr)
    dev = make_scene(objs, tr)
    g = fcpw.GPUScene3D("") if dim == 3 else fcpw.GPUScene2D("")
    g.build_on_gpu(dev)
    for o in range(len(objs)):
        a, b = host._object_bvh_nodes(o), dev._object_bvh_nodes(o)
        for x, y in zip(a[:5], b[:5]):
            np.testing.assert_array_equal(x, y)
    a, b = host._instance_tree_nodes(), dev._instance_tree_nodes()
    for x, y in zip(a[:5], b[:5]):
        np.testing.assert_array_equal(x, y)
    bv = [oracle.Bvh(V, F) for V, F in objs]
    I = oracle.Instances(bv, oracle.make_instances(dim, len(objs), tr))
    q = scenes.queries(objs, tr, 5000, seed=85)
    hits = fcpw.GPUInteractionList()
    obj = np.zeros(len(q), np.int32)
    g.find_closest_points(q, np.full(len(q), np.inf, np.float32), hits, object_indices=obj)
    ref, ref_obj = I.closest_points(q)
    assert_hits_equal(records(hits), obj, ref, ref_obj, fields=("d", "p", "uv"))
