"""The record layouts (option "layout": 0 = DFS order of the internal nodes,
1 = treelets of three levels) give the oracle's results bit for bit: closest
points with normals and finite radii, rays and occlusion rays, sphere
intersection, the traversal counters, refit after a vertex update, the device
build, closest silhouette points before and after a device cone refit, and a
batch whose member was reordered.  Only record indices change between
layouts, so nothing else may."""
import numpy as np
import pytest

import oracle
from tests import meshes
from tests.test_gpu_parity import as_records, assert_bit_equal, make_scene

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
fcpw = pytest.importorskip("fcpw_amd")

MESHES = {
    "bunny": lambda: meshes.load_obj("bunny.obj", 3),
    "torus": lambda: meshes.small_torus(64, 32),
    "soup": lambda: meshes.triangle_soup(500),
    "star2d": lambda: meshes.small_star(4096),
    "walker2d": lambda: meshes.load_obj("walker.obj", 2),
}


@pytest.mark.parametrize("name", list(MESHES))
def test_layouts_give_the_reference_bits(gpu, name):
    V, F = MESHES[name]()
    scene, g = make_scene(V, F)
    ob = oracle.Bvh(V, F)
    q = meshes.queries_around(V, 20000, seed=61)
    r2 = np.full(len(q), np.inf, np.float32)
    r2[::5] = np.float32(0.01 * float(np.abs(V).max()) ** 2)
    o, d = meshes.rays_from(V, 10000, seed=62)
    tmax = np.full(len(o), np.inf, np.float32)
    tmax[::7] = 0.3
    ref_cp = ob.closest_points_gpuorder(q, r2, record_normal=True)
    _, ni, npr, _ = ob.closest_points_gpuorder(q, r2, counters=True)
    ref_ray = ob.intersect_gpuorder(o, d, tmax)
    ref_occ = ob.intersect_gpuorder(o, d, tmax, occlusion=True)
    rn = np.random.default_rng(63).uniform(0, 1, (len(q), 3)).astype(np.float32)
    r2s = np.full(len(q), np.float32(0.05 * float(np.abs(V).max()) ** 2))
    ref_sph = ob.intersect_spheres(q, r2s, rn)
    for layout in (1, 0, 1):
        g.set_option("layout", layout)
        assert g.get_option("layout") == layout
        hits = fcpw.GPUInteractionList()
        g.find_closest_points(q, r2, hits, record_normals=True)
        assert_bit_equal(as_records(hits), ref_cp)
        for occ, ref in ((False, ref_ray), (True, ref_occ)):
            hits = fcpw.GPUInteractionList()
            g.intersect(o, d, tmax, hits, check_for_occlusion=occ)
            assert_bit_equal(as_records(hits), ref)
        hits = fcpw.GPUInteractionList()
        g.intersect(q, r2s, rn, hits)
        assert_bit_equal(as_records(hits), ref_sph)
        g.set_option("stats", 1)
        hits = fcpw.GPUInteractionList()
        g.find_closest_points(q, r2, hits)
        assert (g.info("stat_internal_nodes"), g.info("stat_prim_tests")) == (int(ni.sum()), int(npr.sum()))
        g.set_option("stats", 0)
    # a vertex update and refit on the reordered records
    rng = np.random.default_rng(64)
    V2 = (V + rng.uniform(-0.01, 0.01, V.shape) * float(np.abs(V).max())).astype(np.float32)
    scene.update_object_vertices(V2, 0)
    g.refit(scene)
    ob.update_vertices(V2)
    ob.refit()
    hits = fcpw.GPUInteractionList()
    g.find_closest_points(q, r2, hits)
    assert_bit_equal(as_records(hits), ob.closest_points_gpuorder(q, r2))
    g.set_option("layout", 0)


@pytest.mark.parametrize("name", ["bunny", "star2d"])
def test_layout_of_device_built_trees(gpu, name):
    V, F = MESHES[name]()
    dim = V.shape[1]
    s = fcpw.Scene3D() if dim == 3 else fcpw.Scene2D()
    s.set_object_count(1)
    s.set_object_vertices(V, 0)
    (s.set_object_triangles if dim == 3 else s.set_object_line_segments)(F, 0)
    g = fcpw.GPUScene3D("") if dim == 3 else fcpw.GPUScene2D("")
    q = meshes.queries_around(V, 8000, seed=65)
    ref = oracle.Bvh(V, F).closest_points_gpuorder(q)
    g.build_on_gpu(s)
    g.set_option("layout", 1)  # the device-built records reordered in place
    for rebuild in (False, True):
        if rebuild:
            g.build_on_gpu(s)  # a new build of the context takes its layout
        assert g.get_option("layout") == 1
        hits = fcpw.GPUInteractionList()
        g.find_closest_points(q, np.full(len(q), np.inf, np.float32), hits)
        assert_bit_equal(as_records(hits), ref, fields=("p", "uv", "d"))


@pytest.mark.parametrize("name", ["torus", "walker2d"])
def test_silhouettes_and_cone_refit_in_both_layouts(gpu, name):
    """the cones follow their records: both layouts give the same records,
    before and after a device refit of the cones"""
    from tests.test_oracle_silhouettes import product_scene, queries
    V, F = MESHES[name]()
    dim = V.shape[1]
    s = product_scene(V, F)
    g = fcpw.GPUScene3D("") if dim == 3 else fcpw.GPUScene2D("")
    g.transfer_to_gpu(s)
    q, r2, flip = queries(V, 12000, seed=66)

    def run():
        hits = fcpw.GPUInteractionList()
        g.find_closest_silhouette_points(q, r2, flip.astype(np.int32), hits)
        return as_records(hits).copy().view(np.uint32)
    base = run()
    g.set_option("layout", 1)
    np.testing.assert_array_equal(run(), base)
    V2 = (V + np.random.default_rng(67).uniform(-0.01, 0.01, V.shape) * float(np.abs(V).max())).astype(np.float32)
    s.update_object_vertices(V2, 0)
    g.refit(s)
    moved = run()
    g.set_option("layout", 0)
    np.testing.assert_array_equal(run(), moved)
    assert (moved != base).any()


def test_batch_with_a_reordered_member(gpu):
    from fcpw_amd.batch import BatchedScenes
    items = [MESHES["bunny"](), MESHES["torus"]()]
    scenes = [make_scene(V, F)[1] for V, F in items]
    qs = [meshes.queries_around(V, 7000, seed=68 + k) for k, (V, F) in enumerate(items)]
    sph = torch.zeros((sum(len(x) for x in qs), 4), dtype=torch.float32, device="cuda")
    sph[:, :3] = torch.from_numpy(np.concatenate(qs)).cuda()
    sph[:, 3] = float("inf")
    batch = BatchedScenes(scenes)
    batch.set_queries([len(x) for x in qs])
    out = torch.empty((len(sph), 10), device="cuda")
    batch.find_closest_points_device(sph, out)
    torch.cuda.synchronize()
    base = out.clone()
    scenes[1].set_option("layout", 1)
    out.fill_(0)
    batch.find_closest_points_device(sph, out)
    torch.cuda.synchronize()
    assert torch.equal(out.view(torch.int32), base.view(torch.int32))
