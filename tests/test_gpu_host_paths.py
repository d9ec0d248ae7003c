"""The host-pointer entry points (GPUScene.find_closest_points, intersect,
sphere intersection, closest silhouette points) copy queries in and records out
through pinned bounce buffers in chunks of 8 MiB, and reuse the caller's record
list: with more queries than one chunk holds (and a ragged last chunk), two
calls in a row on the same list give the bits of a fresh list and of the
device entry point, and the second call leaves nothing of the first."""
import numpy as np
import pytest

import oracle
from tests import meshes
from tests.test_gpu_parity import as_records, assert_bit_equal

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
fcpw = pytest.importorskip("fcpw_amd")

N = 300_001  # 12 MB of records: two chunks, the last one partial


@pytest.fixture(scope="module")
def torus():
    from tests.test_oracle_silhouettes import product_scene
    V, F = meshes.small_torus(64, 32)
    s = product_scene(V, F)
    g = fcpw.GPUScene3D("")
    g.transfer_to_gpu(s)
    return V, F, s, g


@pytest.mark.parametrize("n", [524_288, 524_289, N])
def test_closest_points_through_a_reused_list(gpu, torus, n):
    """524,288 spheres fill exactly one 8 MiB chunk (a plain copy in), 524,289
    add a 16-byte second chunk; the records (40 bytes each) span several chunks"""
    V, F, s, g = torus
    ob = oracle.Bvh(V, F)
    qa = meshes.queries_around(V, n, seed=91)
    qb = meshes.queries_around(V, n, seed=92)
    r2 = np.full(n, np.inf, np.float32)
    r2[::3] = 1e-3
    hits = fcpw.GPUInteractionList()
    g.find_closest_points(qa, r2, hits)
    assert_bit_equal(as_records(hits), ob.closest_points_gpuorder(qa, r2))
    g.find_closest_points(qb, r2, hits)  # same list, other queries
    ref = ob.closest_points_gpuorder(qb, r2)
    assert_bit_equal(as_records(hits), ref)
    assert (ref["index"] < 0).any() and (ref["index"] >= 0).any()
    # the device entry point on the same queries
    sph = np.zeros((n, 4), np.float32)
    sph[:, :3] = qb
    sph[:, 3] = r2
    out = torch.empty((n, 10), dtype=torch.float32, device="cuda")
    g.find_closest_points_device(torch.from_numpy(sph).cuda(), out)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out.cpu().numpy().view(np.uint32), as_records(hits).view(np.uint32).reshape(n, 10))


def test_rays_and_occlusion_through_a_reused_list(gpu, torus):
    V, F, s, g = torus
    ob = oracle.Bvh(V, F)
    hits = fcpw.GPUInteractionList()
    for seed, occ in ((93, False), (94, False), (95, True)):
        o, d = meshes.rays_from(V, N, seed=seed)
        tmax = np.full(N, np.inf, np.float32)
        tmax[::5] = 0.2
        g.intersect(o, d, tmax, hits, check_for_occlusion=occ)
        assert_bit_equal(as_records(hits), ob.intersect_gpuorder(o, d, tmax, occlusion=occ))


def test_sphere_intersection_through_a_reused_list(gpu, torus):
    V, F, s, g = torus
    ob = oracle.Bvh(V, F)
    hits = fcpw.GPUInteractionList()
    for seed in (96, 97):
        q = meshes.queries_around(V, N, seed=seed)
        rng = np.random.default_rng(seed)
        r2 = rng.uniform(0.0, 0.05, N).astype(np.float32)
        rn = rng.uniform(0, 1, (N, 3)).astype(np.float32)
        g.intersect(q, r2, rn, hits)
        assert_bit_equal(as_records(hits), ob.intersect_spheres(q, r2, rn))


def test_silhouettes_through_a_reused_list(gpu, torus):
    """host records equal the device entry point's on the same queries"""
    V, F, s, g = torus
    hits = fcpw.GPUInteractionList()
    for seed in (98, 99):
        q = meshes.queries_around(V, N, seed=seed)
        r2 = np.full(N, np.inf, np.float32)
        flip = (np.arange(N) % 2).astype(np.int32)
        g.find_closest_silhouette_points(q, r2, flip, hits)
        sph = np.zeros((N, 4), np.float32)
        sph[:, :3] = q
        sph[:, 3] = r2
        out = torch.empty((N, 10), dtype=torch.float32, device="cuda")
        g.find_closest_silhouette_points_device(torch.from_numpy(sph).cuda(), torch.from_numpy(flip).cuda(), out)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(out.cpu().numpy().view(np.uint32),
                                      as_records(hits).view(np.uint32).reshape(N, 10))
