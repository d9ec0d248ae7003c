"""Sphere intersection (GPUScene::intersect(spheres, randNums), bvh.slang:257-392)
on the GPU: bit-exact against the oracle's restatement for given random
numbers, in 2D and 3D, through all three entry forms; and, independently of
the restatement, the sampling law itself: the sampled point lies on the chosen
primitive, and 1/pdf is an unbiased estimate of the total size (area or
length) of the primitives within the sphere, checked against an FP64 brute
force."""
import numpy as np
import pytest

import oracle
from tests import meshes

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
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


MESHES = {"icosphere": lambda: meshes.small_icosphere(3), "torus": meshes.small_torus,
          "bunny": lambda: meshes.load_obj("bunny.obj", 3), "star2d": meshes.small_star,
          "walker2d": lambda: meshes.load_obj("walker.obj", 2)}


def sphere_queries(V, n, seed):
    rng = np.random.default_rng(seed)
    c = meshes.queries_around(V, n, seed=seed)
    ext = float(np.abs(V).max())
    r = rng.choice([1e-4, 0.02, 0.1, 0.3, 3.0], n) * ext
    u = rng.uniform(0.0, 1.0, (n, 3)).astype(np.float32)
    u[::11] = 0.0
    u[1::13] = np.nextafter(np.float32(1.0), np.float32(0.0))
    return c, (r * r).astype(np.float32), u


@pytest.mark.parametrize("name", list(MESHES))
def test_sphere_intersection_bit_exact(gpu, name):
    V, F = MESHES[name]()
    scene, g = make_scene(V, F)
    c, r2, u = sphere_queries(V, 20000, seed=31)
    hits = fcpw.GPUInteractionList()
    g.intersect(c, r2, u, hits)
    rec = records(hits).copy()
    ref = oracle.Bvh(V, F).intersect_spheres(c, r2, u)
    np.testing.assert_array_equal(rec["index"].view(np.int32), ref["index"])
    for f in ("d", "p", "n", "uv"):
        a, b = rec[f], ref[f]
        assert ((a == b) | (np.isnan(a) & np.isnan(b))).all(), f
    found = ref["index"] >= 0
    assert 0.2 < found.mean() < 0.99
    # the list form and the device form give the same records
    dim = V.shape[1]
    spheres, rnd = fcpw.GPUBoundingSphereList(), fcpw.GPUFloat3List()
    for k in range(0, len(c), 40):
        x = c[k]
        spheres.append(fcpw.GPUBoundingSphere(fcpw.GPUFloat3(float(x[0]), float(x[1]), float(x[2]) if dim == 3 else 0.0),
                                              float(r2[k])))
        rnd.append(fcpw.GPUFloat3(float(u[k, 0]), float(u[k, 1]), float(u[k, 2]) if dim == 3 else 0.0))
    hits2 = fcpw.GPUInteractionList()
    g.intersect(spheres, rnd, hits2)
    sub = records(hits2)
    np.testing.assert_array_equal(sub.view(np.uint32), rec[::40].copy().view(np.uint32))
    sph = torch.zeros((len(c), 4), dtype=torch.float32, device="cuda")
    sph[:, :dim] = torch.from_numpy(c).cuda()
    sph[:, 3] = torch.from_numpy(r2).cuda()
    ut = torch.from_numpy(u).cuda()
    if dim == 2:
        ut[:, 2] = 0.0
    out = torch.empty((len(c), 10), dtype=torch.float32, device="cuda")
    g.intersect_spheres_device(sph, ut, out)
    torch.cuda.synchronize()
    dev = out.cpu().numpy().view(fcpw.INTERACTION_DTYPE).ravel()
    np.testing.assert_array_equal(dev.view(np.uint32), rec.view(np.uint32))  # (the third number is unused in 2D)


def prim_size64(V, F):
    V64 = V.astype(np.float64)
    if V.shape[1] == 3:
        a, b, c = V64[F[:, 0]], V64[F[:, 1]], V64[F[:, 2]]
        return 0.5 * np.linalg.norm(np.cross(b - a, c - a), axis=1)
    return np.linalg.norm(V64[F[:, 1]] - V64[F[:, 0]], axis=1)


@pytest.mark.parametrize("name", ["icosphere", "star2d"])
def test_inverse_pdf_estimates_the_size_inside_the_sphere(gpu, name):
    V, F = MESHES[name]()
    scene, g = make_scene(V, F)
    dim = V.shape[1]
    sizes = prim_size64(V, F)
    rng = np.random.default_rng(41)
    samples = 200000
    for k in range(6):
        centre = V[rng.integers(len(V))] + rng.normal(0, 0.05, dim).astype(np.float32)
        radius = [0.15, 0.3, 0.6, 1.0, 2.5, 0.08][k]
        d64 = oracle.prim_dist_f64(V, F, np.repeat(centre[None].astype(np.float32), len(F), 0),
                                   np.arange(len(F), dtype=np.int32))
        # primitives within FP32 rounding of the sphere may fall either way
        lo_in, hi_in = sizes[d64 <= radius * (1 - 1e-5)].sum(), sizes[d64 <= radius * (1 + 1e-5)].sum()
        sph = torch.zeros((samples, 4), dtype=torch.float32, device="cuda")
        sph[:, :dim] = torch.from_numpy(centre).cuda()
        sph[:, 3] = radius * radius
        ut = torch.rand((samples, 3), device="cuda")
        out = torch.empty((samples, 10), dtype=torch.float32, device="cuda")
        g.intersect_spheres_device(sph, ut, out)
        rec = out.cpu().numpy().view(fcpw.INTERACTION_DTYPE).ravel()
        hit = rec["index"] != fcpw.UINT_MAX
        est = np.where(hit, 1.0 / rec["d"].astype(np.float64), 0.0)
        se = est.std() / np.sqrt(samples)
        assert lo_in - 5 * se - 1e-6 * hi_in <= est.mean() <= hi_in + 5 * se + 1e-6 * hi_in, \
            (k, est.mean(), lo_in, hi_in, se)
        # every sampled point lies on its primitive
        idx = rec["index"][hit].astype(np.int32)
        dd = oracle.prim_dist_f64(V, F, rec["p"][hit][:, :dim].copy(), idx)
        assert dd.max() <= 1e-5 * float(np.abs(V).max())
