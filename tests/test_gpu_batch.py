"""Batched persistent launch over many scenes: per-scene results bitwise equal
to the single-scene launches (which the parity tests pin to the oracle), for
every batch option, in 2D and 3D, forward, compact forward and adjoint; HIP
graph replays equal to eager calls; re-uploaded members picked up."""
import numpy as np
import pytest

import oracle
from tests import meshes

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
fcpw = pytest.importorskip("fcpw_amd")
from fcpw_amd.batch import BatchedScenes  # noqa: E402


def transfer(V, F):
    dim = V.shape[1]
    s = fcpw.Scene3D() if dim == 3 else fcpw.Scene2D()
    s.set_object_count(1)
    s.set_object_vertices(V, 0)
    (s.set_object_triangles if dim == 3 else s.set_object_line_segments)(F, 0)
    s.build(fcpw.AggregateType.Bvh_SurfaceArea, False)
    g = fcpw.GPUScene3D("") if dim == 3 else fcpw.GPUScene2D("")
    g.transfer_to_gpu(s)
    return s, g


MESHES_3D = [lambda: meshes.small_icosphere(1), lambda: meshes.small_torus(8, 4), lambda: meshes.small_icosphere(4),
             lambda: meshes.small_torus(64, 32), lambda: meshes.triangle_soup(300), lambda: meshes.small_icosphere(2),
             lambda: meshes.small_torus(128, 64), lambda: meshes.small_icosphere(3),
             lambda: meshes.load_obj("bunny.obj", 3), lambda: meshes.small_torus(16, 8)]
COUNTS_3D = [17, 0, 20000, 5000, 64, 777, 12345, 3, 4096, 65]
MESHES_2D = [lambda: meshes.small_star(64, 1), lambda: meshes.small_star(4096, 2),
             lambda: meshes.load_obj("walker.obj", 2), lambda: meshes.small_star(300, 3)]
COUNTS_2D = [5000, 129, 0, 9000]


@pytest.fixture(scope="module")
def world3():
    return build_world(MESHES_3D, COUNTS_3D, seed=100)


@pytest.fixture(scope="module")
def world2():
    return build_world(MESHES_2D, COUNTS_2D, seed=200)


def build_world(makers, counts, seed):
    items = []
    for k, make in enumerate(makers):
        V, F = make()
        s, g = transfer(V, F)
        q = meshes.queries_around(V, counts[k], seed=seed + k) if counts[k] else np.zeros((0, V.shape[1]), np.float32)
        items.append((V, F, s, g, q))
    dim = items[0][0].shape[1]
    q_all = np.concatenate([it[4] for it in items]).astype(np.float32)
    sph = np.zeros((len(q_all), 4), np.float32)
    sph[:, :dim] = q_all
    sph[:, 3] = np.inf
    sph[counts[0]:counts[0] + 5, 3] = 0.01  # a few bounded searches
    return items, q_all, torch.from_numpy(sph).cuda()


def single_reference(items, sph):
    out = torch.empty((sph.shape[0], 10), dtype=torch.float32, device="cuda")
    a = 0
    for (V, F, s, g, q) in items:
        b = a + len(q)
        if b > a:
            g.find_closest_points_device(sph[a:b], out[a:b])
        a = b
    torch.cuda.synchronize()
    return out.cpu().numpy()


BATCH_OPTIONS = [dict(), dict(stack=0), dict(stack=4), dict(stack=8), dict(stack=12), dict(stack=16), dict(hold_min=2),
                 dict(hold_min=64),
                 dict(chunk=256), dict(chunk=512), dict(sort=0), dict(waves_per_cu=1), dict(waves_per_cu=3, chunk=128)]
DEFAULTS = dict(stack=6, hold_min=1, chunk=64, sort=1, waves_per_cu=0)


@pytest.mark.parametrize("world_name", ["world3", "world2"])
def test_batch_equals_single_launches_for_every_option(gpu, world_name, request):
    items, q_all, sph = request.getfixturevalue(world_name)
    ref = single_reference(items, sph)
    batch = BatchedScenes([it[3] for it in items])
    batch.set_queries([len(it[4]) for it in items])
    out = torch.empty((sph.shape[0], 10), dtype=torch.float32, device="cuda")
    for opts in BATCH_OPTIONS:
        for k, v in {**DEFAULTS, **opts}.items():
            batch.set_option(k, v)
        out.fill_(-7.0)
        batch.find_closest_points_device(sph, out)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(out.cpu().numpy().view(np.uint32), ref.view(np.uint32), err_msg=str(opts))
    # the single launches themselves against the oracle's reference-order restatement
    a = 0
    for (V, F, s, g, q) in items:
        b = a + len(q)
        if 0 < len(q) <= 6000:
            sp = sph[a:b].cpu().numpy()
            got = ref[a:b].view(fcpw.INTERACTION_DTYPE).ravel()
            exp = oracle.Bvh(V, F).closest_points_gpuorder(q, sp[:, 3].copy())
            np.testing.assert_array_equal(got["d"], exp["d"])
            np.testing.assert_array_equal(got["index"].view(np.int32), exp["index"])
        a = b


def test_batch_compact_forward_and_adjoint_equal_single(gpu, world3):
    items, q_all, sph = world3
    n = len(q_all)
    x = torch.from_numpy(q_all).cuda()
    batch = BatchedScenes([it[3] for it in items])
    batch.set_queries([len(it[4]) for it in items])
    d = torch.empty(n, device="cuda")
    p = torch.empty((n, 3), device="cuda")
    prim = torch.empty(n, dtype=torch.int32, device="cuda")
    slot = torch.empty(n, dtype=torch.int32, device="cuda")
    batch.closest_points_compact_device(x, d, p, prim, slot)
    w = torch.from_numpy(np.random.default_rng(5).uniform(0.5, 1.5, n).astype(np.float32)).cuda()
    gx = torch.empty_like(x)
    nv = int(batch.vertex_offsets[-1])
    gV = torch.zeros((nv, 3), device="cuda")
    batch.distance_backward_device(x, slot, w, gx, gV)
    torch.cuda.synchronize()
    st = torch.cuda.current_stream().cuda_stream
    a = 0
    for k, (V, F, s, g, q) in enumerate(items):
        b = a + len(q)
        if b == a:
            continue
        m = b - a
        d1 = torch.empty(m, device="cuda")
        p1 = torch.empty((m, 3), device="cuda")
        pr1 = torch.empty(m, dtype=torch.int32, device="cuda")
        sl1 = torch.empty(m, dtype=torch.int32, device="cuda")
        xs = x[a:b].contiguous()
        g._closest_points_compact_device(xs.data_ptr(), 0, m, d1.data_ptr(), p1.data_ptr(), pr1.data_ptr(),
                                         sl1.data_ptr(), st)
        gx1 = torch.empty_like(xs)
        gV1 = torch.zeros((len(V), 3), device="cuda")
        ws = w[a:b].contiguous()
        g._distance_backward_device(xs.data_ptr(), sl1.data_ptr(), ws.data_ptr(), m, gx1.data_ptr(), gV1.data_ptr(), st)
        torch.cuda.synchronize()
        assert torch.equal(d[a:b], d1) and torch.equal(p[a:b], p1) and torch.equal(prim[a:b], pr1)
        assert torch.equal(slot[a:b], sl1) and torch.equal(gx[a:b], gx1)
        v0 = int(batch.vertex_offsets[k])
        torch.testing.assert_close(gV[v0:v0 + len(V)], gV1, rtol=1e-5, atol=1e-5)
        a = b


def test_batch_adjoint_in_morton_and_caller_order(gpu, world3):
    """the adjoint walks each scene's queries in the latest forward's Morton
    order, or in the caller's order when set_queries has invalidated that order
    since: the same dL/dx bit for bit, dL/dV up to the FP32 summation order"""
    items, q_all, sph = world3
    n = len(q_all)
    x = torch.from_numpy(q_all).cuda()
    counts = [len(it[4]) for it in items]
    batch = BatchedScenes([it[3] for it in items])
    batch.set_queries(counts)
    d = torch.empty(n, device="cuda")
    slot = torch.empty(n, dtype=torch.int32, device="cuda")
    prim = torch.empty(n, dtype=torch.int32, device="cuda")
    batch.closest_points_compact_device(x, d, prim=prim, slot=slot)
    wn = np.random.default_rng(6).uniform(0.5, 1.5, n).astype(np.float32)
    w = torch.from_numpy(wn).cuda()
    nv = int(batch.vertex_offsets[-1])
    # per vertex, the sum of |terms| landing on it (each term is at most |w| in size):
    # FP32 sums of one vertex in two orders differ by a few ulp of it
    pr = prim.cpu().numpy()
    abs_sum = np.zeros(nv)
    a = 0
    for k, it in enumerate(items):
        b = a + len(it[4])
        ok = pr[a:b] >= 0
        ids = it[1][pr[a:b][ok]].ravel() + int(batch.vertex_offsets[k])
        abs_sum += np.bincount(ids, np.repeat(wn[a:b][ok], 3), nv)
        a = b
    abs_sum = torch.from_numpy(abs_sum.astype(np.float32)).cuda()[:, None]
    runs = []
    for reset in (False, True):
        if reset:
            batch.set_queries(counts)  # the ranges are rebuilt: the order of the forward no longer applies
        gx = torch.empty_like(x)
        gV = torch.zeros((nv, 3), device="cuda")
        batch.distance_backward_device(x, slot, w, gx, gV)
        torch.cuda.synchronize()
        runs.append((gx, gV))
    assert torch.equal(runs[0][0], runs[1][0])
    assert ((runs[0][1] - runs[1][1]).abs() <= 1e-5 * abs_sum + 1e-12).all()
    assert (runs[0][1].abs() > 1e-5 * abs_sum).float().mean() > 0.5


def test_batch_graph_replays_equal_eager(gpu, world3):
    items, q_all, sph = world3
    batch = BatchedScenes([it[3] for it in items])
    batch.set_queries([len(it[4]) for it in items])
    out = torch.empty((sph.shape[0], 10), dtype=torch.float32, device="cuda")
    batch.find_closest_points_device(sph, out)
    torch.cuda.synchronize()
    eager = out.clone()
    graph = batch.capture(lambda: batch.find_closest_points_device(sph, out))
    for rep in range(4):
        out.fill_(float(rep) - 9.0)  # eager work between replays
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(out.view(torch.int32), eager.view(torch.int32)), rep
    # the same for the compact forward with the adjoint after it
    n = len(q_all)
    x = torch.from_numpy(q_all).cuda()
    d = torch.empty(n, device="cuda")
    slot = torch.empty(n, dtype=torch.int32, device="cuda")
    gd = torch.ones(n, device="cuda")
    gx = torch.empty_like(x)

    def step():
        batch.closest_points_compact_device(x, d, slot=slot)
        batch.distance_backward_device(x, slot, gd, gx)
    step()
    torch.cuda.synchronize()
    d_eager, gx_eager = d.clone(), gx.clone()
    graph = batch.capture(step)
    for rep in range(3):
        d.zero_()
        gx.zero_()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(d, d_eager) and torch.equal(gx, gx_eager), rep


def test_batch_picks_up_a_reuploaded_member(gpu):
    V1, F1 = meshes.small_icosphere(2)
    V2, F2 = meshes.small_torus(32, 16)
    s1, g1 = transfer(V1, F1)
    s2, g2 = transfer(V2, F2)
    q1 = meshes.queries_around(V1, 3000, seed=1)
    q2 = meshes.queries_around(V2, 5000, seed=2)
    sph = np.zeros((8000, 4), np.float32)
    sph[:3000, :3] = q1
    sph[3000:, :3] = q2
    sph[:, 3] = np.inf
    sph_t = torch.from_numpy(sph).cuda()
    batch = BatchedScenes([g1, g2])
    batch.set_queries([3000, 5000])
    out = torch.empty((8000, 10), device="cuda")
    batch.find_closest_points_device(sph_t, out)
    # move the torus and transfer it again (new device buffers)
    V2b = (V2 * np.float32(1.3)).astype(np.float32)
    s2.update_object_vertices(V2b, 0)
    s2.build(fcpw.AggregateType.Bvh_SurfaceArea, False)
    g2.transfer_to_gpu(s2)
    batch.find_closest_points_device(sph_t, out)
    torch.cuda.synchronize()
    got = out.cpu().numpy().view(fcpw.INTERACTION_DTYPE).ravel()
    exp = oracle.Bvh(V2b, F2).closest_points_gpuorder(q2)
    np.testing.assert_array_equal(got["d"][3000:], exp["d"])


def test_batch_rejects_a_member_retransferred_as_a_scene_of_several_objects(gpu):
    """a member transferred again from a scene of two objects holds an instance
    tree, not a triangle BVH: the batch refuses it on the host, before any launch"""
    from tests.test_host_instances import make_scene
    V1, F1 = meshes.small_icosphere(2)
    V2, F2 = meshes.small_torus(32, 16)
    s1, g1 = transfer(V1, F1)
    s2, g2 = transfer(V2, F2)
    batch = BatchedScenes([g1, g2])
    batch.set_queries([100, 100])
    sph = torch.zeros((200, 4), device="cuda")
    sph[:, 3] = float("inf")
    out = torch.empty((200, 10), device="cuda")
    batch.find_closest_points_device(sph, out)
    graph = batch.capture(lambda: batch.find_closest_points_device(sph, out))
    graph.replay()
    g2.transfer_to_gpu(make_scene([(V1, F1), (V2, F2)], [[], []]))
    assert g2.info("n_objects") == 2
    with pytest.raises(RuntimeError, match="single-object"):
        batch.find_closest_points_device(sph, out)
    with pytest.raises(RuntimeError, match="capture the graph again"):
        graph.replay()  # the captured table points at the freed single-object buffers
    x = torch.zeros((200, 3), device="cuda")
    with pytest.raises(RuntimeError, match="single-object"):
        batch.closest_points_compact_device(x, torch.empty(200, device="cuda"))
    torch.cuda.synchronize()


def test_batch_graph_refuses_replay_after_set_queries(gpu):
    V1, F1 = meshes.small_icosphere(2)
    _, g1 = transfer(V1, F1)
    batch = BatchedScenes([g1])
    batch.set_queries([64])
    sph = torch.zeros((64, 4), device="cuda")
    out = torch.empty((64, 10), device="cuda")
    graph = batch.capture(lambda: batch.find_closest_points_device(sph, out))
    graph.replay()
    batch.set_queries([5000])
    with pytest.raises(RuntimeError, match="capture the graph again"):
        graph.replay()
    torch.cuda.synchronize()


def test_batch_tensor_arguments_are_checked_before_launch(gpu):
    V, F = meshes.small_icosphere(1)
    _, g = transfer(V, F)
    batch = BatchedScenes([g])
    batch.set_queries([32])
    x = torch.zeros((32, 3), device="cuda")
    f = torch.zeros(32, device="cuda")
    i = torch.zeros(32, dtype=torch.int32, device="cuda")
    nv = g.info("n_vertices")
    bad = [
        lambda: batch.closest_points_compact_device(x, d=torch.zeros(31, device="cuda")),
        lambda: batch.closest_points_compact_device(x, p=torch.zeros((32, 2), device="cuda")),
        lambda: batch.closest_points_compact_device(x, prim=f),
        lambda: batch.closest_points_compact_device(x, slot=torch.zeros(32, dtype=torch.int64, device="cuda")),
        lambda: batch.closest_points_compact_device(x, r2=torch.zeros((32, 1), device="cuda")),
        lambda: batch.closest_points_compact_device(x[:, :2].contiguous()),
        lambda: batch.distance_backward_device(x, f, f),
        lambda: batch.distance_backward_device(x, i, f, grad_points=torch.zeros((32, 3), device="cuda")[::1, :2]),
        lambda: batch.distance_backward_device(x, i, f, grad_vertices=torch.zeros((nv - 1, 3), device="cuda")),
        lambda: batch.distance_backward_device(x.cpu(), i, f),
    ]
    for k, call in enumerate(bad):
        with pytest.raises((TypeError, ValueError)):
            call()
    batch.distance_backward_device(x, i, f, grad_points=torch.zeros((32, 3), device="cuda"),
                                   grad_vertices=torch.zeros((nv, 3), device="cuda"))
    torch.cuda.synchronize()


def test_batch_errors(gpu):
    V, F = meshes.small_icosphere(1)
    _, g3 = transfer(V, F)
    V2, F2 = meshes.small_star(64)
    _, g2 = transfer(V2, F2)
    with pytest.raises(ValueError):
        BatchedScenes([g3, g2])
    batch = BatchedScenes([g3])
    with pytest.raises(ValueError):
        batch.set_queries([1, 2])
    with pytest.raises(ValueError):
        batch.set_option("chunk", 100)
    batch.set_queries([10])
    with pytest.raises(ValueError):
        batch.find_closest_points_device(torch.zeros((9, 4), device="cuda"), torch.zeros((9, 10), device="cuda"))


def test_autotuner_keeps_results_and_caches(gpu, world3):
    from fcpw_amd import autotune
    items, q_all, sph = world3
    V, F, s, g, q = items[2]
    a = sum(len(it[4]) for it in items[:2])
    sp = sph[a:a + len(q)].contiguous()
    out = torch.empty((len(q), 10), device="cuda")
    g.find_closest_points_device(sp, out)
    torch.cuda.synchronize()
    ref = out.clone()
    autotune.clear_cache()
    best, times = autotune.tune_scene(g, sp, reps=1)
    assert times and set(best) <= {"layout", "coop", "stack", "hold_min", "block", "xcds"}
    assert {dict(k)["layout"] for k in times} == {0, 1}
    g.find_closest_points_device(sp, out)
    torch.cuda.synchronize()
    assert torch.equal(out.view(torch.int32), ref.view(torch.int32))  # misses carry UINT_MAX (NaN as float)
    best2, times2 = autotune.tune_scene(g, sp, reps=1)
    assert best2 == best and times2 == {}
    batch = BatchedScenes([it[3] for it in items])
    batch.set_queries([len(it[4]) for it in items])
    bout = torch.empty((sph.shape[0], 10), device="cuda")
    bbest, btimes = autotune.tune_batch(batch, sph, bout, reps=1)
    assert btimes and batch.options()["stack"] == bbest["stack"]
