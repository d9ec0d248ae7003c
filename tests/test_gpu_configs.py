"""The workloads at their full configured sizes, against the oracle:

  * rays on the 1,048,576-triangle torus (16M rays in the benchmark): 65,536
    of them, bit for bit against the oracle's reference GPU order
    (fcpw_gpu.inl:137-156, bvh.slang:144-254), closest hit and occlusion;
  * the differentiable distance on the 200,000-triangle geodesic sphere with
    4,194,304 queries and L = mean d^2: dL/dx and dL/dV against the FP64
    adjoint over all queries, FP64 central differences (h = 1e-4) on 1,024
    points and 1,024 vertices (the vertex differences localized to the
    vertex's one-ring), and the 100-step Adam loop with device refit checked
    against the FP64 brute force along the way;
  * the batch of 256 uneven meshes (20.6M triangles, 35M queries) in one
    persistent launch, eager and replayed from a HIP graph: bitwise equal to
    the 256 separate launches, and 1,024 queries of every mesh bit for bit
    against the oracle."""
import numpy as np
import pytest

import oracle
from fcpw_amd import synth
from tests.test_gpu_autograd import tri_closest64

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
fcpw = pytest.importorskip("fcpw_amd")
from fcpw_amd.batch import BatchedScenes  # noqa: E402
from fcpw_amd.torch_ops import closest_point_distance  # noqa: E402

H = 1e-4
EPS32 = float(np.finfo(np.float32).eps)


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


def rays_tensor(o, d):
    r = torch.zeros((len(o), 10), dtype=torch.float32, device="cuda")
    r[:, 0:3] = torch.from_numpy(o).cuda()
    r[:, 3:6] = torch.from_numpy(d).cuda()
    r[:, 6:9] = 1.0 / r[:, 3:6]
    r[:, 9] = float("inf")
    return r


def test_torus_rays_at_full_size_equal_oracle(gpu):
    """the benchmark's 16M rays, every 256th checked (65,536), through both ray kernels"""
    w = synth.torus_1m_workload(n_queries=16, n_rays=1 << 24)
    V, F = w["positions"], w["indices"]
    s, g = transfer(V, F)
    o, d = w["ray_o"], w["ray_d"]
    rays = rays_tensor(o, d)
    out = torch.empty((len(o), 10), dtype=torch.float32, device="cuda")
    b = oracle.Bvh(V, F)
    sub = np.arange(0, len(o), 256)
    for occlusion in (False, True):
        g.intersect_device(rays, out, check_for_occlusion=occlusion)
        torch.cuda.synchronize()
        got = out[torch.from_numpy(sub).cuda()].cpu().numpy().view(fcpw.INTERACTION_DTYPE).ravel()
        ref = b.intersect_gpuorder(o[sub], d[sub], np.full(len(sub), np.inf, np.float32), occlusion=occlusion)
        np.testing.assert_array_equal(got["index"].view(np.int32), ref["index"])
        hit = ref["index"] >= 0
        assert 0.2 < hit.mean() < 0.9
        for f in ("d", "p", "n", "uv"):
            np.testing.assert_array_equal(got[f], ref[f], err_msg=f)
        # the one-ray-per-lane kernel gives the same records on all 16M rays
        first = out.clone()
        g.set_option("ray_coop", 0)
        g.intersect_device(rays, out, check_for_occlusion=occlusion)
        g.set_option("ray_coop", 1)
        torch.cuda.synchronize()
        assert torch.equal(first.view(torch.int32), out.view(torch.int32))


@pytest.fixture(scope="module")
def geodesic():
    w = synth.geodesic_workload()
    V, F, q = w["positions"], w["indices"], w["queries"]
    s, g = transfer(V, F)
    return V, F, q, s, g


def one_rings(F, nv):
    """incident triangles of every vertex (CSR)"""
    tri = np.repeat(np.arange(len(F)), 3)
    ver = F.ravel()
    o = np.argsort(ver, kind="stable")
    start = np.searchsorted(ver[o], np.arange(nv + 1))
    return tri[o], start


def test_adjoint_at_config_size_against_fp64(gpu, geodesic):
    """L = mean d^2 over 4,194,304 queries on the 200k-triangle sphere:
      dL/dx against 2/N (x - p64) on the kernel's primitive (all queries), and
      against FP64 central differences of d on 1,024 points; dL/dV against the
      FP64 scatter over all queries, and against FP64 central differences of L
      on 1,024 vertices, localized to the queries that can involve the vertex's
      one-ring (a vertex is skipped only if a query is within 2h of a tie
      between its ring and a triangle outside it)"""
    V, F, q, s, g = geodesic
    N = len(q)
    x = torch.from_numpy(q).cuda().requires_grad_(True)
    Vt = torch.from_numpy(V).cuda().requires_grad_(True)
    d, p, prim = closest_point_distance(g, x, Vt)
    loss = (d * d).mean()
    loss.backward()
    gx = x.grad.cpu().numpy().astype(np.float64)
    gV = Vt.grad.cpu().numpy().astype(np.float64)
    pr = prim.cpu().numpy()
    dg = d.detach().cpu().numpy().astype(np.float64)
    assert (pr >= 0).all()
    V64, q64 = V.astype(np.float64), q.astype(np.float64)
    Fp = F[pr]
    p64, bw, d64c = tri_closest64(q64, V64[Fp[:, 0]], V64[Fp[:, 1]], V64[Fp[:, 2]])

    # dL/dx = 2/N (x - p): the FP32 p carries a few ulp of |x|
    floor = 8 * EPS32 * np.abs(q64).max()
    assert np.abs(gx - 2.0 / N * (q64 - p64)).max() <= 2.0 / N * floor
    # dL/dV = -2/N sum b_k (x - p) over the queries on each vertex's triangles
    ref = np.zeros_like(V64)
    mag = np.zeros(len(V64))
    for k in range(3):
        for c in range(3):
            ref[:, c] += np.bincount(Fp[:, k], -(2.0 / N) * bw[:, k] * (q64[:, c] - p64[:, c]), len(V64))
        mag += np.bincount(Fp[:, k], (2.0 / N) * bw[:, k] * d64c, len(V64))
    err = np.abs(gV - ref).max(axis=1)
    assert (err <= 1e-4 * mag + 1e-12).all(), np.max(err - 1e-4 * mag)

    # the kernel's primitive is FP64-optimal (to the FP32 distance error) on a sample
    rng = np.random.default_rng(41)
    pts = rng.choice(N, 3072, replace=False)
    dbest, _, ibest = oracle.brute_closest_f64(V, F, q[pts])
    assert (d64c[pts] - dbest <= 1e-6 * dbest + floor).all()
    assert np.abs(dg[pts] - dbest).max() <= 1e-6 * dbest.max() + floor
    # the finite differences below use the first 1,024 of them with d > 0.01
    keep = np.nonzero(dbest > 0.01)[0][:1024]
    pts, dbest, ibest = pts[keep], dbest[keep], ibest[keep]

    # FP64 central differences of d at 1,024 points.  Within one Voronoi region
    # of one primitive the error is O((h/d)^2).  Across a region boundary d
    # stays C1 and the difference is the mean of the gradient (x - p)/d along
    # the stencil, within (4h + |p+ - p-|)/d of its value at x.  Where the
    # closest point jumps (|p+ - p-| > 20h) the stencil straddles a medial
    # ridge -- a tie between distant parts of the noisy surface, where d is not
    # differentiable -- and the point is excluded (a tie is not a gradient
    # error), counted.
    far = pts[dbest > 0.01]
    qq = np.concatenate([q64[far] + s_ * H * e for e in np.eye(3) for s_ in (1.0, -1.0)])
    d_pm, p_pm, i_pm = oracle.brute_closest_f64(V, F, qq.astype(np.float32))
    # the perturbed points are rounded to FP32: use their exact FP32 offsets
    qq32 = qq.astype(np.float32).astype(np.float64)
    n_far = len(far)
    i0_p = ibest[dbest > 0.01]
    dd = dbest[dbest > 0.01]
    grad_d = gx[far] * N / (2.0 * np.maximum(dg[far], 1e-30))[:, None]  # dL/dx = 2/N d grad d

    def region(pts, tris):
        """Voronoi region of the closest point: which barycentric weights are nonzero"""
        w = tri_closest64(pts, V64[F[tris, 0]], V64[F[tris, 1]], V64[F[tris, 2]])[1]
        return (w > 1e-12) @ np.array([1, 2, 4])
    r0 = region(q64[far], i0_p)
    n_region = n_ridge = 0
    for k in range(3):
        a, b_ = slice((2 * k) * n_far, (2 * k + 1) * n_far), slice((2 * k + 1) * n_far, (2 * k + 2) * n_far)
        step = qq32[a, k] - qq32[b_, k]
        fd = (d_pm[a] - d_pm[b_]) / step
        same = ((i_pm[a] == i0_p) & (i_pm[b_] == i0_p) & (region(qq32[a], i0_p) == r0) &
                (region(qq32[b_], i0_p) == r0))
        jump = np.linalg.norm(p_pm[a] - p_pm[b_], axis=1)
        ridge = jump > 20 * H
        tol = 1e-5 + floor / dd + 3.0 * (H / dd) ** 2
        tol = np.where(same, tol, tol + (4 * H + jump) / dd)
        ok = ridge | (np.abs(grad_d[:, k] - fd) <= tol)
        n_region += int((~same & ~ridge).sum())
        n_ridge += int(ridge.sum())
        assert ok.all(), (k, np.nonzero(~ok)[0][:5])
    assert n_far == 1024 and n_ridge < 0.05 * 3 * n_far and n_region < 0.2 * 3 * n_far, (n_far, n_ridge, n_region)

    # FP64 central differences of L on 1,024 vertices, localized to the one-ring:
    # L changes only through the queries whose distance involves a triangle of
    # the ring.  The users (queries whose kernel primitive is in the ring) are
    # recomputed against the ring.  A user within 2h of a triangle just outside
    # the ring, or a query of such a triangle within 2h of the ring, sits at a
    # tie whose one-sided derivative lies between 0 and the ring's own: the
    # tolerance takes that query's |2/N d b_v| as slack (zero when the tie is a
    # shared edge or vertex, where b_v = 0), and the slack is counted.
    tri_of, start = one_rings(F, len(V))
    verts = rng.choice(len(V), 1024, replace=False)
    order = np.argsort(pr, kind="stable")
    qstart = np.searchsorted(pr[order], np.arange(len(F) + 1))  # queries of each primitive (CSR)

    def queries_on(tris):
        return np.concatenate([order[qstart[t]:qstart[t + 1]] for t in tris])

    def ring_dist(qs, ring, Vp, v):
        """min distance of each query to the ring and the weight of v at that point"""
        qr = np.repeat(qs, len(ring), 0)
        tr = np.tile(ring, len(qs))
        _, wr, dr = tri_closest64(qr, Vp[F[tr, 0]], Vp[F[tr, 1]], Vp[F[tr, 2]])
        dr = dr.reshape(len(qs), -1)
        wv = (wr * (F[tr] == v)).sum(1).reshape(len(qs), -1)
        k = dr.argmin(1)
        return dr[np.arange(len(qs)), k], wv[np.arange(len(qs)), k]
    checked = with_slack = 0
    rel = []
    for v in verts:
        ring = tri_of[start[v]:start[v + 1]]
        users = queries_on(ring)
        if len(users) == 0:
            continue
        qu = q64[users]
        d_ring, wv = ring_dist(qu, ring, V64, v)
        nb = np.unique(np.concatenate([tri_of[start[u]:start[u + 1]] for u in np.unique(F[ring])]))
        outside = np.setdiff1d(nb, ring)
        slack = 0.0
        if len(outside):
            qo = np.repeat(qu, len(outside), 0)
            to = np.tile(outside, len(users))
            do = tri_closest64(qo, V64[F[to, 0]], V64[F[to, 1]], V64[F[to, 2]])[2].reshape(len(users), -1).min(1)
            tie = do <= d_ring + 2 * H
            slack += 2.0 / N * float((d_ring[tie] * wv[tie]).sum())
            others = queries_on(outside)
            if len(others):
                dn, wn = ring_dist(q64[others], ring, V64, v)
                tie_in = dn <= d64c[others] + 2 * H
                slack += 2.0 / N * float((dn[tie_in] * wn[tie_in]).sum())

        def L_ring(Vp):
            dr, _ = ring_dist(qu, ring, Vp, v)
            return (dr * dr).sum() / N
        fd = np.zeros(3)
        for k in range(3):
            Vp, Vm = V64.copy(), V64.copy()
            Vp[v, k] += H
            Vm[v, k] -= H
            fd[k] = (L_ring(Vp) - L_ring(Vm)) / (2 * H)
        err = np.abs(gV[v] - fd).max()
        assert err <= 1e-3 * mag[v] + slack + 1e-12, (v, gV[v], fd, mag[v], slack)
        with_slack += int(err > 1e-3 * mag[v] + 1e-12)
        rel.append(err / mag[v])
        checked += 1
    # most vertices hold the tight bound; a gradient with a dropped or
    # mis-signed term would miss it on nearly all of them
    assert checked >= 1000 and with_slack < 0.25 * checked and np.median(rel) < 1e-4, (
        checked, with_slack, np.median(rel))


def test_hundred_step_adam_loop_with_device_refit(gpu, geodesic):
    """config 4's loop: 100 Adam steps on the vertex positions of the
    200k-triangle sphere with the BVH refit on the device before every
    forward; at steps 0, 50 and 99 the forward distances of 2,048 queries
    equal the FP64 brute force on the current mesh (which holds only if the
    refit kept the tree valid), and the loss decreases"""
    V, F, q, s, g = geodesic
    x = torch.from_numpy(q).cuda()
    Vt = torch.nn.Parameter(torch.from_numpy(V).cuda())
    opt = torch.optim.Adam([Vt], lr=1e-4)
    sub = np.arange(0, len(q), len(q) // 2048)[:2048]
    subt = torch.from_numpy(sub).cuda()
    losses = []
    for step in range(100):
        opt.zero_grad()
        d, p, prim = closest_point_distance(g, x, Vt)
        loss = (d * d).mean()
        loss.backward()
        if step in (0, 50, 99):
            Vc = Vt.detach().cpu().numpy()
            d64, _, _ = oracle.brute_closest_f64(Vc, F, q[sub])
            dn = d.detach()[subt].cpu().numpy()
            assert np.abs(dn - d64).max() <= 1e-6 * d64.max() + 8 * EPS32 * np.abs(q).max(), step
        opt.step()
        losses.append(loss.item())
    assert all(b <= a for a, b in zip(losses[::10], losses[10::10])), losses[::10]
    assert losses[-1] < 0.9 * losses[0]
    # the scene holds the last V: restore the original geometry for later tests
    V0 = torch.from_numpy(V).cuda()
    g._update_positions_device(V0.data_ptr(), len(V), torch.cuda.current_stream().cuda_stream)
    g._refit_device(torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()


def test_batch_of_256_meshes_at_config_size(gpu):
    meshes = synth.batch_workload()
    assert len(meshes) == 256
    scenes = [transfer(m["positions"], m["indices"])[1] for m in meshes]
    counts = [len(m["queries"]) for m in meshes]
    q_all = np.concatenate([m["queries"] for m in meshes])
    n = len(q_all)
    assert n > 3e7 and sum(len(m["indices"]) for m in meshes) > 2e7
    sph = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
    sph[:, :3] = torch.from_numpy(q_all).cuda()
    sph[:, 3] = float("inf")
    out = torch.empty((n, 10), dtype=torch.float32, device="cuda")
    batch = BatchedScenes(scenes)
    batch.set_queries(counts)
    batch.find_closest_points_device(sph, out)
    torch.cuda.synchronize()
    eager = out.clone()
    graph = batch.capture(lambda: batch.find_closest_points_device(sph, out))
    for rep in range(2):
        out.fill_(-1.0)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(out.view(torch.int32), eager.view(torch.int32)), rep
    out.fill_(-1.0)
    batch.separate_find_closest_points_device(sph, out)
    torch.cuda.synchronize()
    assert torch.equal(out.view(torch.int32), eager.view(torch.int32))
    # 1,024 queries of every mesh (all of a mesh with fewer) against the oracle
    rec = eager.cpu().numpy().view(fcpw.INTERACTION_DTYPE).ravel()
    a = 0
    for m, c in zip(meshes, counts):
        pick = np.linspace(0, c - 1, min(c, 1024)).astype(np.int64)
        ref = oracle.Bvh(m["positions"], m["indices"]).closest_points_gpuorder(m["queries"][pick])
        got = rec[a + pick]
        np.testing.assert_array_equal(got["index"].view(np.int32), ref["index"])
        for f in ("d", "p", "uv"):
            np.testing.assert_array_equal(got[f], ref[f], err_msg=f)
        a += c
    # the compact forward + adjoint of the batch equal the separate scenes' (a graph replay)
    x = sph[:, :3].contiguous()
    d = torch.empty(n, device="cuda")
    slot = torch.empty(n, dtype=torch.int32, device="cuda")
    gd = torch.ones(n, device="cuda")
    gx = torch.empty_like(x)

    def step():
        batch.closest_points_compact_device(x, d, slot=slot)
        batch.distance_backward_device(x, slot, gd, gx)
    gr = batch.capture(step)
    gr.replay()
    torch.cuda.synchronize()
    assert torch.equal(d, eager[:, 8])
    diff = (x - torch.from_numpy(rec["p"].copy()).cuda()) / d[:, None]
    ok = d > 1e-6
    torch.testing.assert_close(gx[ok], diff[ok], rtol=1e-5, atol=1e-6)
