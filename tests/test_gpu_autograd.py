"""Differentiable distance on the GPU: the adjoint kernel behind
torch.autograd.Function against FP64 central finite differences (h = 1e-4)
and the analytic (x - p)/d, plus an optimisation loop that refits the BVH
on the device between steps."""
import numpy as np
import pytest

import oracle
from fcpw_amd import synth
from tests import meshes

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
fcpw = pytest.importorskip("fcpw_amd")
from fcpw_amd.torch_ops import closest_point_distance  # noqa: E402

H = 1e-4


def make(V, F):
    dim = V.shape[1]
    s = fcpw.Scene3D() if dim == 3 else fcpw.Scene2D()
    s.set_object_count(1)
    s.set_object_vertices(V, 0)
    (s.set_object_triangles if dim == 3 else s.set_object_line_segments)(F, 0)
    s.build(fcpw.AggregateType.Bvh_SurfaceArea, False)
    g = fcpw.GPUScene3D("") if dim == 3 else fcpw.GPUScene2D("")
    g.transfer_to_gpu(s)
    return s, g


def near_queries(V, F, n, radius, seed):
    if V.shape[1] == 3:
        return synth.f32(synth.near_surface(n, V.astype(np.float64), F, radius, synth.Rng(seed)))
    rng = np.random.default_rng(seed)
    k = rng.integers(0, len(F), n)
    t = rng.uniform(0, 1, n)
    p = V[F[k, 0]] * (1 - t)[:, None] + V[F[k, 1]] * t[:, None]
    return (p + rng.normal(0, radius, (n, 2))).astype(np.float32)


@pytest.mark.parametrize("name", ["icosphere", "star2d"])
def test_forward_compact_matches_records(gpu, name):
    V, F = meshes.small_icosphere(4) if name == "icosphere" else meshes.small_star(4096)
    s, g = make(V, F)
    q = near_queries(V, F, 50000, 0.1, 3)
    x = torch.from_numpy(q).cuda()
    d, p, prim = closest_point_distance(g, x)
    hits = fcpw.GPUInteractionList()
    g.find_closest_points(q, np.full(len(q), np.inf, np.float32), hits)
    np.testing.assert_array_equal(d.cpu().numpy(), hits.get_distances())
    np.testing.assert_array_equal(prim.cpu().numpy().astype(np.uint32), hits.get_indices())
    np.testing.assert_array_equal(p.cpu().numpy(), hits.get_positions()[:, :V.shape[1]])
    ref = oracle.Bvh(V, F).closest_points_gpuorder(q)
    np.testing.assert_array_equal(d.cpu().numpy(), ref["d"])


@pytest.mark.parametrize("name", ["icosphere", "star2d"])
def test_grad_x_is_unit_direction(gpu, name):
    V, F = meshes.small_icosphere(4) if name == "icosphere" else meshes.small_star(4096)
    s, g = make(V, F)
    q = near_queries(V, F, 40000, 0.1, 4)
    x = torch.from_numpy(q).cuda().requires_grad_(True)
    d, p, prim = closest_point_distance(g, x)
    w = torch.rand(len(q), device="cuda")
    (w * d).sum().backward()
    xd = x.detach()
    expect = w[:, None] * (xd - p) / d[:, None]
    ok = d > 1e-6
    torch.testing.assert_close(x.grad[ok], expect[ok], rtol=1e-5, atol=1e-7)


def tri_closest64(x, a, b, c):
    """FP64 closest point on each triangle (a, b, c) to x, with its barycentric
    weights: the plane projection when it falls inside, else the nearest
    clamped edge point (a different method from the kernels' Voronoi-region
    test)."""
    n = np.cross(b - a, c - a)
    nn = np.einsum("ij,ij->i", n, n)
    s = np.einsum("ij,ij->i", x - a, n) / nn
    q = x - s[:, None] * n
    la = np.einsum("ij,ij->i", np.cross(c - b, q - b), n) / nn
    lb = np.einsum("ij,ij->i", np.cross(a - c, q - c), n) / nn
    lc = 1.0 - la - lb
    best_p, best_w, best_d = q.copy(), np.stack([la, lb, lc], 1), np.abs(s) * np.sqrt(nn)
    inside = (la >= 0) & (lb >= 0) & (lc >= 0)
    best_d = np.where(inside, best_d, np.inf)
    for (u, v, iu, iv) in ((a, b, 0, 1), (b, c, 1, 2), (c, a, 2, 0)):
        e = v - u
        t = np.clip(np.einsum("ij,ij->i", x - u, e) / np.einsum("ij,ij->i", e, e), 0.0, 1.0)
        pe = u + t[:, None] * e
        de = np.linalg.norm(x - pe, axis=1)
        take = de < best_d
        w = np.zeros((len(x), 3))
        w[:, iu] = 1.0 - t
        w[:, iv] = t
        best_p[take], best_w[take], best_d[take] = pe[take], w[take], de[take]
    return best_p, best_w, best_d


EPS32 = float(np.finfo(np.float32).eps)


def test_grad_x_matches_fp64_adjoint_and_fd(gpu):
    """x.grad against (x - p)/d evaluated in FP64 on the primitive the kernel
    chose (which must be optimal in FP64 up to the FP32 distance error), and
    against FP64 central differences where those are accurate (d >> H)"""
    V, F = meshes.small_icosphere(3)
    s, g = make(V, F)
    q = near_queries(V, F, 4096, 0.1, 5)
    x = torch.from_numpy(q).cuda().requires_grad_(True)
    d, p, prim = closest_point_distance(g, x)
    d.sum().backward()
    gx = x.grad.cpu().numpy().astype(np.float64)
    pr = prim.cpu().numpy()
    V64, q64 = V.astype(np.float64), q.astype(np.float64)
    Fp = F[pr]
    p64, _, dch = tri_closest64(q64, V64[Fp[:, 0]], V64[Fp[:, 1]], V64[Fp[:, 2]])
    dbest, _, _ = oracle.brute_closest_f64(V, F, q)
    floor = 8 * EPS32 * np.abs(q64).max()
    assert (dch - dbest <= 1e-6 * dbest + floor).all()
    adj = (q64 - p64) / dch[:, None]
    # p carries ~eps32 * |x| absolute error, so the unit vector carries that over d
    tol = 1e-5 + floor / dch[:, None]
    assert (np.abs(gx - adj) <= tol).all(), np.max(np.abs(gx - adj) - tol)
    # central differences of the FP64 distance field, away from the surface
    far = np.nonzero(dbest > 0.02)[0][:256]
    assert len(far) > 100
    A, B, C = V64[F[:, 0]], V64[F[:, 1]], V64[F[:, 2]]

    def dist(pt):
        return tri_closest64(np.repeat(pt[None], len(F), 0), A, B, C)[2].min()
    for i in far:
        fd = np.array([(dist(q64[i] + H * e) - dist(q64[i] - H * e)) / (2 * H) for e in np.eye(3)])
        # the field is C1 but its curvature jumps across face/edge region
        # boundaries, where a central difference errs by O(H / d)
        assert np.abs(gx[i] - fd).max() < 2 * H / dbest[i], (i, gx[i], fd)


def test_grad_vertices_matches_fp64_adjoint_and_fd(gpu):
    """V.grad against the FP64 scatter of -w b_k (x - p)/d over the chosen
    primitives, and against FP64 central differences of the weighted loss for
    sampled vertices; queries closer to the surface than 100 H get weight 0 so
    neither the |d| kink nor the O((H/d)^2) truncation enters the differences"""
    V, F = meshes.small_icosphere(3)
    s, g = make(V, F)
    q = near_queries(V, F, 20000, 0.05, 6)
    V64, q64 = V.astype(np.float64), q.astype(np.float64)
    dbest, _, _ = oracle.brute_closest_f64(V, F, q)
    w64 = np.random.default_rng(7).uniform(0.5, 1.5, len(q)) * (dbest > 100 * H)
    x = torch.from_numpy(q).cuda()
    Vt = torch.from_numpy(V).cuda().requires_grad_(True)
    d, p, prim = closest_point_distance(g, x, Vt)
    wts = torch.from_numpy(w64.astype(np.float32)).cuda()
    (wts * d).sum().backward()
    gV = Vt.grad.cpu().numpy().astype(np.float64)
    pr = prim.cpu().numpy()
    Fp = F[pr]
    p64, bw, dch = tri_closest64(q64, V64[Fp[:, 0]], V64[Fp[:, 1]], V64[Fp[:, 2]])
    unit = (q64 - p64) / dch[:, None]
    ref = np.zeros_like(V64)
    mag = np.zeros(len(V64))
    wf = w64.astype(np.float32).astype(np.float64)
    for k in range(3):
        np.add.at(ref, Fp[:, k], -(wf * bw[:, k])[:, None] * unit)
        np.add.at(mag, Fp[:, k], wf * bw[:, k])
    err = np.abs(gV - ref).max(axis=1)
    assert (err <= 1e-4 * mag + 1e-5).all(), np.max(err - 1e-4 * mag)

    rng = np.random.default_rng(8)
    checked = 0
    for v in rng.choice(len(V), 24, replace=False):
        inc = np.nonzero((F == v).any(1))[0]
        users = np.nonzero(np.isin(pr, inc) & (w64 > 0))[0]
        if len(users) == 0:
            continue

        def f(Vp):
            A, B, C = Vp[F[inc, 0]], Vp[F[inc, 1]], Vp[F[inc, 2]]
            tot = 0.0
            for i in users:
                tot += wf[i] * tri_closest64(np.repeat(q64[i][None], len(inc), 0), A, B, C)[2].min()
            return tot
        fd = np.zeros(3)
        for k in range(3):
            Vp, Vm = V64.copy(), V64.copy()
            Vp[v, k] += H
            Vm[v, k] -= H
            fd[k] = (f(Vp) - f(Vm)) / (2 * H)
        assert np.abs(gV[v] - fd).max() <= 2e-3 * max(mag[v], 1.0), (v, gV[v], fd)
        checked += 1
    assert checked >= 12


def test_optimisation_loop_with_device_refit(gpu):
    """mean d^2 of near-surface points, minimised over the vertex positions;
    after each step the forward distances match an FP64 brute force on the
    current mesh, which holds only if the on-device refit kept the BVH valid"""
    V, F = meshes.small_icosphere(3)
    s, g = make(V, F)
    q = near_queries(V * 1.1, F, 20000, 0.05, 9)
    x = torch.from_numpy(q).cuda()
    Vt = torch.nn.Parameter(torch.from_numpy(V).cuda())
    opt = torch.optim.Adam([Vt], lr=5e-3)
    losses = []
    for step in range(12):
        opt.zero_grad()
        d, p, prim = closest_point_distance(g, x, Vt)
        loss = (d * d).mean()
        loss.backward()
        opt.step()
        losses.append(loss.item())
        if step % 4 == 3:
            Vc = Vt.detach().cpu().numpy()
            sub = np.arange(0, len(q), 97)
            d64, _, _ = oracle.brute_closest_f64(Vc, F, q[sub])
            with torch.no_grad():
                dn, _, _ = closest_point_distance(g, x[sub], Vt)
            assert np.abs(dn.cpu().numpy() - d64).max() < 1e-5
    assert losses[-1] < 0.6 * losses[0], losses


def test_forward_is_deterministic_and_backward_accumulates(gpu):
    V, F = meshes.small_icosphere(4)
    s, g = make(V, F)
    q = near_queries(V, F, 100000, 0.1, 10)
    x = torch.from_numpy(q).cuda()
    Vt = torch.from_numpy(V).cuda().requires_grad_(True)
    d1, _, _ = closest_point_distance(g, x, Vt)
    d2, _, _ = closest_point_distance(g, x, Vt)
    assert torch.equal(d1, d2)
    d1.sum().backward()
    g1 = Vt.grad.clone()
    Vt.grad = None
    d2.sum().backward()
    torch.testing.assert_close(Vt.grad, g1, rtol=1e-5, atol=1e-4)


@pytest.mark.parametrize("name", ["icosphere", "star2d"])
def test_adjoint_orders_agree(gpu, name):
    """the adjoint walked in the caller's order, in the forward's Morton order and
    in that order with the wave pre-reduction: dL/dx bit for bit, dL/dV equal up
    to FP32 summation order; a backward after a forward of another size falls
    back to the caller's order"""
    V, F = meshes.small_icosphere(4) if name == "icosphere" else meshes.small_star(4096)
    s, g = make(V, F)
    q = near_queries(V, F, 60000, 0.1, 12)
    x = torch.from_numpy(q).cuda()
    dim = V.shape[1]
    w = torch.rand(len(q), device="cuda")
    grads = []
    for order, reduce in ((0, 0), (1, 0), (1, 1), (1, 2)):
        g.set_option("bwd_order", order)
        g.set_option("bwd_reduce", reduce)
        xg = x.clone().requires_grad_(True)
        Vt = torch.from_numpy(V).cuda().requires_grad_(True)
        d, _, prim = closest_point_distance(g, xg, Vt)
        (w * d).sum().backward()
        grads.append((xg.grad.clone(), Vt.grad.clone()))
    g.set_option("bwd_order", 1)
    g.set_option("bwd_reduce", 2)
    # every vertex term has magnitude <= |w| (a unit direction times b_k <= 1): the
    # FP32 sums of one vertex in two orders differ by a few ulp of the sum of the
    # |terms| that land on it, which bounds the difference however much they cancel
    pr = prim.cpu().numpy()
    ids = F[pr[pr >= 0]].ravel()
    wt = np.repeat(np.abs(w.cpu().numpy())[pr >= 0], dim)
    abs_sum = torch.from_numpy(np.bincount(ids, wt, len(V)).astype(np.float32)).cuda()[:, None]
    for gx, gV in grads[1:]:
        assert torch.equal(gx, grads[0][0])
        assert ((gV - grads[0][1]).abs() <= 1e-5 * abs_sum + 1e-12).all()
    assert (grads[0][1].abs() > 1e-5 * abs_sum).float().mean() > 0.5  # the bound is not vacuous
    # another query call in between, of another size: the permutation is not used
    xg = x.clone().requires_grad_(True)
    Vt = torch.from_numpy(V).cuda().requires_grad_(True)
    d, _, _ = closest_point_distance(g, xg, Vt)
    closest_point_distance(g, x[:5000])
    (w * d).sum().backward()
    assert torch.equal(xg.grad, grads[0][0])
    assert ((Vt.grad - grads[0][1]).abs() <= 1e-5 * abs_sum + 1e-12).all()
    assert gV.shape == (len(V), dim)
