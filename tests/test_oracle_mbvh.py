"""The vectorized CPU aggregate (oracle.VectorizedBvh, oracle/fcpw_oracle_mbvh.cpp),
the second leg of the CPU baseline: the SAH tree with leaves packed to the SIMD
width (fcpw.inl:812-827, bvh.inl:3-37,113-140), collapsed into 4-wide nodes
(mbvh.inl:56-147) with SIMD leaf groups (mbvh.inl:171-245), queried by the wide
traversal (mbvh.inl:745-803,2034-2197; wide_query_operations.h:127-268).

Pinned by the collapsed tree's invariants, closed-form answers on one triangle
and one segment, and the reference's own check of the vectorized path: equal to
the scalar path and the brute force within 1e-6 (aggregate_tests.cpp:562,
790-807), indices equal except at FP64 ties."""
import numpy as np
import pytest

import oracle
from tests import meshes

EPS = np.float32(np.finfo(np.float32).eps)
WIDTHS = [8] + ([16] if oracle.host_simd_width() == 16 else [])
CASES = {
    "icosphere": meshes.small_icosphere(),
    "torus": meshes.small_torus(),
    "soup": meshes.triangle_soup(),
    "bunny": meshes.load_obj("bunny.obj", 3),
    "star2d": meshes.small_star(),
    "walker2d": meshes.load_obj("walker.obj", 2),
}


def _prim_boxes(V, F):
    P = V[F]
    return (P - EPS).min(axis=1), (P + EPS).max(axis=1)


@pytest.mark.parametrize("width", WIDTHS)
@pytest.mark.parametrize("name", list(CASES))
def test_packed_and_collapsed_tree_structure(name, width):
    V, F = CASES[name]
    dim = V.shape[1]
    vb = oracle.VectorizedBvh(V, F, width=width)
    # the packed binary tree: leaves of at most `width` references, DFS layout
    pmin, pmax, nrefs, offset = vb.binary_nodes()
    assert (nrefs[nrefs > 0] <= width).all()
    bmin, bmax, child, leaf_prim = vb.export()
    n = len(child)
    is_leaf = child[:, 0] < 0
    lo, hi = _prim_boxes(V, F)
    seen = np.zeros(len(F), np.int32)

    def leaf_prims(i):
        first, groups, refs = -child[i, 0] - 1, child[i, 1], child[i, 3]
        assert groups == (refs + width - 1) // width
        lanes = leaf_prim[first:first + groups].ravel()
        assert (lanes[refs:] == lanes[refs - 1]).all()  # padding repeats the last primitive
        return lanes[:refs]

    def node_box(i):
        """the exact union of what node i holds"""
        if is_leaf[i]:
            p = leaf_prims(i)
            return lo[p].min(0), hi[p].max(0)
        used = child[i] != 0x7fffffff
        return bmin[i][:dim, used].min(1), bmax[i][:dim, used].max(1)

    for i in range(n):
        if is_leaf[i]:
            seen[leaf_prims(i)] += 1
            continue
        used = child[i] != 0x7fffffff
        k = int(used.sum())
        assert 2 <= k <= 4 and used[:k].all()
        if k < 4:  # collapse stops early only when every entry is a leaf (mbvh.inl:88-121)
            assert is_leaf[child[i, :k]].all()
        for c in range(k):
            cl, ch = node_box(child[i, c])
            np.testing.assert_array_equal(bmin[i][:dim, c], cl)
            np.testing.assert_array_equal(bmax[i][:dim, c], ch)
    assert (seen == 1).all()
    assert vb.info()["n_leaf_groups"] == len(leaf_prim)


def _check_against(V, F, q, out, ref, tol=1e-6):
    np.testing.assert_array_equal(out["index"] >= 0, ref["index"] >= 0)
    f = ref["index"] >= 0
    assert np.abs(out["d"][f].astype(np.float64) - ref["d"][f]).max() <= tol
    m = f & (out["index"] != ref["index"])
    if m.any():
        tie = oracle.classify_index_mismatch(V, F, q[m], out["index"][m], ref["index"][m])
        assert tie.all(), f"{(~tie).sum()} non-tie index mismatches"


@pytest.mark.parametrize("width", WIDTHS)
@pytest.mark.parametrize("name", list(CASES))
def test_equals_scalar_path_and_brute_force(name, width):
    V, F = CASES[name]
    dim = V.shape[1]
    q = meshes.queries_around(V, 4000, seed=31)
    vb = oracle.VectorizedBvh(V, F, width=width)
    out = vb.closest_points(q)
    b = oracle.Bvh(V, F)
    _check_against(V, F, q, out, b.brute_closest_f32(q))
    _check_against(V, F, q, out, b.closest_points_cpu(q))
    # the closest point is consistent with d and lies on the returned primitive
    scale = np.abs(q).max() + np.abs(V).max()
    p = out["p"][:, :dim].astype(np.float64)
    assert np.abs(np.linalg.norm(q - p, axis=1) - out["d"]).max() <= 8 * EPS * scale
    assert oracle.prim_dist_f64(V, F, out["p"][:, :dim], out["index"]).max() <= 16 * EPS * scale
    # the error against the FP64 brute force
    d64, _, _ = oracle.brute_closest_f64(V, F, q)
    assert np.abs(out["d"] - d64).max() <= 1e-6


@pytest.mark.parametrize("width", WIDTHS)
@pytest.mark.parametrize("heuristic", [oracle.OVERLAP_SURFACE_AREA, oracle.LONGEST_AXIS_CENTER, oracle.VOLUME])
def test_other_heuristics(heuristic, width):
    V, F = CASES["bunny"]
    q = meshes.queries_around(V, 2000, seed=32)
    out = oracle.VectorizedBvh(V, F, width=width, heuristic=heuristic).closest_points(q)
    _check_against(V, F, q, out, oracle.Bvh(V, F).brute_closest_f32(q))


@pytest.mark.parametrize("width", WIDTHS)
@pytest.mark.parametrize("name", ["icosphere", "star2d"])
def test_bounded_radii(name, width):
    """finite radii, a third of them too small to reach the surface"""
    V, F = CASES[name]
    q = meshes.queries_around(V, 3000, seed=33)
    b = oracle.Bvh(V, F)
    full = b.closest_points_cpu(q)
    r2 = (full["d"] * np.where(np.arange(len(q)) % 3 == 0, 0.5, 1.5)) ** 2
    r2 = r2.astype(np.float32)
    out = oracle.VectorizedBvh(V, F, width=width).closest_points(q, r2)
    ref = b.brute_closest_f32(q, r2)
    _check_against(V, F, q, out, ref)
    assert (out["index"] < 0).sum() > 500 and (out["index"] >= 0).sum() > 1500


@pytest.mark.parametrize("width", WIDTHS)
def test_closed_form_triangle_regions(width):
    """one triangle (0,0,0) (1,0,0) (0,1,0): queries in each of the 7 Voronoi
    regions of the wide region test have known p, uv = (bary_a, bary_b) and d"""
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    F = np.array([[0, 1, 2]], np.int32)
    q = np.array([[-1, -1, 2], [2, -0.5, 1], [-0.5, 2, -1], [0.5, -1, 0], [-1, 0.5, 0],
                  [1, 1, 3], [0.25, 0.25, -2]], np.float32)
    p = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0.5, 0, 0], [0, 0.5, 0], [0.5, 0.5, 0],
                  [0.25, 0.25, 0]], np.float64)
    uv = np.array([[1, 0], [0, 1], [0, 0], [0.5, 0.5], [0.5, 0], [0, 0.5], [0.5, 0.25]], np.float64)
    out = oracle.VectorizedBvh(V, F, width=width).closest_points(q)
    assert (out["index"] == 0).all()
    np.testing.assert_allclose(out["p"], p, atol=1e-6)
    np.testing.assert_allclose(out["uv"], uv, atol=1e-6)
    np.testing.assert_allclose(out["d"], np.linalg.norm(q - p, axis=1), rtol=1e-6)


@pytest.mark.parametrize("width", WIDTHS)
def test_closed_form_segment(width):
    V = np.array([[0, 0], [2, 0]], np.float32)
    F = np.array([[0, 1]], np.int32)
    q = np.array([[-1, 1], [1, 3], [3, -2], [0.5, -0.5]], np.float32)
    p = np.array([[0, 0], [1, 0], [2, 0], [0.5, 0]], np.float64)
    t = np.array([0, 0.5, 1, 0.25])
    out = oracle.VectorizedBvh(V, F, width=width).closest_points(q)
    np.testing.assert_allclose(out["p"][:, :2], p, atol=1e-6)
    np.testing.assert_allclose(out["uv"][:, 0], t, atol=1e-6)
    np.testing.assert_allclose(out["d"], np.linalg.norm(q - p, axis=1), rtol=1e-6)


def test_widths_and_thread_counts_agree():
    V, F = CASES["torus"]
    q = meshes.queries_around(V, 5000, seed=34)
    outs = [oracle.VectorizedBvh(V, F, width=w).closest_points(q, nthreads=t) for w in WIDTHS for t in (1, 3)]
    for o in outs[1:]:
        assert np.abs(o["d"] - outs[0]["d"]).max() <= 1e-6
