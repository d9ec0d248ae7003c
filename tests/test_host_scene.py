"""Host side of the drop-in: the product's Scene build reproduces the reference
BVH (checked against the oracle's restatement, bit for bit) and the Scene API
keeps the reference's semantics and error cases."""
import numpy as np
import pytest

import fcpw_amd as fcpw
import oracle
from tests import meshes

HEUR = {
    fcpw.AggregateType.Bvh_SurfaceArea: oracle.SURFACE_AREA,
    fcpw.AggregateType.Bvh_OverlapSurfaceArea: oracle.OVERLAP_SURFACE_AREA,
    fcpw.AggregateType.Bvh_LongestAxisCenter: oracle.LONGEST_AXIS_CENTER,
    fcpw.AggregateType.Bvh_Volume: oracle.VOLUME,
    fcpw.AggregateType.Bvh_OverlapVolume: oracle.OVERLAP_VOLUME,
}

CASES = {
    "icosphere": meshes.small_icosphere,
    "torus": meshes.small_torus,
    "soup": meshes.triangle_soup,
    "bunny": lambda: meshes.load_obj("bunny.obj", 3),
    "star2d": meshes.small_star,
    "walker2d": lambda: meshes.load_obj("walker.obj", 2),
}


def scene_of(V, F, agg=fcpw.AggregateType.Bvh_SurfaceArea):
    s = fcpw.Scene3D() if V.shape[1] == 3 else fcpw.Scene2D()
    s.set_object_count(1)
    s.set_object_vertices(V, 0)
    (s.set_object_triangles if V.shape[1] == 3 else s.set_object_line_segments)(F, 0)
    s.build(agg, False)
    return s


def assert_same_tree(s, ob):
    lo, hi, nr, off, order, depth, leaves = s._bvh_nodes()
    olo, ohi, onr, ooff = ob.nodes()
    np.testing.assert_array_equal(lo, olo)
    np.testing.assert_array_equal(hi, ohi)
    np.testing.assert_array_equal(nr, onr)
    np.testing.assert_array_equal(off, ooff)
    np.testing.assert_array_equal(order, ob.prim_order())
    assert depth == ob.max_depth and leaves == ob.n_leaves


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("agg", list(HEUR))
def test_build_matches_oracle(name, agg):
    V, F = CASES[name]()
    assert_same_tree(scene_of(V, F, agg), oracle.Bvh(V, F, heuristic=HEUR[agg]))


@pytest.mark.slow
def test_parallel_build_matches_oracle_at_scale():
    """above the parallel cutoff the subtrees are built on separate threads"""
    from fcpw_amd import synth
    v, f = synth.torus_mesh(256, 128)
    V, F = synth.f32(v), f.astype(np.int32)
    assert_same_tree(scene_of(V, F), oracle.Bvh(V, F))


@pytest.mark.parametrize("name", ["torus", "star2d"])
def test_cpu_refit_matches_oracle(name):
    V, F = CASES[name]()
    s = scene_of(V, F)
    ob = oracle.Bvh(V, F)
    V2 = (V + np.random.default_rng(2).uniform(-0.05, 0.05, V.shape)).astype(np.float32)
    s.update_object_vertices(V2, 0)
    s.refit()
    ob.update_vertices(V2)
    ob.refit()
    assert_same_tree(s, ob)
    # update_object_vertex takes the caller's vertex index
    s.update_object_vertex(V[0], 0, 0)
    V3 = V2.copy()
    V3[0] = V[0]
    ob.update_vertices(V3)
    s.refit()
    ob.refit()
    assert_same_tree(s, ob)


def test_scene_errors():
    V, F = meshes.small_icosphere()
    s = fcpw.Scene3D()
    s.set_object_count(2)
    s.set_object_vertices(V, 0)
    s.set_object_triangles(F, 0)
    s.set_object_vertices(V, 1)
    s.set_object_triangles(F, 1)
    # two objects build a tree per object and one over the instances; silhouettes
    # need a scene of one object
    s.build(fcpw.AggregateType.Bvh_SurfaceArea, False)
    assert s.is_instanced and s.n_objects == 2
    with pytest.raises(RuntimeError, match="exactly one object"):
        s.compute_silhouettes()
    s = fcpw.Scene3D()
    s.set_object_count(1)
    s.set_object_vertices(V, 0)
    with pytest.raises(RuntimeError, match="without primitives"):
        s.build(fcpw.AggregateType.Bvh_SurfaceArea, False)
    s.set_object_triangles(F, 0)
    with pytest.raises(RuntimeError, match="Already set triangles"):
        s.set_object_triangles(F, 0)
    with pytest.raises(ValueError):
        s.build(fcpw.AggregateType.Baseline, False)
    bad = F.copy()
    bad[0, 0] = len(V)
    s2 = fcpw.Scene3D()
    s2.set_object_count(1)
    s2.set_object_vertices(V, 0)
    s2.set_object_triangles(bad, 0)
    with pytest.raises(IndexError):
        s2.build(fcpw.AggregateType.Bvh_SurfaceArea, False)
    with pytest.raises(ValueError):
        s.set_object_vertices(np.zeros((4, 2), np.float32), 0)
    with pytest.raises(RuntimeError, match="not been built"):
        fcpw.Scene3D().refit()


def test_lower_level_setters_match_bulk():
    V, F = meshes.small_icosphere()
    s = fcpw.Scene3D()
    s.set_object_count(1)
    s.set_object_vertex_count(len(V), 0)
    for i, v in enumerate(V):
        s.set_object_vertex(v, i, 0)
    s.set_object_triangle_count(len(F), 0)
    for i, f in enumerate(F):
        s.set_object_triangle(f, i, 0)
    s.build(fcpw.AggregateType.Bvh_SurfaceArea, False)
    assert_same_tree(s, oracle.Bvh(V, F))
