"""Host build of scenes with several objects and instance transforms
(Scene::build, fcpw.inl:1052-1126): the per-object trees, the instance list
with each transform's inverse, and the tree over the instances equal the
oracle's restatement bit for bit, for every heuristic, before and after a
refit."""
import numpy as np
import pytest

import fcpw_amd as fcpw
import oracle
from tests import scenes
from tests.test_host_scene import HEUR


def make_scene(objs, tr, agg=fcpw.AggregateType.Bvh_SurfaceArea):
    dim = objs[0][0].shape[1]
    s = fcpw.Scene3D() if dim == 3 else fcpw.Scene2D()
    s.set_object_count(len(objs))
    for o, (V, F) in enumerate(objs):
        s.set_object_vertices(V, o)
        (s.set_object_triangles if dim == 3 else s.set_object_line_segments)(F, o)
        if len(tr[o]):
            s.set_object_instance_transforms(np.asarray(tr[o], np.float32), o)
    s.build(agg, False)
    return s


def instances_of(s):
    return np.frombuffer(s._instances(), dtype=oracle.INSTANCE_DTYPE)


def assert_same_nodes(got, ob):
    lo, hi, nr, off, order, depth, leaves = got
    olo, ohi, onr, ooff = ob.nodes()
    np.testing.assert_array_equal(lo, olo)
    np.testing.assert_array_equal(hi, ohi)
    np.testing.assert_array_equal(nr, onr)
    np.testing.assert_array_equal(off, ooff)
    np.testing.assert_array_equal(order, ob.prim_order())


@pytest.mark.parametrize("name", scenes.NAMES)
def test_instance_list_matches_oracle(name):
    objs, tr = scenes.scene(name)
    dim = objs[0][0].shape[1]
    s = make_scene(objs, tr)
    assert s.is_instanced and s.n_objects == len(objs)
    got = instances_of(s)
    ref = oracle.make_instances(dim, len(objs), tr)
    np.testing.assert_array_equal(got["object"], ref["object"])
    np.testing.assert_array_equal(got["transformed"], ref["transformed"])
    np.testing.assert_array_equal(got["t"], ref["t"])
    np.testing.assert_array_equal(got["tinv"], ref["tinv"])


@pytest.mark.parametrize("agg", list(HEUR))
@pytest.mark.parametrize("name", scenes.NAMES)
def test_trees_match_oracle(name, agg):
    objs, tr = scenes.scene(name)
    dim = objs[0][0].shape[1]
    s = make_scene(objs, tr, agg)
    bv = [oracle.Bvh(V, F, heuristic=HEUR[agg]) for V, F in objs]
    for o, b in enumerate(bv):
        assert_same_nodes(s._object_bvh_nodes(o), b)
    I = oracle.Instances(bv, oracle.make_instances(dim, len(objs), tr), heuristic=HEUR[agg])
    assert_same_nodes(s._instance_tree_nodes(), I)


@pytest.mark.parametrize("name", ["mixed3d", "many2d"])
def test_refit_matches_oracle(name):
    objs, tr = scenes.scene(name)
    dim = objs[0][0].shape[1]
    s = make_scene(objs, tr)
    bv = [oracle.Bvh(V, F) for V, F in objs]
    I = oracle.Instances(bv, oracle.make_instances(dim, len(objs), tr))
    rng = np.random.default_rng(41)
    for o, (V, F) in enumerate(objs):
        V2 = (V + rng.uniform(-0.03, 0.03, V.shape)).astype(np.float32)
        s.update_object_vertices(V2, o)
        bv[o].update_vertices(V2)
    s.refit()
    I.refit()
    for o, b in enumerate(bv):
        assert_same_nodes(s._object_bvh_nodes(o), b)
    assert_same_nodes(s._instance_tree_nodes(), I)


def test_one_object_with_one_transform_is_instanced():
    objs, tr = scenes.scene("single3d")
    s = make_scene(objs, tr)
    assert s.is_instanced
    lo, hi, nr, off, order, depth, leaves = s._instance_tree_nodes()
    assert len(nr) == 1 and nr[0] == 1 and order.tolist() == [0]


def test_transform_errors():
    objs, tr = scenes.scene("single3d")
    V, F = objs[0]
    s = fcpw.Scene3D()
    s.set_object_count(1)
    s.set_object_vertices(V, 0)
    s.set_object_triangles(F, 0)
    with pytest.raises(ValueError, match="affine"):
        s.set_object_instance_transforms(np.eye(3, dtype=np.float32), 0)
    with pytest.raises(IndexError):
        s.set_object_instance_transforms(np.eye(4, dtype=np.float32), 1)
    s.set_object_instance_transforms(np.zeros((4, 4), np.float32), 0)
    with pytest.raises(ValueError, match="singular"):
        s.build(fcpw.AggregateType.Bvh_SurfaceArea, False)


def test_crowd_instance_tree_matches_oracle():
    objs, tr = scenes.scene("crowd3d")
    s = make_scene(objs, tr)
    bv = [oracle.Bvh(V, F) for V, F in objs]
    I = oracle.Instances(bv, oracle.make_instances(3, len(objs), tr))
    assert_same_nodes(s._instance_tree_nodes(), I)
    assert s._instance_tree_nodes()[5] >= 10   # a deep instance tree
