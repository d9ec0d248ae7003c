"""Scenes of several objects and transformed instances for the instancing
tests: the objects, their instance transforms ((dim+1) x (dim+1) affine
matrices with uniform scaling, as fcpw.h:68 requires) and the FP64 world-space
soup every instance spans."""
import numpy as np

from fcpw_amd import synth
from tests import meshes


def rotation(dim, rng, angle=None):
    """a random rotation (3D: axis-angle from the generator; 2D: angle)"""
    if dim == 2:
        a = rng.uniform(1, 0.0, 2 * np.pi)[0] if angle is None else angle
        return np.array([[np.cos(a), -np.sin(a)], [np.sin(a), np.cos(a)]])
    ax = rng.sphere(1, 3)[0]
    a = rng.uniform(1, 0.0, 2 * np.pi)[0] if angle is None else angle
    K = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
    return np.eye(3) + np.sin(a) * K + (1 - np.cos(a)) * K @ K


def affine(dim, R, scale, shift):
    M = np.eye(dim + 1)
    M[:dim, :dim] = scale * R
    M[:dim, dim] = shift
    return M


def _random_transforms(dim, n, rng, box=3.0, scales=(0.4, 1.6)):
    out = []
    for _ in range(n):
        R = rotation(dim, rng)
        s = rng.uniform(1, scales[0], scales[1])[0]
        t = rng.uniform(dim, -box, box)
        out.append(affine(dim, R, s, t))
    return out


def _scene(name):
    rng = synth.Rng(synth.seed(90, sum(map(ord, name))))
    if name == "mixed3d":
        objs = [meshes.small_torus(), meshes.small_icosphere(2), meshes.triangle_soup(200)]
        tr = [[], _random_transforms(3, 3, rng) + [affine(3, np.eye(3), 1.0, (0.5, -2.0, 1.0))],
              _random_transforms(3, 2, rng)]
    elif name == "objects3d":
        objs = [meshes.small_torus(), meshes.small_icosphere(2), meshes.triangle_soup(200)]
        tr = [[], [], []]
    elif name == "many3d":
        objs = [meshes.small_icosphere(1), meshes.small_torus(16, 8)]
        tr = [_random_transforms(3, 40, rng, box=6.0), _random_transforms(3, 25, rng, box=6.0)]
    elif name == "single3d":
        objs = [meshes.small_icosphere(2)]
        tr = [_random_transforms(3, 1, rng)]
    elif name == "mixed2d":
        objs = [meshes.small_star(512), meshes.load_obj("walker.obj", 2), meshes.load_obj("plus-shape.obj", 2)]
        tr = [_random_transforms(2, 1, rng), [], _random_transforms(2, 3, rng)]
    elif name == "many2d":
        objs = [meshes.small_star(256)]
        tr = [_random_transforms(2, 30, rng, box=5.0)]
    elif name == "crowd3d":
        # 4,000 instances of two small meshes (a deep instance tree) and 50
        # untransformed objects
        objs = [meshes.small_icosphere(0), meshes.small_torus(8, 4)]
        tr = [_random_transforms(3, 2500, rng, box=40.0, scales=(0.2, 1.0)),
              _random_transforms(3, 1500, rng, box=40.0, scales=(0.2, 1.0))]
        for k in range(50):
            V, F = meshes.small_icosphere(0)
            objs.append(((V * np.float32(0.5) + np.float32([45.0 + 2.0 * (k % 10), 2.0 * (k // 10), 0.0]))
                         .astype(np.float32), F))
            tr.append([])
    else:
        raise KeyError(name)
    return objs, tr


NAMES = ["mixed3d", "objects3d", "many3d", "single3d", "mixed2d", "many2d"]


def scene(name):
    """(objects [(V, F)], transforms [list of matrices per object])"""
    return _scene(name)


def world_soup(objs, transforms):
    """All instances' primitives in world space (positions in FP64, rounded to
    FP32 once), with each soup primitive's (object, primitive) pair, in the
    instance order of Scene::build (fcpw.inl:1066-1085)."""
    dim = objs[0][0].shape[1]
    pos, idx, obj, prim = [], [], [], []
    base = 0
    for o, (V, F) in enumerate(objs):
        ms = transforms[o] if len(transforms[o]) else [np.eye(dim + 1)]
        for M in ms:
            W = V.astype(np.float64) @ np.asarray(M)[:dim, :dim].T + np.asarray(M)[:dim, dim]
            pos.append(W)
            idx.append(F + base)
            obj.append(np.full(len(F), o, np.int32))
            prim.append(np.arange(len(F), dtype=np.int32))
            base += len(V)
    return (np.vstack(pos).astype(np.float32), np.vstack(idx).astype(np.int32), np.concatenate(obj),
            np.concatenate(prim))


def queries(objs, transforms, n, seed, scale=1.2):
    W, _, _, _ = world_soup(objs, transforms)
    return meshes.queries_around(W, n, scale=scale, seed=seed)


def rays(objs, transforms, n, seed):
    W, _, _, _ = world_soup(objs, transforms)
    return meshes.rays_from(W, n, seed=seed)
