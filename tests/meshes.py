"""Test meshes: the reference's own input files (tests/golden/meshes, copied
from the reference tests/input/) and small seeded synthetic meshes."""
import os

import numpy as np

from fcpw_amd import synth

MESH_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "meshes")


def _face_index(tok):
    # parseFaceIndex (scene_loader.inl:25-43): first '/'-separated field, 1-based
    return int(tok.split("/")[0]) - 1


def load_obj(name, dim):
    """OBJ reader following loadTriangleSoupFromOBJFile (scene_loader.inl:100-150)
    for dim 3 and loadLineSegmentSoupFromOBJFile (scene_loader.inl:46-98) for
    dim 2: an 'f' polygon in 2D becomes segments (i, i+1) for i < F-1, an 'l'
    entry contributes its indices directly."""
    pos, idx = [], []
    with open(os.path.join(MESH_DIR, name)) as fh:
        for line in fh:
            tok = line.split()
            if not tok:
                continue
            if tok[0] == "v":
                xyz = [float(t) for t in tok[1:4]]
                pos.append(xyz[:dim])
            elif tok[0] == "f" and dim == 3:
                idx.extend(_face_index(t) for t in tok[1:])
            elif tok[0] in ("f", "l") and dim == 2:
                ids = [_face_index(t) for t in tok[1:]]
                if tok[0] == "f":
                    for i in range(len(ids) - 1):
                        idx.extend([ids[i], ids[(i + 1) % len(ids)]])
                else:
                    idx.extend(ids)
    pos = np.asarray(pos, dtype=np.float32)
    idx = np.asarray(idx, dtype=np.int32).reshape(-1, dim)
    return pos, idx


def small_torus(U=48, V=24, seed=77):
    v, f = synth.torus_mesh(U, V, stream_seed=seed)
    return synth.f32(v), f.astype(np.int32)


def small_icosphere(sub=3, seed=78):
    v, f = synth.icosphere(sub)
    v = synth.radial_noise(v, 0.01, synth.Rng(seed))
    return synth.f32(v), f.astype(np.int32)


def small_star(n=2048, seed=79):
    v, i = synth.star_polygon(n, synth.Rng(seed))
    return synth.f32(v), i.astype(np.int32)


def triangle_soup(n=500, seed=80):
    """independent random triangles (overlapping, degenerate-free)"""
    rng = synth.Rng(seed)
    c = rng.uniform(n * 3, -1.0, 1.0).reshape(n, 1, 3)
    e = rng.uniform(n * 9, -0.2, 0.2).reshape(n, 3, 3)
    v = (c + e).reshape(-1, 3)
    return synth.f32(v), np.arange(3 * n, dtype=np.int32).reshape(n, 3)


def queries_around(positions, n, scale=1.5, seed=81):
    lo, hi = positions.min(0).astype(np.float64), positions.max(0).astype(np.float64)
    return synth.f32(synth.uniform_in_box(n, lo, hi, scale, synth.Rng(seed)))


def rays_from(positions, n, seed=82):
    rng = synth.Rng(seed)
    lo, hi = positions.min(0).astype(np.float64), positions.max(0).astype(np.float64)
    o = synth.uniform_in_box(n, lo, hi, 1.0, rng)
    d = rng.sphere(n, positions.shape[1])
    return synth.f32(o), synth.f32(d)


def degenerate_soup():
    """An icosphere with degenerate and duplicate triangles appended: one
    collapsed to a point, one with a sub-denormal extent, a collinear one, one
    with a repeated vertex, and three duplicates."""
    V, F = small_icosphere(2)
    n = len(V)
    extra_v = [V[0], V[0] + np.float32(1e-30), V[5], (V[5] + V[6]) / 2, V[6]]
    V = np.vstack([V, np.asarray(extra_v, np.float32)]).astype(np.float32)
    extra_f = [[0, 0, 0], [n, n + 1, 0], [n + 2, n + 3, n + 4], [5, 6, 5], F[0], F[0], F[1]]
    F = np.vstack([F, np.asarray(extra_f, np.int32)]).astype(np.int32)
    return V, F
