"""Many meshes queried together on one GPU.

``BatchedScenes(gpu_scenes)`` binds already transferred ``GPUScene2D/3D``
objects into one batch.  Queries of all scenes are concatenated in scene
order (``set_queries(counts)``: counts[s] queries for scene s).  One call runs
a single persistent launch in which every wave pops work items (``chunk``
queries of one scene) from a device counter until none are left; scenes are
handed out longest-first (queries x log2 primitives).  Results index each
scene's own primitives; vertex gradients are the scenes' gradients
concatenated in scene order.

After ``set_queries`` the query calls allocate nothing and only enqueue
kernels on the current stream, so ``capture`` records them in a HIP graph
(``torch.cuda.CUDAGraph``) for replay.  A member scene that is transferred
again is picked up by the next eager call (not by an existing graph).

The reference GPU path serves one object per ``GPUScene``
(include/fcpw/fcpw_gpu.h:17-20); this batch has no counterpart there.
``separate_find_closest_points_device`` runs the same queries as one plain
launch per scene, the comparison the batch is measured against.
"""
import numpy as np
import torch

from . import _fcpw_amd as _native

__all__ = ["BatchedScenes"]


def _stream(stream):
    if stream is None:
        return torch.cuda.current_stream().cuda_stream
    return stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream)


def _ptr(t):
    return 0 if t is None else t.data_ptr()


class BatchedScenes:
    def __init__(self, gpu_scenes):
        scenes = list(gpu_scenes)
        if not scenes:
            raise ValueError("BatchedScenes needs at least one scene")
        dim = scenes[0].DIM
        if any(g.DIM != dim for g in scenes):
            raise ValueError("all scenes of a batch must have the same dimension")
        self.DIM = dim
        self.scenes = scenes
        self._b = (_native.GPUSceneBatch3D if dim == 3 else _native.GPUSceneBatch2D)(scenes)
        self.counts = None
        self.offsets = None
        self.vertex_offsets = None

    # ---- configuration
    def set_queries(self, counts):
        counts = [int(c) for c in counts]
        if len(counts) != len(self.scenes):
            raise ValueError("set_queries needs one count per scene")
        self._b.set_queries(counts)
        self.counts = np.asarray(counts, np.int64)
        self.offsets = np.concatenate([[0], np.cumsum(self.counts)]).astype(np.int64)
        nv = [g.info("n_vertices") for g in self.scenes]
        self.vertex_offsets = np.concatenate([[0], np.cumsum(nv)]).astype(np.int64)

    @property
    def n_queries(self):
        return int(self.offsets[-1]) if self.offsets is not None else 0

    def set_option(self, key, value):
        self._b._set_option(key, int(value))

    def get_option(self, key):
        return self._b._get_option(key)

    def options(self):
        return {k: self.get_option(k) for k in ("stack", "hold_min", "chunk", "sort", "waves_per_cu")}

    # ---- queries (device tensors; every call is stream-ordered)
    def _check(self, t, rows, cols, dtype, name):
        if t.device.type != "cuda" or t.dtype != dtype or not t.is_contiguous():
            raise TypeError(f"{name} must be a contiguous {dtype} CUDA tensor")
        if t.shape[0] != rows or (cols is not None and (t.dim() != 2 or t.shape[1] != cols)):
            raise ValueError(f"{name} must have {rows} rows" + (f" and {cols} columns" if cols else ""))

    def find_closest_points_device(self, spheres, out, record_normals=False, stream=None):
        """spheres: (n, 4) GPUBoundingSphere records; out: (n, 10) GPUInteraction records."""
        n = self.n_queries
        self._check(spheres, n, 4, torch.float32, "spheres")
        self._check(out, n, 10, torch.float32, "out")
        self._b._closest_points_device(spheres.data_ptr(), bool(record_normals), out.data_ptr(), _stream(stream))

    def closest_points_compact_device(self, points, d=None, p=None, prim=None, slot=None, r2=None, stream=None):
        n = self.n_queries
        self._check(points, n, self.DIM, torch.float32, "points")
        self._b._closest_points_compact_device(points.data_ptr(), _ptr(r2), _ptr(d), _ptr(p), _ptr(prim), _ptr(slot),
                                               _stream(stream))

    def distance_backward_device(self, points, slot, grad_d, grad_points=None, grad_vertices=None, stream=None):
        """grad_vertices (sum of the scenes' vertex counts x dim) is accumulated into."""
        self._b._distance_backward_device(points.data_ptr(), slot.data_ptr(), grad_d.data_ptr(), _ptr(grad_points),
                                          _ptr(grad_vertices), _stream(stream))

    # ---- graphs
    def capture(self, fn, warmup=1):
        """Run fn eagerly `warmup` times (sizing and table upload happen there),
        then record it into a torch.cuda.CUDAGraph and return the graph."""
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            fn()
        return graph

    # ---- the comparison: one plain launch per scene
    def separate_find_closest_points_device(self, spheres, out, record_normals=False, stream=None):
        for s, g in enumerate(self.scenes):
            a, b = int(self.offsets[s]), int(self.offsets[s + 1])
            if b > a:
                g.find_closest_points_device(spheres[a:b], out[a:b], record_normals=record_normals, stream=stream)
