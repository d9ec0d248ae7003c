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
kernels on the current stream, so ``capture`` records them in a HIP graph for
replay.  A graph holds the device addresses the batch had at capture: a member
scene transferred again frees its old buffers, and a ``set_queries`` that grows
the batch's workspaces frees the old ones.  The ``BatchGraph`` that ``capture``
returns therefore refuses to replay (``RuntimeError``) after either happened;
the next eager call picks the new state up, and a new ``capture`` records it.

The reference GPU path serves one object per ``GPUScene``
(include/fcpw/fcpw_gpu.h:17-20); this batch has no counterpart there.
``separate_find_closest_points_device`` runs the same queries as one plain
launch per scene, the comparison the batch is measured against.
"""
import numpy as np
import torch

from . import _fcpw_amd as _native

__all__ = ["BatchedScenes", "BatchGraph"]


def _stream(stream):
    if stream is None:
        return torch.cuda.current_stream().cuda_stream
    return stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream)


def _ptr(t):
    return 0 if t is None else t.data_ptr()


class BatchGraph:
    """A captured HIP graph of batch calls (``BatchedScenes.capture``).
    ``replay()`` raises RuntimeError when the batch's device state moved since
    the capture (a member transferred again, or a new ``set_queries``)."""

    def __init__(self, batch, graph):
        self._batch = batch
        self._token = batch._state_token()
        self.graph = graph

    def replay(self):
        if self._batch._state_token() != self._token:
            raise RuntimeError("BatchGraph.replay: a member scene was transferred again or set_queries was called "
                               "after the capture; capture the graph again")
        self.graph.replay()


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
        self._device = int(scenes[0].info("device"))
        self.counts = None
        self.offsets = None
        self.vertex_offsets = None
        self._epoch = 0

    # ---- configuration
    def set_queries(self, counts):
        counts = [int(c) for c in counts]
        if len(counts) != len(self.scenes):
            raise ValueError("set_queries needs one count per scene")
        self._b.set_queries(counts)
        self._epoch += 1
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
        """t: a contiguous `dtype` tensor on the scenes' device with `rows` rows and
        `cols` columns (None: a 1-D tensor)."""
        if not isinstance(t, torch.Tensor) or t.device.type != "cuda" or t.dtype != dtype or not t.is_contiguous():
            raise TypeError(f"{name} must be a contiguous {dtype} CUDA tensor")
        if t.device.index != self._device:
            raise ValueError(f"{name} is on cuda:{t.device.index}, the scenes on cuda:{self._device}")
        if t.shape[0] != rows or (cols is None and t.dim() != 1) or (
                cols is not None and (t.dim() != 2 or t.shape[1] != cols)):
            raise ValueError(f"{name} must have {rows} rows" + (f" and {cols} columns" if cols else " and one dimension"))

    def _state_token(self):
        return (self._epoch, tuple(int(g.info("generation")) for g in self.scenes))

    def find_closest_points_device(self, spheres, out, record_normals=False, stream=None):
        """spheres: (n, 4) GPUBoundingSphere records; out: (n, 10) GPUInteraction records."""
        n = self.n_queries
        self._check(spheres, n, 4, torch.float32, "spheres")
        self._check(out, n, 10, torch.float32, "out")
        self._b._closest_points_device(spheres.data_ptr(), bool(record_normals), out.data_ptr(), _stream(stream))

    def closest_points_compact_device(self, points, d=None, p=None, prim=None, slot=None, r2=None, stream=None):
        """points: (n, dim) float32; optional r2 (n,) float32 squared radii; outputs
        (each optional): d (n,) float32, p (n, dim) float32, prim and slot (n,) int32."""
        n = self.n_queries
        self._check(points, n, self.DIM, torch.float32, "points")
        for t, cols, dtype, name in ((r2, None, torch.float32, "r2"), (d, None, torch.float32, "d"),
                                     (p, self.DIM, torch.float32, "p"), (prim, None, torch.int32, "prim"),
                                     (slot, None, torch.int32, "slot")):
            if t is not None:
                self._check(t, n, cols, dtype, name)
        self._b._closest_points_compact_device(points.data_ptr(), _ptr(r2), _ptr(d), _ptr(p), _ptr(prim), _ptr(slot),
                                               _stream(stream))

    def distance_backward_device(self, points, slot, grad_d, grad_points=None, grad_vertices=None, stream=None):
        """points (n, dim), slot (n,) int32 from the compact forward, grad_d (n,);
        grad_points (n, dim) is written, grad_vertices (sum of the scenes' vertex
        counts x dim) is accumulated into."""
        n = self.n_queries
        self._check(points, n, self.DIM, torch.float32, "points")
        self._check(slot, n, None, torch.int32, "slot")
        self._check(grad_d, n, None, torch.float32, "grad_d")
        if grad_points is not None:
            self._check(grad_points, n, self.DIM, torch.float32, "grad_points")
        if grad_vertices is not None:
            self._check(grad_vertices, int(self.vertex_offsets[-1]), self.DIM, torch.float32, "grad_vertices")
        self._b._distance_backward_device(points.data_ptr(), slot.data_ptr(), grad_d.data_ptr(), _ptr(grad_points),
                                          _ptr(grad_vertices), _stream(stream))

    # ---- graphs
    def capture(self, fn, warmup=1):
        """Run fn eagerly `warmup` times (sizing and table upload happen there),
        then record it into a torch.cuda.CUDAGraph and return it as a BatchGraph,
        whose replay() checks that the batch's device state is still the captured one."""
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            fn()
        return BatchGraph(self, graph)

    # ---- the comparison: one plain launch per scene
    def separate_find_closest_points_device(self, spheres, out, record_normals=False, stream=None):
        for s, g in enumerate(self.scenes):
            a, b = int(self.offsets[s]), int(self.offsets[s + 1])
            if b > a:
                g.find_closest_points_device(spheres[a:b], out[a:b], record_normals=record_normals, stream=stream)
