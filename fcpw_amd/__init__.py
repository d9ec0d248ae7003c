"""fcpw_amd -- MI355X query backend for FCPW.

Drop-in for the reference's GPU Python surface (python/fcpw_bindings.cpp:493-673):
``GPUScene2D/3D`` with ``transfer_to_gpu``, ``refit``, ``intersect`` (rays and
spheres), ``find_closest_points`` and ``find_closest_silhouette_points``, the ``GPU*`` record types and ``GPUInteractionList``
with its bulk ``get_*`` extraction, plus the ``Scene2D/3D`` subset that builds
what the GPU consumes.  Queries run in hand-written HIP kernels for gfx950
(``lib/libfcpw_hip.so``, C ABI in ``include/fcpw_hip.h``); there is no CPU
fallback: importing fails if the native library is missing.

Usage is the reference's::

    import fcpw_amd as fcpw
    scene = fcpw.Scene3D()
    scene.set_object_count(1)
    scene.set_object_vertices(V, 0)
    scene.set_object_triangles(F, 0)
    scene.build(fcpw.AggregateType.Bvh_SurfaceArea, False)
    gpu = fcpw.GPUScene3D(fcpw.slang_directory_path)
    gpu.transfer_to_gpu(scene)
    hits = fcpw.GPUInteractionList()
    gpu.find_closest_points(Q, r2, hits)
    d = hits.get_distances()
"""
import os as _os

import numpy as _np

_HERE = _os.path.dirname(_os.path.abspath(__file__))

# One HIP runtime per process: torch's libtorch_hip needs the unversioned
# libamdhip64.so from its own directory, while libfcpw_hip.so needs
# libamdhip64.so.7.  Loading torch first lets the dynamic loader satisfy
# libfcpw_hip.so with the runtime torch already mapped (same soname); the other
# order maps two runtimes, and the one initialised second finds no device.
try:
    import torch as _torch  # noqa: F401
except ImportError:
    _torch = None

try:
    from . import _fcpw_amd as _native
except ImportError as e:  # no silent fallback: the HIP library is the product
    raise ImportError(
        f"fcpw_amd: the native module is not built ({e}). "
        "Run `python -c 'import __graft_entry__ as g; g.build()'` or "
        "`python fcpw_amd/_build.py` in the repository root.") from e

PrimitiveType = _native.PrimitiveType
AggregateType = _native.AggregateType
Scene2D = _native.Scene2D
Scene3D = _native.Scene3D
GPUFloat2 = _native.GPUFloat2
GPUFloat3 = _native.GPUFloat3
GPURay = _native.GPURay
GPUBoundingSphere = _native.GPUBoundingSphere
GPUInteraction = _native.GPUInteraction
GPUFloat3List = _native.GPUFloat3List
GPURayList = _native.GPURayList
GPUBoundingSphereList = _native.GPUBoundingSphereList
GPUInteractionList = _native.GPUInteractionList
Uint32List = _native.Uint32List
UINT_MAX = _native.UINT_MAX

# The reference exports the directory of its packaged Slang shaders
# (python/fcpw/__init__.py:16-20) for GPUScene's first argument.  No shader is
# compiled at run time here; the constant is kept so callers work unchanged.
slang_directory_path = _HERE

# numpy view of GPUInteraction / GPURay / GPUBoundingSphere records
INTERACTION_DTYPE = _np.dtype([("p", "<f4", (3,)), ("n", "<f4", (3,)), ("uv", "<f4", (2,)),
                               ("d", "<f4"), ("index", "<u4")])
RAY_DTYPE = _np.dtype([("o", "<f4", (3,)), ("d", "<f4", (3,)), ("dinv", "<f4", (3,)), ("tmax", "<f4")])
SPHERE_DTYPE = _np.dtype([("c", "<f4", (3,)), ("r2", "<f4")])


def _is_list(x, cls):
    return isinstance(x, cls)


class _GPUSceneMixin:
    """Reference signatures on top of the native GPUScene classes."""

    def intersect(self, *args, **kwargs):
        """intersect(ray_origins, ray_directions, ray_distance_bounds, interactions, check_for_occlusion=False)
        intersect(rays, interactions, check_for_occlusion=False)
        intersect(sphere_centers, sphere_squared_radii, rand_nums, interactions)
        intersect(bounding_spheres, rand_nums, interactions)
        (python/fcpw_bindings.cpp:596-611 overloads)"""
        if "rays" in kwargs or (args and _is_list(args[0], GPURayList)):
            return self._intersect_rays(*args, **kwargs)
        if "bounding_spheres" in kwargs or (args and _is_list(args[0], GPUBoundingSphereList)):
            return self._intersect_spheres(*args, **kwargs)
        if "sphere_centers" in kwargs or (len(args) >= 3 and _np.ndim(args[1]) == 1 and _np.ndim(args[2]) == 2):
            return self._intersect_spheres_arrays(*args, **kwargs)
        return self._intersect_arrays(*args, **kwargs)

    def find_closest_points(self, *args, **kwargs):
        """find_closest_points(query_points, squared_max_radii, interactions, record_normals=False)
        find_closest_points(bounding_spheres, interactions, record_normals=False)
        object_indices= (optional, int32 array with one entry per query) receives each
        result's object index (Interaction::objectIndex) or -1 on a miss."""
        if "bounding_spheres" in kwargs or (args and _is_list(args[0], GPUBoundingSphereList)):
            return self._find_closest_points_spheres(*args, **kwargs)
        return self._find_closest_points_arrays(*args, **kwargs)

    def find_closest_silhouette_points(self, *args, **kwargs):
        """find_closest_silhouette_points(query_points, squared_max_radii, flip_normal_orientation, interactions,
                                          squared_min_radius=0.0, precision=1e-3)
        find_closest_silhouette_points(bounding_spheres, flip_normal_orientation, interactions,
                                       squared_min_radius=0.0, precision=1e-3)
        (python/fcpw_bindings.cpp:620-627,664-671).  Needs a scene on which
        compute_silhouettes() was called before build()."""
        if "bounding_spheres" in kwargs or (args and _is_list(args[0], GPUBoundingSphereList)):
            args = list(args)
            if len(args) > 1 and not isinstance(args[1], Uint32List):
                args[1] = Uint32List([int(v) for v in _np.asarray(args[1]).ravel()])
            if "flip_normal_orientation" in kwargs and not isinstance(kwargs["flip_normal_orientation"], Uint32List):
                kwargs["flip_normal_orientation"] = Uint32List(
                    [int(v) for v in _np.asarray(kwargs["flip_normal_orientation"]).ravel()])
            return self._find_closest_silhouette_points(*args, **kwargs)
        return self._find_closest_silhouette_points_arrays(*args, **kwargs)

    # -- device-resident queries (torch tensors) ---------------------------------
    # Every tensor is checked on the host before its pointer reaches the C ABI,
    # which can only check alignment: a wrong shape, dtype or device would
    # otherwise become an out-of-bounds device access.
    def find_closest_points_device(self, spheres, out, record_normals=False, stream=None, object_index=None):
        """spheres: (n, 4) float32 CUDA tensor of GPUBoundingSphere records;
        out: (n, 10) float32 CUDA tensor receiving GPUInteraction records;
        object_index: optional (n,) int32 CUDA tensor receiving object indices."""
        dev = self._device()
        n = _check_tensor(spheres, None, 4, _F32, "spheres", dev)
        _check_tensor(out, n, 10, _F32, "out", dev)
        self._closest_points_device(spheres.data_ptr(), n, bool(record_normals), out.data_ptr(),
                                    _stream_handle(spheres, stream), _object_ptr(object_index, n, dev))

    def intersect_device(self, rays, out, check_for_occlusion=False, stream=None, object_index=None):
        """rays: (n, 10) float32 CUDA tensor of GPURay records; out: (n, 10) float32;
        object_index: optional (n,) int32 CUDA tensor receiving object indices."""
        dev = self._device()
        n = _check_tensor(rays, None, 10, _F32, "rays", dev)
        _check_tensor(out, n, 10, _F32, "out", dev)
        self._intersect_device(rays.data_ptr(), n, bool(check_for_occlusion), out.data_ptr(),
                               _stream_handle(rays, stream), _object_ptr(object_index, n, dev))

    def find_closest_silhouette_points_device(self, spheres, flip_normal_orientation, out, squared_min_radius=0.0,
                                              precision=1e-3, stream=None):
        """spheres: (n, 4) GPUBoundingSphere records, flip_normal_orientation: (n,) int32/uint32,
        out: (n, 10) GPUInteraction records (CUDA tensors)."""
        dev = self._device()
        n = _check_tensor(spheres, None, 4, _F32, "spheres", dev)
        fl = flip_normal_orientation
        _check_tensor(fl, n, 0, _I32, "flip_normal_orientation", dev)
        _check_tensor(out, n, 10, _F32, "out", dev)
        self._closest_silhouette_points_device(spheres.data_ptr(), fl.data_ptr(), n, float(squared_min_radius),
                                               float(precision), out.data_ptr(), _stream_handle(spheres, stream))

    def intersect_spheres_device(self, spheres, rand_nums, out, stream=None, object_index=None):
        """spheres: (n, 4) GPUBoundingSphere records, rand_nums: (n, 3) float32 in [0, 1),
        out: (n, 10) GPUInteraction records, object_index: optional (n,) int32 (CUDA tensors)."""
        dev = self._device()
        n = _check_tensor(spheres, None, 4, _F32, "spheres", dev)
        _check_tensor(rand_nums, n, 3, _F32, "rand_nums", dev)
        _check_tensor(out, n, 10, _F32, "out", dev)
        self._intersect_spheres_device(spheres.data_ptr(), rand_nums.data_ptr(), n, out.data_ptr(),
                                       _stream_handle(spheres, stream), _object_ptr(object_index, n, dev))

    def _device(self):
        return int(self._info("device"))

    def set_option(self, key, value):
        self._set_option(key, int(value))

    def get_option(self, key):
        return self._get_option(key)

    def info(self, key):
        return self._info(key)


_F32 = ("float32",)
_I32 = ("int32", "uint32")


def _check_tensor(t, rows, cols, dtypes, name, device=None):
    """Checks that t is a contiguous CUDA tensor of one of `dtypes` (names) on
    `device`, with `rows` rows (None: any) and `cols` columns (0: a 1-D tensor).
    Returns its row count.  Raises TypeError / ValueError before any launch."""
    if not hasattr(t, "data_ptr") or not hasattr(t, "is_cuda"):
        raise TypeError(f"{name} must be a torch tensor")
    if not t.is_cuda:
        raise TypeError(f"{name} must be a CUDA tensor")
    if str(t.dtype).replace("torch.", "") not in dtypes:
        raise TypeError(f"{name} must have dtype {' or '.join(dtypes)}, not {t.dtype}")
    if device is not None and t.device.index != device:
        raise ValueError(f"{name} is on cuda:{t.device.index}, the scene on cuda:{device}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if cols == 0:
        if t.dim() != 1:
            raise ValueError(f"{name} must be a 1-D tensor")
    elif t.dim() != 2 or t.shape[1] != cols:
        raise ValueError(f"{name} must have shape (n, {cols}), not {tuple(t.shape)}")
    if rows is not None and t.shape[0] != rows:
        raise ValueError(f"{name} must have {rows} rows, not {t.shape[0]}")
    return int(t.shape[0])


def _object_ptr(t, n, device=None):
    if t is None:
        return 0
    _check_tensor(t, n, 0, _I32, "object_index", device)
    return t.data_ptr()


def _stream_handle(t, stream):
    import torch
    if stream is None:
        stream = torch.cuda.current_stream(t.device)
    return stream.cuda_stream


class GPUScene2D(_GPUSceneMixin, _native.GPUScene2D):
    DIM = 2


class GPUScene3D(_GPUSceneMixin, _native.GPUScene3D):
    DIM = 3


__all__ = ["PrimitiveType", "AggregateType", "Scene2D", "Scene3D", "GPUFloat2", "GPUFloat3", "GPURay",
           "GPUBoundingSphere", "GPUInteraction", "GPUFloat3List", "GPURayList", "GPUBoundingSphereList",
           "GPUInteractionList", "Uint32List", "GPUScene2D", "GPUScene3D", "slang_directory_path", "UINT_MAX",
           "INTERACTION_DTYPE", "RAY_DTYPE", "SPHERE_DTYPE"]
