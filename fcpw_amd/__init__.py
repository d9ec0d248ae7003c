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
    def find_closest_points_device(self, spheres, out, record_normals=False, stream=None, object_index=None):
        """spheres: (n, 4) float32 CUDA tensor of GPUBoundingSphere records;
        out: (n, 10) float32 CUDA tensor receiving GPUInteraction records;
        object_index: optional (n,) int32 CUDA tensor receiving object indices."""
        n = spheres.shape[0]
        self._closest_points_device(spheres.data_ptr(), n, bool(record_normals), out.data_ptr(),
                                    _stream_handle(spheres, stream), _object_ptr(object_index, n))

    def intersect_device(self, rays, out, check_for_occlusion=False, stream=None, object_index=None):
        """rays: (n, 10) float32 CUDA tensor of GPURay records; out: (n, 10) float32;
        object_index: optional (n,) int32 CUDA tensor receiving object indices."""
        n = rays.shape[0]
        self._intersect_device(rays.data_ptr(), n, bool(check_for_occlusion), out.data_ptr(),
                               _stream_handle(rays, stream), _object_ptr(object_index, n))

    def find_closest_silhouette_points_device(self, spheres, flip_normal_orientation, out, squared_min_radius=0.0,
                                              precision=1e-3, stream=None):
        """spheres: (n, 4) GPUBoundingSphere records, flip_normal_orientation: (n,) int32/uint32,
        out: (n, 10) GPUInteraction records (CUDA tensors)."""
        n = spheres.shape[0]
        fl = flip_normal_orientation
        if fl.shape[0] != n or not fl.is_contiguous() or fl.element_size() != 4:
            raise ValueError("flip_normal_orientation must be a contiguous (n,) 32-bit tensor")
        self._closest_silhouette_points_device(spheres.data_ptr(), fl.data_ptr(), n, float(squared_min_radius),
                                               float(precision), out.data_ptr(), _stream_handle(spheres, stream))

    def intersect_spheres_device(self, spheres, rand_nums, out, stream=None, object_index=None):
        """spheres: (n, 4) GPUBoundingSphere records, rand_nums: (n, 3) float32 in [0, 1),
        out: (n, 10) GPUInteraction records, object_index: optional (n,) int32 (CUDA tensors)."""
        n = spheres.shape[0]
        if rand_nums.shape[0] != n or rand_nums.shape[-1] != 3 or not rand_nums.is_contiguous():
            raise ValueError("rand_nums must be a contiguous (n, 3) tensor")
        self._intersect_spheres_device(spheres.data_ptr(), rand_nums.data_ptr(), n, out.data_ptr(),
                                       _stream_handle(spheres, stream), _object_ptr(object_index, n))

    def set_option(self, key, value):
        self._set_option(key, int(value))

    def get_option(self, key):
        return self._get_option(key)

    def info(self, key):
        return self._info(key)


def _object_ptr(t, n):
    if t is None:
        return 0
    if t.shape[0] != n or not t.is_contiguous() or t.element_size() != 4:
        raise ValueError("object_index must be a contiguous (n,) 32-bit tensor")
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
