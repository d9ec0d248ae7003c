"""The allocation-free closest-point entry point
(fcpw_hip_closest_points_device_noalloc) through the C ABI: on two calls in a
row, the second with a smaller batch, its records equal those of
fcpw_hip_closest_points_device bit for bit (3D and 2D, with and without
normals); a batch above its scratch maximum is rejected before any launch."""
import ctypes
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")  # map torch's HIP runtime before the library
from fcpw_amd import synth  # noqa: E402

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "fcpw_amd", "lib", "libfcpw_hip.so")
MAX_QUERIES = 1 << 21
VP, I64 = ctypes.c_void_p, ctypes.c_int64


@pytest.fixture(scope="module")
def lib(gpu):
    L = ctypes.CDLL(LIB)
    L.fcpw_hip_create.argtypes = [ctypes.c_int, ctypes.POINTER(VP)]
    L.fcpw_hip_destroy.argtypes = [VP]
    L.fcpw_hip_last_error.restype = ctypes.c_char_p
    L.fcpw_hip_last_error.argtypes = [VP]
    L.fcpw_hip_build_bvh.argtypes = [VP, ctypes.c_int, VP, ctypes.c_int, VP, ctypes.c_int, ctypes.c_int,
                                     VP, VP, VP, VP, VP, VP]
    for name in ("fcpw_hip_closest_points_device", "fcpw_hip_closest_points_device_noalloc"):
        getattr(L, name).argtypes = [VP, VP, I64, ctypes.c_int, VP, VP]
    return L


def make_context(L, V, F):
    ctx = VP()
    assert L.fcpw_hip_create(torch.cuda.current_device(), ctypes.byref(ctx)) == 0, L.fcpw_hip_last_error(None)
    V = np.ascontiguousarray(V, np.float32)
    F = np.ascontiguousarray(F, np.int32)
    st = L.fcpw_hip_build_bvh(ctx, V.shape[1], V.ctypes.data, len(V), F.ctypes.data, len(F), 1,
                              None, None, None, None, None, None)
    assert st == 0, L.fcpw_hip_last_error(ctx)
    return ctx


def spheres_around(V, n, seed):
    rng = synth.Rng(seed)
    lo, hi = V.min(0).astype(np.float64), V.max(0).astype(np.float64)
    q = synth.f32(synth.uniform_in_box(n, lo, hi, 1.3, rng))
    s = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
    s[:, :V.shape[1]] = torch.from_numpy(q).cuda()
    s[:, 3] = float("inf")
    s[::7, 3] = 0.01  # some bounded searches, most of them misses
    return s


@pytest.mark.parametrize("dim", [3, 2])
def test_noalloc_equals_the_allocating_entry_point(lib, dim):
    if dim == 3:
        v, f = synth.torus_mesh(256, 128, stream_seed=901)
    else:
        v, f = synth.star_polygon(1 << 14, synth.Rng(902))
    ctx = make_context(lib, synth.f32(v), f)
    stream = torch.cuda.current_stream().cuda_stream
    try:
        for k, (n, normals) in enumerate(((600_000, 0), (70_000, 1))):  # the second call is smaller
            sph = spheres_around(synth.f32(v), n, 903 + k)
            ref = torch.full((n, 10), -3.0, dtype=torch.float32, device="cuda")
            got = torch.full((n, 10), -7.0, dtype=torch.float32, device="cuda")
            assert lib.fcpw_hip_closest_points_device(ctx, sph.data_ptr(), n, normals, ref.data_ptr(), stream) == 0
            st = lib.fcpw_hip_closest_points_device_noalloc(ctx, sph.data_ptr(), n, normals, got.data_ptr(), stream)
            assert st == 0, lib.fcpw_hip_last_error(ctx)
            torch.cuda.synchronize()
            assert torch.equal(got.view(torch.int32), ref.view(torch.int32)), (dim, n)
            found = (ref[:, 9].view(torch.int32) != -1).float().mean().item()
            assert 0.5 < found < 1.0
        # above the scratch maximum: rejected on the host, nothing launched
        st = lib.fcpw_hip_closest_points_device_noalloc(ctx, sph.data_ptr(), MAX_QUERIES + 1, 0, got.data_ptr(),
                                                         stream)
        assert st == 1 and b"scratch maximum" in lib.fcpw_hip_last_error(ctx)
        assert lib.fcpw_hip_closest_points_device_noalloc(ctx, sph.data_ptr(), 0, 0, got.data_ptr(), stream) == 0
        torch.cuda.synchronize()
    finally:
        lib.fcpw_hip_destroy(ctx)
