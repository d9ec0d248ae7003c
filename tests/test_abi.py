"""The C ABI library loads, exports every function include/fcpw_hip.h declares,
and reports errors without a GPU (no compute calls here)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fcpw_hip.h")
LIB = os.path.join(ROOT, "fcpw_amd", "lib", "libfcpw_hip.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fcpw_hip_\w+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.fail("libfcpw_hip.so is not built; run __graft_entry__.build()")
    L = ctypes.CDLL(LIB)
    L.fcpw_hip_last_error.restype = ctypes.c_char_p
    L.fcpw_hip_last_error.argtypes = [ctypes.c_void_p]
    return L


def test_every_declared_symbol_is_exported(lib):
    names = declared_functions()
    assert len(names) >= 17
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_create_without_gpu_reports_an_error(lib):
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is visible")
    except ImportError:
        pass
    ctx = ctypes.c_void_p()
    st = lib.fcpw_hip_create(0, ctypes.byref(ctx))
    assert st == 3  # FCPW_HIP_ERR_HIP
    assert not ctx.value
    assert b"no HIP device" in lib.fcpw_hip_last_error(None)


def test_null_context_is_rejected(lib):
    assert lib.fcpw_hip_destroy(None) == 1
    assert lib.fcpw_hip_refit(None, None) == 1
    assert lib.fcpw_hip_closest_points(None, None, ctypes.c_int64(0), 0, None) == 1
    assert lib.fcpw_hip_set_option(None, b"block", 64) == 1


def test_record_layouts_match_the_reference():
    import fcpw_amd
    assert fcpw_amd.INTERACTION_DTYPE.itemsize == 40  # GPUInteraction
    assert fcpw_amd.RAY_DTYPE.itemsize == 40          # GPURay
    assert fcpw_amd.SPHERE_DTYPE.itemsize == 16       # GPUBoundingSphere
    text = open(HEADER).read()
    assert "#define FCPW_HIP_UINT_MAX 4294967295u" in text
    assert fcpw_amd.UINT_MAX == 4294967295


def test_interaction_list_defaults_and_bulk_getters():
    import fcpw_amd as fcpw
    lst = fcpw.GPUInteractionList()
    lst.append(fcpw.GPUInteraction())
    i = fcpw.GPUInteraction()
    i.d = 2.5
    i.index = 7
    i.p = fcpw.GPUFloat3(1, 2, 3)
    i.uv = fcpw.GPUFloat2(0.25, 0.5)
    lst.append(i)
    assert lst.get_distances().tolist() == [np.finfo(np.float32).max, 2.5]
    assert lst.get_indices().tolist() == [4294967295, 7]
    assert lst.get_positions().tolist() == [[0, 0, 0], [1, 2, 3]]
    assert lst.get_uvs().tolist() == [[0, 0], [0.25, 0.5]]
    assert lst.get_normals().shape == (2, 3)
    r = fcpw.GPURay(fcpw.GPUFloat3(0, 0, 0), fcpw.GPUFloat3(2, -4, 0.5))
    assert (r.inv_d.x, r.inv_d.y, r.inv_d.z) == (0.5, -0.25, 2.0)
    assert r.t_max == np.finfo(np.float32).max
