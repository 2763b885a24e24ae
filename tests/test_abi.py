"""The C ABI loads, exports every entry point include/phx.h declares, and reports errors with
status codes instead of crashing (CPU only; no compute call needs a GPU here)."""
import ctypes as C
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from phantasyenginetestbed_amd import phx

HEADER = ROOT / "include" / "phx.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(phx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 20
    L = phx.lib()
    for name in names:
        assert hasattr(L, name), f"libphx.so does not export {name}"
    # nm agrees (dynamic symbol table)
    out = subprocess.run(["nm", "-D", "--defined-only", str(phx.LIB_PATH)], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (phx_[a-z0-9_]+)", out))
    assert set(names) <= exported


def test_python_binding_covers_header():
    assert set(declared_functions()) == set(phx.exported_symbols())


def test_struct_sizes_match_header_layouts():
    # the reference's static_asserts: BVHNode 64, TriangleVertices 48, TriangleData 64,
    # RayIn 32, RayHit 16, Material 48; SphereLight is 36 with padding
    assert phx.NODE.itemsize == 64
    assert phx.TRI.itemsize == 48
    assert phx.TRIDATA.itemsize == 64
    assert phx.RAY.itemsize == 32
    assert phx.HIT.itemsize == 16
    assert phx.MATERIAL.itemsize == 48
    assert phx.LIGHT.itemsize == 36
    assert C.sizeof(phx.CameraDef) == 48


def test_version_and_last_error():
    assert phx.lib().phx_version().decode().startswith("phx ")
    with pytest.raises(phx.PhxError) as e:
        phx.BVH(np.zeros(0, phx.TRI))
    assert e.value.status == phx.ERR_INVALID_SCENE
    assert "no triangles" in phx.lib().phx_last_error().decode()


def test_invalid_handles_are_reported_not_dereferenced():
    L = phx.lib()
    bogus = C.c_void_p(0x1234)
    rays = np.zeros(4, phx.RAY)
    hits = np.zeros(4, phx.HIT)
    assert L.phx_cast_closest(None, None, rays.ctypes.data, 4, hits.ctypes.data, 0) == \
        phx.ERR_INVALID_HANDLE
    assert L.phx_cast_closest(bogus, bogus, rays.ctypes.data, 4, hits.ctypes.data, 0) == \
        phx.ERR_INVALID_HANDLE
    assert L.phx_cast_anyhit(bogus, bogus, rays.ctypes.data, 4, hits.ctypes.data, 0) == \
        phx.ERR_INVALID_HANDLE
    assert L.phx_synchronize(bogus) == phx.ERR_INVALID_HANDLE
    assert L.phx_context_stream(bogus) is None
    L.phx_scene_release(bogus)  # ignored
    L.phx_destroy(bogus)        # ignored
    L.phx_bvh_release(bogus)    # ignored
    assert L.phx_bvh_arrays(bogus, None, None, None, None, None) == phx.ERR_INVALID_HANDLE


def test_create_without_device_reports_no_device():
    import torch  # noqa: F401  (only to learn whether a GPU is visible here)
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    h = C.c_void_p()
    assert phx.lib().phx_create(0, C.byref(h)) == phx.ERR_NO_DEVICE
    assert h.value is None


def test_argument_checks_need_no_device():
    L = phx.lib()
    assert L.phx_bvh_build(None, None, 10, 0, C.byref(C.c_void_p())) == phx.ERR_INVALID_ARGUMENT
    assert L.phx_scene_random_triangles(0, 1, 0.1, 0.2, None, None) == phx.ERR_INVALID_ARGUMENT
    cam = phx.CameraDef()
    f3 = (C.c_float * 3)(0, 0, 0)
    assert L.phx_camera_def(f3, f3, f3, 1.0, 0.0, 1.0, C.byref(cam)) == phx.ERR_INVALID_ARGUMENT
