"""The C-ABI libraries load and export every function their headers declare
(no compute calls: no GPU needed)."""
import ctypes
import os
import re

import pytest

from fooskagpu_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fkm?_[a-z_0-9]+)\s*\(", text)))


@pytest.mark.parametrize("header,loader", [("fooska_mesh.h", _lib.mesh_lib), ("fooska.h", _lib.hip_lib)])
def test_library_exports_every_declared_symbol(header, loader):
    lib = loader()
    names = declared(header)
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_bindings_cover_the_headers():
    bound = {n for n, _, _ in _lib.MESH_API} | {n for n, _, _ in _lib.HIP_API}
    for h in ("fooska_mesh.h", "fooska.h"):
        assert set(declared(h)) <= bound, set(declared(h)) - bound


def test_params_struct_layout_matches_header():
    # 18 int32 + 9 double + 7 double[2]
    assert ctypes.sizeof(_lib.FkParams) == 18 * 4 + 9 * 8 + 14 * 8


def test_create_without_device_fails_cleanly():
    """fk_create reports an error code (never crashes) when the device index
    does not exist."""
    L = _lib.hip_lib()
    p = _lib.FkParams.from_dict({"model": 0, "precision": 0, "rkSteps": 3, "adaptDt": 1, "adaptDtItr": 1})
    h = ctypes.c_void_p()
    rc = L.fk_create(9999, ctypes.byref(p), ctypes.byref(h))
    assert rc < 0 and not h.value
    assert L.fk_last_error()
