"""The C-ABI libraries load and export every function their headers declare
(no compute calls: no GPU needed)."""
import ctypes
import os
import re
import subprocess

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


def test_params_struct_layout_matches_header(tmp_path):
    """Every FkParams field sits at the offset the C compiler gives it."""
    names = [n for n, _ in _lib.FkParams._fields_]
    src = tmp_path / "layout.c"
    src.write_text('#include <stddef.h>\n#include <stdio.h>\n#include "fooska.h"\nint main(void) {\n'
                   + "".join(f'  printf("%zu\\n", offsetof(fk_params, {n}));\n' for n in names)
                   + '  printf("%zu\\n", sizeof(fk_params));\n  return 0;\n}\n')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    want = [getattr(_lib.FkParams, n).offset for n in names] + [ctypes.sizeof(_lib.FkParams)]
    assert got == want


def test_create_without_device_fails_cleanly():
    """fk_create reports an error code (never crashes) when the device index
    does not exist."""
    L = _lib.hip_lib()
    p = _lib.FkParams.from_dict({"model": 0, "precision": 0, "rkSteps": 3, "adaptDt": 1, "adaptDtItr": 1})
    h = ctypes.c_void_p()
    rc = L.fk_create(9999, ctypes.byref(p), ctypes.byref(h))
    assert rc < 0 and not h.value
    assert L.fk_last_error()


@pytest.mark.parametrize("over,msg", [
    (dict(thermFntRlxFlag=1, relaxMode=_lib.RELAX_MLP), "finite-rate"),
    (dict(thermFntRlxFlag=2), "thermFntRlxFlag"),
    (dict(inletP=2e6), "inletT"),
    (dict(inletP=2e6, inletT=-5.0), "inletT"),
])
def test_create_rejects_bad_relaxation_and_inlet_parameters(over, msg):
    """Parameter checks run before any device call: finite-rate relaxation
    only with the plain Newton relaxation, an inlet needs T > 0."""
    from fooskagpu_amd import cases
    L = _lib.hip_lib()
    p = _lib.FkParams.from_dict(dict(cases._two_phase_params(), **over))
    h = ctypes.c_void_p()
    rc = L.fk_create(0, ctypes.byref(p), ctypes.byref(h))
    assert rc == _lib.EINVAL and not h.value
    assert msg in L.fk_last_error().decode()
