"""The C-ABI library loads and exports every function declared in include/*.h,
and the ctypes binding declares each of them.  No GPU compute calls."""
import os
import re
import subprocess

import pytest

from openems_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in ("ems_hip.h", "ems_operator.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(ems[a-z]*_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_every_declared_symbol_is_exported(native):
    out = subprocess.run(["nm", "-D", "--defined-only", _native.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    decl = declared_functions()
    assert len(decl) > 40
    missing = sorted(decl - exported)
    assert not missing, f"declared but not exported: {missing}"


def test_binding_covers_every_declaration(native):
    bound = {n for n, _, _ in _native.OPERATOR_API + _native.ENGINE_API}
    missing = sorted(declared_functions() - bound)
    assert not missing, f"declared but not bound: {missing}"
    for name in bound:
        assert hasattr(native, name)


def test_errors_are_reported_not_fatal(native):
    import ctypes as C
    h = C.c_void_p()
    assert native.emsop_create(C.byref(h)) == 0
    assert native.emsop_build(h, 10) < 0                  # no grid, no signal
    assert b"" != native.emsop_last_error()
    native.emsop_destroy(h)
    assert native.ems_create(None, None, None, None, None, None, 0) == -1
    assert b"null" in native.ems_last_error()


def test_engine_refuses_without_device_or_raises_cleanly(native):
    """On a host without a GPU the engine must fail loudly (no CPU fallback)."""
    from openems_amd import scenes
    from openems_amd.engine import EngineHIP
    from openems_amd.operator import Operator
    if native.ems_device_count() > 0:
        pytest.skip("a GPU is visible; covered by the GPU tests")
    op = Operator(scenes.pec_cavity(12, 10))
    with pytest.raises(RuntimeError):
        EngineHIP(op)


def test_cpp_host_classes_compile(native):
    """include/openems_hip.hpp (Operator_HIP / Engine_HIP over the C ABI)
    compiles on its own, and the check program the GPU tests run is built
    against the library."""
    import subprocess
    import tempfile
    from openems_amd import build as B
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "tu.cpp")
        with open(src, "w") as f:
            f.write('#include "openems_hip.hpp"\nint main() { return openems_hip::Operator_HIP::New() ? 0 : 1; }\n')
        r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                            src], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
    assert os.path.exists(B.MIRROR_BIN)
