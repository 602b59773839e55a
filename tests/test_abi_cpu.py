"""CPU checks of the C ABI: the built library loads, exports every function
that include/*.h declares, and reports errors (not crashes) without a GPU."""
import ctypes as C
import glob
import os
import re

import pytest

from binarytpm_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        src = re.sub(r"//[^\n]*", "", src)
        src = re.sub(r"^\s*#[^\n]*", "", src, flags=re.M)   # preprocessor lines
        for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\([^;{]*\)\s*;", src):
            name = m.group(1)
            if name not in ("if", "while", "for", "return", "sizeof"):
                names.add(name)
    return sorted(names)


def test_headers_declare_entry_points():
    names = declared_functions()
    assert "btpm_shape_create" in names and "btpm_batch_run" in names
    assert len(names) >= 25


def test_library_exports_every_declared_symbol():
    lib = _lib.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_prototypes_cover_native_api():
    lib = _lib.lib()
    for name, _, _ in _lib.PROTOTYPES:
        assert hasattr(lib, name)


def test_version_and_defaults():
    lib = _lib.lib()
    assert lib.btpm_version() >= 100
    p = _lib.Params()
    lib.btpm_params_default(C.byref(p), 650, 15480.0, 0.01)
    assert p.nx == 56 and p.max_rev == 301 and p.check_after == 5
    assert abs(p.newton_tol - 0.001) < 1e-9


def test_context_without_gpu_reports_error():
    """No GPU in this container: creating a context fails with an error
    code and message, never a crash or a silent CPU fallback."""
    from binarytpm_amd.thermal import Context
    try:
        c = Context(0)
    except _lib.BtpmError as e:
        assert e.code in (_lib.BTPM_ERR_HIP, _lib.BTPM_ERR_ARG)
        assert "device" in str(e) or "hip" in str(e).lower()
        return
    c.close()
    pytest.skip("a GPU is present")
