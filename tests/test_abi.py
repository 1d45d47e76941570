"""The C-ABI library without a GPU: it loads, exports every function include/rp_hip.h
declares, the Python binding covers them with matching arity, and a compute call on a
machine without a HIP device fails loudly instead of falling back to the CPU."""
import ctypes as C
import os
import re

import pytest

from rptunnel import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rp_hip.h")


def _declared():
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    out = {}
    for m in re.finditer(r"\bint\s+(rp_[a-z0-9_]+)\s*\(([^)]*)\)\s*;", text):
        args = m.group(2).strip()
        out[m.group(1)] = 0 if args in ("", "void") else args.count(",") + 1
    return out


def test_library_exports_every_declared_function():
    decl = _declared()
    assert len(decl) >= 50
    lib = C.CDLL(L.LIB_PATH)
    missing = [n for n in decl if getattr(lib, n, None) is None]
    assert not missing, missing


def test_python_binding_matches_header():
    decl = _declared()
    assert set(L._SIGS) == set(decl)
    for name, nargs in decl.items():
        assert len(L._SIGS[name]) == nargs, name


def test_version_and_error_channel():
    lib = L.lib()
    assert lib.rp_version() == 1
    buf = C.create_string_buffer(8)
    assert lib.rp_last_error(buf, 8) == 0


def _have_gpu():
    import torch
    return torch.cuda.is_available()


@pytest.mark.skipif(_have_gpu(), reason="checks the no-device error path")
def test_no_device_fails_loudly():
    lib = L.lib()
    rc = lib.rp_init(0)
    assert rc == L.RP_ERR_HIP
    assert L.last_error()
    with pytest.raises(L.RpError):
        from rptunnel import Pes, specs
        Pes(specs.quartic())


def test_null_handles_are_argument_errors():
    lib = L.lib()
    assert lib.rp_pes_destroy(None) in (L.RP_OK, L.RP_ERR_ARG)
    assert lib.rp_step_v(None, 1.0) == L.RP_ERR_ARG
    assert lib.rp_batch_run(None, 1, 0, None, 0) == L.RP_ERR_ARG
    assert lib.rp_ring_um(None, None, None) == L.RP_ERR_ARG
    assert "NULL" in L.last_error() or L.last_error()
