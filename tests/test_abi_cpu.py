"""CPU checks of the C-ABI library: it loads and exports every declared symbol.

No compute call is made (there is no GPU here)."""
import ctypes
import os
import re

import pytest

from openqg_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "oqg.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(oqg_[a-z_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("oqg_create", "oqg_destroy", "oqg_step", "oqg_suppress", "oqg_set_state", "oqg_get_state",
                 "oqg_solve_inhomog", "oqg_run", "oqg_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail(f"{_lib.LIB_PATH} not built (run python -m openqg_amd.build)")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"missing exports: {missing}"


def test_python_binding_covers_header():
    bound = {s[0] for s in _lib.SIGNATURES}
    assert set(declared_functions()) == bound


def test_desc_struct_layout_matches_header():
    # the C struct has 4 ints, 4 doubles, 4x8 doubles, 2 doubles, 4 ints and 10 pointers
    assert ctypes.sizeof(_lib.OqgDesc) == 4 * 4 + 4 * 8 + 32 * 8 + 2 * 8 + 4 * 4 + 10 * 8


def test_load_fails_loudly_without_library(tmp_path, monkeypatch):
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "liboqg.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(_lib.OqgError):
        _lib.load()


def test_create_without_device_reports_error():
    # On a machine with no HIP device oqg_create must fail with a message, not crash.
    try:
        import torch

        if torch.cuda.is_available():
            pytest.skip("device present")
    except ImportError:
        pass
    lib = _lib.load()
    d = _lib.OqgDesc()
    d.nxp = d.nyp = 33
    d.nl = 3
    d.dx = d.dy = 1.0e4
    d.fnot = 1e-4
    for k in range(3):
        d.h[k] = 1000.0
    d.gp[0] = d.gp[1] = 0.02
    h = ctypes.c_void_p()
    rc = lib.oqg_create(ctypes.byref(d), ctypes.byref(h))
    assert rc != 0
    assert lib.oqg_last_error(None)
