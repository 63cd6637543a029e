"""The C-ABI library builds for gfx950 and exports every symbol include/pm3d_hip.h declares (CPU:
no compute calls)."""
import ctypes
import re
from pathlib import Path

import pytest

from pbh_amd import _lib

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    names = set()
    for hdr in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", hdr.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(pm3d_[a-z0-9_]+)\s*\(", text))
    return names


def test_header_declares_entry_points():
    names = declared_symbols()
    assert {"pm3d_create", "pm3d_step", "pm3d_deposit", "pm3d_interp", "pm3d_forces"} <= names


def test_library_exports_every_declared_symbol():
    if not _lib.SO_PATH.exists():
        pytest.skip("library not built (python -m pbh_amd.build)")
    lib = ctypes.CDLL(str(_lib.SO_PATH))
    missing = [n for n in sorted(declared_symbols()) if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_signatures_cover_header():
    assert declared_symbols() == set(_lib.SIGNATURES)


def test_create_error_path_without_gpu():
    """Invalid configurations are rejected before any device call."""
    if not _lib.SO_PATH.exists():
        pytest.skip("library not built")
    lib = _lib.load()
    h = ctypes.c_void_p()
    cfg = _lib.Pm3dCfg(N=12, Np=12, L=1.0, smooth=0.0, precision=0, frac_bits=0)
    assert lib.pm3d_create(ctypes.byref(cfg), ctypes.byref(h)) == -4
    assert b"power of two" in lib.pm3d_last_error(None)
    cfg = _lib.Pm3dCfg(N=16, Np=16, L=-1.0, smooth=0.0, precision=0, frac_bits=0)
    assert lib.pm3d_create(ctypes.byref(cfg), ctypes.byref(h)) == -1
    # fraction bits: FP64 accepts 52 and 40 (40 only up to 2^23 particles), mixed only 31
    for prec, fb, np_side in ((0, 40, 256), (0, 31, 16), (1, 40, 16), (1, 52, 16)):
        cfg = _lib.Pm3dCfg(N=16, Np=np_side, L=1.0, smooth=0.0, precision=prec, frac_bits=fb)
        assert lib.pm3d_create(ctypes.byref(cfg), ctypes.byref(h)) == -1, (prec, fb, np_side)
        assert b"frac_bits" in lib.pm3d_last_error(None)
