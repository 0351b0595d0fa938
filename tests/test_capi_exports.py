"""CPU: the C-ABI library builds for gfx950, loads, and exports every symbol
include/echo_amd.h declares (no compute calls: no GPU here)."""
import ctypes
import os
import re

import pytest

from libs import REPO

HEADER = os.path.join(REPO, "include", "echo_amd.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(echo_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    syms = declared_symbols()
    for s in ("echo_create", "echo_steps", "echo_rk_stage", "echo_cons2prim_batch",
              "echo_upload_state", "echo_download_state", "echo_destroy"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    from echoamd import LIB_PATH, EXPORTS
    if not os.path.exists(LIB_PATH):
        pytest.skip("libechoamd.so not built")
    lib = ctypes.CDLL(LIB_PATH)
    syms = declared_symbols()
    assert set(syms) == set(EXPORTS), set(syms) ^ set(EXPORTS)
    for s in syms:
        assert hasattr(lib, s), s


def test_config_struct_layout():
    import numpy as np
    from echoamd import EchoConfig
    from libs import make_config
    assert ctypes.sizeof(EchoConfig) == 16 * 4 + 19 * 8
    ic, rc = make_config(nx=7, ny=5, nz=3, tmin=0.4, cfl=0.25)
    c = EchoConfig.from_arrays(ic, rc)
    assert (c.nx, c.ny, c.nz) == (7, 5, 3)
    assert c.tmin == 0.4 and c.cfl == 0.25 and c.tmax == rc[18]


def test_missing_library_fails_loudly(tmp_path):
    import echoamd
    with pytest.raises(RuntimeError):
        echoamd._lib = None
        echoamd.load_library(str(tmp_path / "nope.so"))
    echoamd._lib = None
