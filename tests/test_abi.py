"""CPU-side checks of the C-ABI boundary: the library loads, exports every symbol include/dgx.h
declares, and rejects bad arguments before touching the device."""
import ctypes as C
import os
import re

import pytest

from dpgsolver_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "dgx.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(dgx_[a-z0-9_]+)\s*\(", src)))


def test_header_declarations_match_binding_list():
    assert _declared() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    for name in _declared():
        assert hasattr(L, name), name
    assert L.dgx_version().startswith(b"dgx")


def test_no_oracle_symbols_in_product_library():
    L = C.CDLL(_lib.LIB_PATH)
    for name in ("or_create", "or_compute_rhs_d"):
        with pytest.raises(AttributeError):
            getattr(L, name)


@pytest.mark.parametrize("cfg", [
    dict(d=4, p=2, pde=1, num_flux=1),      # bad dimension
    dict(d=2, p=9, pde=1, num_flux=1),      # p out of range
    dict(d=2, p=2, pde=7, num_flux=1),      # unknown pde
    dict(d=2, p=2, pde=1, num_flux=0),      # upwind for Euler
    dict(d=3, p=5, pde=1, num_flux=1),      # HEX p=5 not compiled
])
def test_create_rejects_bad_config_without_device(cfg):
    L = _lib.load()
    c = _lib.DgxConfig()
    c.d, c.p, c.pde, c.num_flux = cfg["d"], cfg["p"], cfg["pde"], cfg["num_flux"]
    h = C.c_void_p()
    rc = L.dgx_create(C.byref(h), C.byref(c))
    assert rc in (-1, -3)
    assert not h.value


def test_null_arguments_are_errors_not_crashes():
    L = _lib.load()
    assert L.dgx_compute_rhs(None, None, None) == -1
    assert L.dgx_ssp_rk3_step(None, None, 0.1, 1, None) == -1
    assert L.dgx_upload_mesh(None, None) == -1
    assert L.dgx_destroy(None) == 0
