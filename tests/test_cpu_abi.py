"""The C-ABI library on the host: every declared entry point is exported,
the host-side setup (wdg_host_setup) matches the oracle, and invalid input
is rejected with a status code and a message.  No GPU is touched."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from oracle import wigner1d as W
from real_space_dg_amd import _lib
from tests._cases import build_case

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "wigner_dg.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(wdg_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    path = ROOT / "real_space_dg_amd" / "libwdg.so"
    if not path.exists():
        pytest.fail("libwdg.so is not built: run __graft_entry__.build() first")
    return _lib.load()


def test_every_declared_symbol_is_exported(lib):
    names = declared_functions()
    assert len(names) >= 20, names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding knows each of them
    assert set(names) <= set(_lib.SIGNATURES), sorted(set(names) - set(_lib.SIGNATURES))


def _desc_pot(c):
    p = c["p"]
    mat = c["mat"]
    desc = _lib.wdg_desc(p["N_chi"], p["N_K_chi"], p["N_xi"], p["Nk"], p["L_chi"], p["L_xi"])
    x = np.ascontiguousarray(mat["x"] * 1e-9)
    v = np.ascontiguousarray(c["E"][0, :, 0].astype(float))
    pot = _lib.wdg_potential(x.size, _lib.WDG_FERMI_MSA_3D, _lib.dptr(x), _lib.dptr(v), c["m"][0], c["m"][-1],
                             c["efl"], c["efr"], float(mat["Temp"]), 1.0, 0.0)
    return desc, pot, (x, v)


@pytest.mark.parametrize("ne,deg,nk,bias", [(8, 2, 64, 0.1), (6, 3, 96, 0.0)])
def test_host_setup_matches_oracle(lib, ne, deg, nk, bias):
    c = build_case(ne, deg, nk, bias)
    desc, pot, keep = _desc_pot(c)
    N = ne * (deg + 1) * nk
    rhs = np.zeros(2 * N)
    n = deg + 1
    kmats = np.zeros(5 * n * n)
    assert lib.wdg_host_setup(C.byref(desc), C.byref(pot), _lib.dptr(rhs), _lib.dptr(kmats)) == 0
    _, rhs_ref = W.get_diff(c["p"])
    got = rhs[0::2] + 1j * rhs[1::2]
    assert np.linalg.norm(got - rhs_ref) <= 1e-12 * np.linalg.norm(rhs_ref)
    # K1 = invM (S + P1/2 - P3/2): transport of a constant vanishes in the
    # interior (rows of K1 sum to the boundary terms)
    K1 = kmats[:n * n].reshape(n, n)
    assert np.isfinite(K1).all() and np.abs(K1).max() > 0


def test_invalid_input_is_rejected_with_a_message(lib):
    c = build_case(8, 2, 64, 0.0)
    desc, pot, keep = _desc_pot(c)
    bad = _lib.wdg_desc(desc.n_chi, desc.n_k_chi, desc.n_xi + 1, desc.nk, desc.l_chi, desc.l_xi)
    rc = lib.wdg_host_setup(C.byref(bad), C.byref(pot), None, None)
    assert rc == -1
    msg = lib.wdg_last_error(None).decode()
    assert "n_xi" in msg or "nk" in msg
    bad2 = _lib.wdg_desc(desc.n_chi, 9, desc.n_xi, desc.nk, desc.l_chi, desc.l_xi)
    assert lib.wdg_host_setup(C.byref(bad2), C.byref(pot), None, None) == -1
    assert lib.wdg_host_setup(None, None, None, None) == -1
    assert lib.wdg_theta_method(None) == -1


def test_full2d_poisson_entry_points_reject_null_arguments(lib):
    """wdg2_poisson_setup / wdg2_poisson_step validate their arguments before
    touching a device."""
    assert lib.wdg2_poisson_setup(None, None) == _lib.WDG_ERR_ARG
    info = _lib.wdg2_poisson_info()
    assert lib.wdg2_poisson_step(None, 1, None, None, 0, 1.0, None, None, None, C.byref(info)) == _lib.WDG_ERR_ARG
    assert "null" in lib.wdg_last_error(None).decode()


def test_time_apply_rejects_null_arguments(lib):
    """wdg_time_apply validates its arguments before touching a device."""
    ms = C.c_double(0.0)
    assert lib.wdg_time_apply(None, _lib.WDG_PREC_MIXED, 10, C.byref(ms)) == _lib.WDG_ERR_ARG
    assert "null" in lib.wdg_last_error(None).decode()
