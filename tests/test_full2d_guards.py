"""Host-side guards of the full-2D mirror (no GPU): the reference's
DG:Full2D:InvalidInitialRho (DG_code/solve_transport_DG_full2D.m:508-525),
DG:Full2D:GmresMemoryLimit (:183-203) and the length checks of every
Full2DOperator method that hands a host vector to the library."""
import ctypes as C

import numpy as np
import pytest

from real_space_dg_amd import _lib
from real_space_dg_amd.full2d import Full2DOperator, gmres_basis_guard, validate_initial_rho


def test_initial_rho_must_be_a_finite_vector_of_n_total():
    n = 24
    ok = validate_initial_rho(np.arange(n, dtype=float), n)
    assert ok.dtype == np.complex128 and ok.shape == (n,)
    assert validate_initial_rho(np.ones((n, 1), complex), n).shape == (n,)      # a column vector is a vector
    for bad in (np.ones(n - 1), np.ones((2, 3, 2, 2)), np.ones((4, 6)), np.array(["a"] * n)):
        with pytest.raises(_lib.WdgError) as ei:
            validate_initial_rho(bad, n)
        assert ei.value.identifier == "DG:Full2D:InvalidInitialRho"
        assert ei.value.code == _lib.WDG_ERR_INITIAL_RHO
    x = np.ones(n, complex)
    x[3] = np.nan
    with pytest.raises(_lib.WdgError, match="finite"):
        validate_initial_rho(x, n)


def test_gmres_basis_guard_follows_the_reference():
    n = 3_136_320                      # the reference's production RTD grid
    est, lim = gmres_basis_guard({}, n, 40, "line-transverse")
    assert est == 16 * n * 41 and lim == 9 * 1024 ** 3
    with pytest.raises(_lib.WdgError) as ei:
        gmres_basis_guard({}, 20 * n, 40, "line-transverse")
    assert ei.value.identifier == "DG:Full2D:GmresMemoryLimit"
    gmres_basis_guard({}, 20 * n, 40, "jacobi")          # only the line-transverse path is guarded
    gmres_basis_guard({"full2D_gmresMaxBasisBytes": 1e12}, 20 * n, 40, "line-transverse")
    with pytest.raises(_lib.WdgError):
        gmres_basis_guard({"full2D_gmresMaxBasisBytes": 1e3}, 100, 5, "line-transverse")


def test_operator_methods_reject_wrong_lengths_before_the_library():
    op = object.__new__(Full2DOperator)      # no device context: the checks come first
    op.n = 12
    short = np.zeros(5, complex)
    with pytest.raises(ValueError):
        op.apply(short)
    with pytest.raises(ValueError):
        op.precond(short)
    with pytest.raises(ValueError):
        op.observables(rho=short)
    with pytest.raises(_lib.WdgError, match="InvalidInitialRho"):
        op.solve(x0=short)


def test_memory_limit_entry_point_rejects_bad_input():
    lib = _lib.load()
    assert lib.wdg_set_memory_limits(None, 0.0, 1.0) == _lib.WDG_ERR_ARG
