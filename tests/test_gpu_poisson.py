"""Device Poisson step (wdg_poisson_step) against the oracle, and the Gummel loop."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import poisson as P
from oracle import wigner1d as W
from real_space_dg_amd.devices import rtd_self_consistent
from real_space_dg_amd.solver import SolveOptions, WignerOperator, _node_values

M0 = 9.109e-31
pytestmark = pytest.mark.gpu


def _state(n_elem=24, degree=3, nk=64, bias=0.1):
    mat = rtd_self_consistent(n_elem=n_elem, degree=degree, nk=nk, n_bias=2, bias_max=bias)
    x_nm = mat["x"]
    ramp = np.clip((x_nm - mat["Lc1"]) / (x_nm[-1] - mat["Lc1"] - mat["Lc2"]), 0, 1)
    phi0 = bias * ramp + 0.01 * np.sin(np.pi * x_nm / x_nm[-1])
    V = mat["Vc"] - phi0
    m = mat["me_x"][0, :, 0] * M0
    efl = mat["Vc"][0] + 0.05
    efr = mat["Vc"][-1] + 0.05 - bias
    return mat, phi0, V, m, efl, efr


@pytest.mark.parametrize("method", ["newton-raphson", "direct"])
def test_poisson_step_matches_oracle(method):
    bias = 0.1
    mat, phi0, V, m, efl, efr = _state(bias=bias)
    p = W.init_params(mat, efl, efr, V, m, 1.0)
    A, rhs = W.get_diff(p)
    G = W.get_drift(mat, p, V)
    rho = spla.spsolve((A + G).tocsc(), rhs)
    n_chi, n = p["N_chi"], p["N_K_chi"]
    m_nodes = _node_values(n_chi, n, m)
    n_dg, _, _ = W.observables(p, rho, m_nodes)
    T, widths = P.projection_matrix(n_chi, p["r"], mat["x"] * 1e-9, p["L_chi"])
    n_cv = P.project_density(T, widths, n_dg)
    M = P.poisson_matrix(mat["eps"], mat["dx"])
    ref_phi, ref_info = P.poisson_step(M, n_cv, mat["Nd"], phi0, bias, mat["Temp"], method=method)

    op = WignerOperator.from_mat(mat)
    op.set_potential(mat["x"] * 1e-9, V, efl, efr, m[0], m[-1], temp=mat["Temp"], fermi_model="TSU_ESAKI")
    _, info = op.solve(SolveOptions(tol=1e-13, precision="fp64"), return_x=False)
    assert info["flag"] == 0
    op.poisson_setup(mat["x"] * 1e-9, mat["eps"], mat["Nd"], temp=mat["Temp"], method=method)
    phi, ncv_gpu, pinfo = op.poisson_step(phi0, bias)
    assert np.max(np.abs(ncv_gpu - n_cv)) <= 1e-9 * np.max(np.abs(n_cv))
    assert np.max(np.abs(phi - ref_phi)) <= 1e-9 * max(ref_info["unmixed_step"], 1e-3)
    assert phi[0] == 0.0 and phi[-1] == bias
    assert abs(pinfo["physical_relres_before"] - ref_info["physical_relres_before"]) <= \
        1e-8 * ref_info["physical_relres_before"]
    assert abs(pinfo["unmixed_step"] - ref_info["unmixed_step"]) <= 1e-8 * ref_info["unmixed_step"]
    # the projection conserves the DG charge
    assert abs(pinfo["charge_integral"] - pinfo["dg_integral"]) <= 1e-12 * abs(pinfo["dg_integral"])
    op.close()


def test_gummel_converges_and_matches_individual_newton_iterations():
    from real_space_dg_amd.gummel import GummelOptions, solve_rtd_self_consistent
    mat = rtd_self_consistent(n_elem=24, degree=3, nk=64, n_bias=3, bias_max=0.2)
    opt = GummelOptions(max_iterations=60, potential_tolerance=1e-7, acceleration="anderson", mixing=0.5)
    opt.transport = SolveOptions(tol=1e-11, precision="fp64", restart=30)
    res = solve_rtd_self_consistent(mat, opt)
    assert all(res["converged"]), res["status"]
    # zero bias: no current; current grows with bias
    j = res["current"]
    assert abs(j[0]) < 1e-6 * abs(j[-1])
    assert j[1] > 0 and j[2] > j[1]
    # the converged potential is a fixed point of the oracle's Poisson step
    # on the oracle's own transport density
    for ib in (1, 2):
        bias = res["bias"][ib]
        phi = res["phi"][ib]
        V = mat["Vc"] - phi
        m = mat["me_x"][0, :, 0] * M0
        p = W.init_params(mat, mat["Vc"][0] + 0.05, mat["Vc"][-1] + 0.05 - bias, V, m, 1.0)
        A, rhs = W.get_diff(p)
        G = W.get_drift(mat, p, V)
        rho = spla.spsolve((A + G).tocsc(), rhs)
        n_dg, _, _ = W.observables(p, rho, _node_values(p["N_chi"], p["N_K_chi"], m))
        T, widths = P.projection_matrix(p["N_chi"], p["r"], mat["x"] * 1e-9, p["L_chi"])
        M = P.poisson_matrix(mat["eps"], mat["dx"])
        cand, _ = P.poisson_step(M, P.project_density(T, widths, n_dg), mat["Nd"], phi, bias, mat["Temp"])
        assert np.max(np.abs(cand - phi)) < 1e-5
