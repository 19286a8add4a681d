"""Device RK4 transient (wdg_transient_rk4) against the oracle's RK4."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import transient as TR
from real_space_dg_amd.solver import SolveOptions
from tests._cases import build_case, gpu_operator, oracle_system, rand_c

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case():
    c = build_case(8, 2, 64, 0.1)
    A, G, rhs = oracle_system(c)
    c["K"] = (A + G).tocsc()
    c["rhs"] = rhs
    c["lam"] = np.max(np.abs(np.linalg.eigvals(c["K"].toarray())))
    c["op"] = gpu_operator(c)
    c["m_nodes"] = np.full(c["p"]["N_chi"] * c["p"]["N_K_chi"], c["m"][0])
    yield c
    c["op"].close()


def test_rk4_from_perturbed_state_matches_oracle(case):
    a = 2e-10
    x_stat = spla.spsolve(case["K"], case["rhs"])
    rho0 = x_stat + 0.05 * np.linalg.norm(x_stat) / np.sqrt(x_stat.size) * rand_c(x_stat.size, 7)
    dt = 0.5 / (a * case["lam"])
    ref, n_ref, j_ref = TR.rk4(case["K"], case["rhs"], rho0, dt, 24, a_scale=a, b_scale=a, save_every=4,
                               p=case["p"], m_nodes=case["m_nodes"])
    rho, n_t, j_t = case["op"].transient_rk4(dt, 24, case["m_nodes"], save_every=4, a_scale=a, b_scale=a,
                                             rho0=rho0)
    assert n_t.shape == n_ref.shape == (case["m_nodes"].size, 6)
    assert np.linalg.norm(rho - ref) <= 1e-11 * np.linalg.norm(ref)
    assert np.max(np.abs(n_t - n_ref)) <= 1e-11 * np.max(np.abs(n_ref))
    assert np.max(np.abs(j_t - j_ref)) <= 1e-9 * np.max(np.abs(j_ref))
    # the perturbation decays towards the stationary state
    assert np.linalg.norm(rho - x_stat) < np.linalg.norm(rho0 - x_stat)


def test_rk4_from_device_solution_with_reference_scales(case):
    op = case["op"]
    x, info = op.solve(SolveOptions(tol=1e-13, precision="fp64"))
    assert info["flag"] == 0
    a, b = 2e-10, 1.0   # solve_transport_DG_2D_transient_SC_new.m:148 scales A only
    dt = 0.25 / (a * case["lam"])
    ref, n_ref, j_ref = TR.rk4(case["K"], case["rhs"], x, dt, 10, a_scale=a, b_scale=b, save_every=5,
                               p=case["p"], m_nodes=case["m_nodes"])
    rho, n_t, j_t = op.transient_rk4(dt, 10, case["m_nodes"], save_every=5, a_scale=a, b_scale=b)
    assert np.linalg.norm(rho - ref) <= 1e-10 * np.linalg.norm(ref)
    assert np.max(np.abs(n_t - n_ref)) <= 1e-10 * np.max(np.abs(n_ref))
    # the state stays on the device: observables of the last solution are those of rho
    n_last, _, _ = op.observables(case["m_nodes"])
    assert np.allclose(n_last, n_t[:, -1], rtol=1e-12, atol=0)


def test_rk4_rejects_bad_input(case):
    from real_space_dg_amd._lib import WdgError
    with pytest.raises(WdgError):
        case["op"].transient_rk4(-1.0, 3, case["m_nodes"])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_crank_nicolson_1d_matches_oracle(case, precision):
    """wdg_transient_cn against the oracle's directly solved trapezoidal steps."""
    a = 2e-10
    x_stat = spla.spsolve(case["K"], case["rhs"])
    rho0 = x_stat + 0.05 * np.linalg.norm(x_stat) / np.sqrt(x_stat.size) * rand_c(x_stat.size, 11)
    dt = 4.0 / (a * case["lam"])      # 8x the explicit stability limit of the RK4 step above
    ref = TR.crank_nicolson(case["K"], case["rhs"], rho0, dt, 6, a_scale=a, b_scale=a)
    tol = 1e-12 if precision == "fp64" else 1e-10
    rho, info = case["op"].transient_cn(dt, 6, a_scale=a, b_scale=a, tol=tol, precision=precision, rho0=rho0)
    assert info["steps"] == 6 and info["flag"] == 0 and info["max_relres"] <= tol
    assert np.linalg.norm(rho - ref) <= 1e-9 * np.linalg.norm(ref)
    n_last, _, _ = case["op"].observables(case["m_nodes"])
    n_ref, _, _ = TR.W.observables(case["p"], ref, case["m_nodes"])
    assert np.max(np.abs(n_last - n_ref)) <= 1e-8 * np.max(np.abs(n_ref))


def test_crank_nicolson_batch_mixes_1d_and_full2d_instances(case):
    """wdg_transient_cn_batch over a 1D and a full-2D operator (one option
    set for both): each instance ends where its own wdg_transient_cn run
    ends, and on the oracle's directly solved steps.  The shared step is the
    1D one; the full-2D operator's scale is ~1e21 smaller in these units,
    so that instance only accumulates its source (the batching, not the
    2D dynamics, is what this checks; those have their own tests)."""
    from oracle import full2d as F
    from real_space_dg_amd import full2d as P
    from real_space_dg_amd.solver import transient_cn_batch
    from test_full2d_setup import CONFIGS
    a = 2e-10
    x_stat = spla.spsolve(case["K"], case["rhs"])
    rho0_1d = x_stat + 0.05 * np.linalg.norm(x_stat) / np.sqrt(x_stat.size) * rand_c(x_stat.size, 13)
    dt = 4.0 / (a * case["lam"])
    m, EfL, EfR = CONFIGS["upwind-zero-inflow"]
    p, A2, rhs2, _ = F.solve_direct_full2d(m, m["V"], EfL, EfR)
    rho0_2d = np.zeros(p["nTotal"], complex)
    kw = dict(a_scale=a, b_scale=a, tol=1e-12, restart=40)
    with P.Full2DOperator(P.Full2DSetup(m, m["V"], EfL, EfR)) as op2:
        op2.upload_rhs()
        assert op2.set_preconditioner("rowabs") == "rowabs"
        r1, _ = case["op"].transient_cn(dt, 4, rho0=rho0_1d, **kw)
        r2, _ = op2.transient_cn(dt, 4, rho0=rho0_2d, **kw)
        outs, infos = transient_cn_batch([case["op"], op2], dt, 4, rho0=[rho0_1d, rho0_2d], **kw)
        print("CN batch 1D+2D:", infos, flush=True)
        assert all(i["flag"] == 0 and i["steps"] == 4 for i in infos)
        assert np.linalg.norm(outs[0] - r1) <= 1e-12 * np.linalg.norm(r1)
        assert np.linalg.norm(outs[1] - r2) <= 1e-12 * np.linalg.norm(r2)
        ref1 = TR.crank_nicolson(case["K"], case["rhs"], rho0_1d, dt, 4, a_scale=a, b_scale=a)
        assert np.linalg.norm(outs[0] - ref1) <= 1e-9 * np.linalg.norm(ref1)
        ref2 = TR.crank_nicolson(A2, rhs2, rho0_2d, dt, 4, a_scale=a, b_scale=a)
        assert np.linalg.norm(outs[1] - ref2) <= 1e-8 * np.linalg.norm(ref2)
