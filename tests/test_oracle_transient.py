"""Pins of the oracle's RK4 transient against the exact linear-ODE solution."""
import numpy as np
import scipy.linalg as sla
import scipy.sparse.linalg as spla

from oracle import transient as TR
from tests._cases import build_case, oracle_system


def _small():
    c = build_case(4, 2, 32, 0.1)
    A, G, rhs = oracle_system(c)
    K = (A + G).tocsc()
    return c, K, rhs


def test_rk4_matches_matrix_exponential_with_fourth_order_error():
    c, K, rhs = _small()
    a = 2e-10
    Kd = K.toarray() * a
    lam = np.max(np.abs(np.linalg.eigvals(Kd)))
    rho_inf = np.linalg.solve(Kd, rhs)   # fixed point of rhs - a K rho
    rng = np.random.default_rng(0)
    rho0 = rho_inf + 0.1 * np.linalg.norm(rho_inf) / np.sqrt(rho_inf.size) * (
        rng.standard_normal(rho_inf.size) + 1j * rng.standard_normal(rho_inf.size))
    T = 1.0 / lam * 8
    exact = rho_inf + sla.expm(-Kd * T) @ (rho0 - rho_inf)
    errs = []
    for nsteps in (40, 80):
        dt = T / nsteps
        rho, _, _ = TR.rk4(K, rhs, rho0, dt, nsteps, a_scale=a)
        errs.append(np.linalg.norm(rho - exact) / np.linalg.norm(exact))
    assert errs[1] < 1e-6
    ratio = errs[0] / errs[1]
    assert 12 < ratio < 20, (errs, ratio)


def test_stationary_state_is_a_fixed_point():
    c, K, rhs = _small()
    a, bsc = 2e-10, 3.0
    rho_inf = spla.spsolve(K, rhs) * bsc / a
    lam = np.max(np.abs(np.linalg.eigvals(K.toarray() * a)))
    rho, n_t, j_t = TR.rk4(K, rhs, rho_inf, 0.5 / lam, 5, a_scale=a, b_scale=bsc, p=c["p"],
                           m_nodes=np.full(c["p"]["N_chi"] * c["p"]["N_K_chi"], c["m"][0]))
    assert np.linalg.norm(rho - rho_inf) <= 1e-9 * np.linalg.norm(rho_inf)
    assert n_t.shape == (c["p"]["N_chi"] * c["p"]["N_K_chi"], 5)
    assert np.allclose(n_t[:, 0], n_t[:, -1], rtol=1e-9)


def test_crank_nicolson_second_order_and_amplification_factor():
    c, K, rhs = _small()
    a = 2e-10
    Kd = K.toarray() * a
    lam = np.max(np.abs(np.linalg.eigvals(Kd)))
    rho_inf = np.linalg.solve(Kd, rhs)
    rng = np.random.default_rng(1)
    rho0 = rho_inf + 0.1 * np.linalg.norm(rho_inf) / np.sqrt(rho_inf.size) * (
        rng.standard_normal(rho_inf.size) + 1j * rng.standard_normal(rho_inf.size))
    T = 8.0 / lam
    exact = rho_inf + sla.expm(-Kd * T) @ (rho0 - rho_inf)
    errs = []
    for nsteps in (200, 400):
        rho = TR.crank_nicolson(K, rhs, rho0, T / nsteps, nsteps, a_scale=a)
        errs.append(np.linalg.norm(rho - exact) / np.linalg.norm(exact))
    ratio = errs[0] / errs[1]
    assert 3.5 < ratio < 4.5, (errs, ratio)
    # stationary state is a fixed point
    rho = TR.crank_nicolson(K, rhs, rho_inf, 3.0 / lam, 4, a_scale=a)
    assert np.linalg.norm(rho - rho_inf) <= 1e-10 * np.linalg.norm(rho_inf)
    # diagonal A: per-mode amplification (1 - z/2)/(1 + z/2)
    import scipy.sparse as sp
    d = np.array([1.0, 2.0 + 1j, 0.3j, 5.0])
    z = 0.7 * d
    r = TR.crank_nicolson(sp.diags(d), np.zeros(4), np.ones(4), 0.7, 3)
    assert np.allclose(r, ((1 - z / 2) / (1 + z / 2)) ** 3, rtol=1e-14, atol=0)


def test_matrix_free_crank_nicolson_equals_direct():
    c, K, rhs = _small()
    a = 2e-10
    Kd = K * a
    lam = abs(Kd).sum(axis=1).max()
    rng = np.random.default_rng(2)
    rho0 = rng.standard_normal(K.shape[0]) + 1j * rng.standard_normal(K.shape[0])
    dt = 5.0 / lam
    ref = TR.crank_nicolson(K, rhs, rho0, dt, 3, a_scale=a)
    mf = TR.crank_nicolson_mf(lambda v: K @ v, rhs, rho0, dt, 3, a_scale=a, row_scale=abs(K).sum(axis=1).A1)
    assert np.linalg.norm(mf - ref) <= 1e-11 * np.linalg.norm(ref)
