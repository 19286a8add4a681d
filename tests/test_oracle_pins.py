"""Pins of the 1D oracle by mathematics that a plausible slip would fail.

* Theta[V] scale and spacings: for a uniform field V = -F x the potential
  difference is B(xi) = -F xi for every x, and the Wigner potential term must
  reduce to (q F / hbar) d f / d k on a smooth f(k).  The oracle's discrete
  operator (get_drift, DG_code/core/get_Drift_accel_trial_permutedShell.m
  :64-66,95-151 with the DG projection of :200-219) carries two known
  factors: the projection Q_drift invM o (W~^T W) of a state constant over an
  element's nodes gives (2/dxi) sum_b (M^-1 o M)_ab (the xi-scaled mass and the
  Hadamard product of get_Drift.m:141), and the FV sum over N_xi cells of
  width dxi = L/(N_xi - 1) integrates over N_xi/(N_xi - 1) L.  With both, the
  residual must converge at second order; dxi/4 -> dxi/2 (a factor 2) or
  dxi = L/N_xi (a 1/N error) fails.
* V_transient: x >= L takes min(EM) and x < 0 takes EM(1)
  (DG_code/core/V_transient.m:4-6), on a potential whose minimum is interior.
* DG_MSA_2D inflow (DG_code/core/verteilung.m:118-153): the 1001-point k_z sum
  equals (2/pi) [sqrt(pi)/2 F_{-1/2}(eta) / sqrt(a) + dkz g(0)/2] with the
  complete Fermi-Dirac integral F_{-1/2}(eta) = -Li_{1/2}(-e^eta) (mpmath)
  and the trapezoid end correction of the left-endpoint-inclusive sum.
"""
import mpmath
import numpy as np
import pytest

from oracle import wigner1d as W
from real_space_dg_amd.devices import rtd_1d

HBAR = 1.054571817e-34
QE = 1.602176634e-19
M0 = 9.109e-31


def _uniform_field_error(nk, F=1e6, l_xi=60e-9, length_nm=600.0, slip=None):
    mat, _, _, efl, efr = rtd_1d(4, 2, nk, bias_v=0.0, length_nm=length_nm, l_xi=l_xi)
    x = mat["x"] * 1e-9
    m = mat["me_x"][0, :, 0] * M0
    em = -F * x
    p = W.init_params(mat, efl, efr, em, m, 1.0)
    if slip is not None:
        p = slip(dict(p))
    G = W.get_drift(mat, p, em) - W.get_drift(mat, p, np.zeros_like(em))   # potential part (CAP cancels)
    k = p["k"]
    s = 8 * np.pi / l_xi                       # rho(xi) decays well inside [-L/2, L/2]
    f = np.exp(-(k / s) ** 2 / 2)
    df = -(k / s ** 2) * f
    n, ne = p["N_K_chi"], p["N_chi"]
    y = (G @ np.kron(np.ones(ne * n), f)).reshape(ne, n, nk)
    Mr = p["M"]
    proj = (2 / p["delta_xi"]) * (np.linalg.inv(Mr) * Mr).sum(axis=1)
    e = ne // 2                                # x +- L_xi/2 stays inside the device
    errs = []
    for a in range(n):
        expect = proj[a] * (QE * F / HBAR) * df * nk / (nk - 1)
        errs.append(np.linalg.norm(y[e, a] - expect) / np.linalg.norm(expect))
    return max(errs)


def test_uniform_field_potential_is_the_k_derivative():
    e1, e2 = _uniform_field_error(128), _uniform_field_error(256)
    print("uniform field residual", e1, e2)
    assert e2 < 4e-3
    assert 3.5 < e1 / e2 < 4.5          # second order


@pytest.mark.parametrize("slip", ["dxi_half", "dxi_edges"])
def test_uniform_field_check_catches_scale_slips(slip):
    """The same check on an oracle with a planted slip fails."""
    if slip == "dxi_half":
        def f(p):               # C = (dxi/2) tridiag(...) instead of dxi/4
            p["delta_xi"] = 2 * p["delta_xi"]
            p["M"] = p["M"] / 2.0      # keep the projection's xi-scaled mass unchanged
            return p
    else:
        def f(p):               # dxi = L/N instead of L/(N-1)
            n = p["N_xi"]
            p["delta_xi"] = p["delta_xi"] * (n - 1) / n
            p["M"] = p["M"] * n / (n - 1)
            return p
    e2 = _uniform_field_error(256, slip=f)
    assert e2 > 5e-3


def test_v_transient_clamps_to_min_beyond_the_device():
    L = 60e-9
    xg = np.linspace(0.0, L, 7)
    ug = np.array([0.3, 0.1, -0.2, 0.0, 0.05, 0.1, 0.05])   # min -0.2 is interior, EM(end) = 0.05
    x = np.array([-1e-9, 0.0, 0.5 * L, L - 1e-12, L, L + 5e-9])
    v = W.v_transient(x, xg, ug, None, L)
    assert v[0] == ug[0]                        # x < 0 -> EM(1)
    assert v[4] == -0.2 and v[5] == -0.2       # x >= L -> min(EM), not EM(end)
    assert np.isclose(v[2], np.interp(x[2], xg, ug))
    assert np.isclose(v[3], np.interp(x[3], xg, ug), rtol=1e-14)


def test_drift_samples_use_the_min_clamp():
    mat, E, _, efl, efr = rtd_1d(8, 2, 32, bias_v=-0.15)   # potential rises to the right: min(EM) = EM(1) != EM(end)
    em = E[0, :, 0]
    assert em.min() != em[-1]
    m = mat["me_x"][0, :, 0] * M0
    p = W.init_params(mat, efl, efr, em, m, 1.0)
    B, xq, xe = W.drift_samples(mat, p, em)
    xg = mat["x"] * 1e-9
    i = np.argmax(xq)                          # the quadrature point nearest x = L
    l = np.argmax(xe)                          # the largest xi edge: x + xi/2 beyond L
    assert xq[i] + xe[l] / 2 >= p["L_chi"]
    vm = np.interp(xq[i] - xe[l] / 2, xg, em) if xq[i] - xe[l] / 2 >= 0 else em[0]
    assert np.isclose(B[i, l].real, em.min() - vm, rtol=1e-12, atol=1e-15)


def test_dg_msa_2d_inflow_is_the_half_order_fermi_integral():
    mat, E, _, efl, efr = rtd_1d(8, 2, 64, bias_v=0.1, fermi_model="DG_MSA_2D")
    em = E[0, :, 0]
    m = mat["me_x"][0, :, 0] * M0
    p = W.init_params(mat, efl, efr, em, m, 1.0)
    fL, fR, k, dk = W.inflow_k(mat, p, efl, efr, em, m)
    hb = W.H_VT / 2 / np.pi
    q, kB, T = W.Q_VT, W.KB_VT, float(mat["Temp"])
    me_z = float(mat["me_z_ch"])
    a = hb ** 2 / (2 * W.M0_VT * me_z * kB * T)          # E_kz / kT = a kz^2
    kz_max = np.sqrt(2 * 2 * W.M0_VT * me_z * q * np.max(em + 0.3)) / hb
    dkz = kz_max / 1000
    for f, v0, ef, mm in ((fL, em[0], efl, m[0]), (fR, em[-1], efr, m[-1])):
        e_k = v0 + hb ** 2 * k ** 2 / 2 / mm / q
        eta = q * (ef - e_k) / (kB * T)
        fd = np.array([float(mpmath.re(-mpmath.polylog(0.5, -mpmath.exp(t)))) for t in eta])
        g0 = 1.0 / (1.0 + np.exp(-eta))
        closed = 2 / np.pi * (np.sqrt(np.pi) / 2 * fd / np.sqrt(a) + dkz * g0 / 2)
        assert np.max(np.abs(f - closed) / np.abs(closed)) < 1e-9
