"""Pins of the oracle's DG primitives and basis against closed forms."""
import math

import numpy as np
import pytest
from numpy.polynomial import legendre as L

from oracle import wigner1d as W


@pytest.mark.parametrize("n", [2, 3, 4, 5, 8, 10])
def test_gll_nodes_weights(n):
    x, w = W.jacobi_gl(0, 0, n - 1)
    N = n - 1
    assert x[0] == -1.0 and x[-1] == 1.0
    assert np.all(np.diff(x) > 0)
    # interior nodes are the roots of P'_N
    c = np.zeros(N + 1); c[N] = 1
    dP = L.legder(c)
    if n > 2:
        assert np.max(np.abs(L.legval(x[1:-1], dP))) < 1e-12
    # weights: 2/(N(N+1) P_N^2) and exactness up to degree 2N-1
    assert np.allclose(w, 2.0 / (N * (N + 1) * L.legval(x, c) ** 2), rtol=1e-13)
    for k in range(2 * N):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(np.sum(w * x ** k) - exact) < 1e-13


def test_gll_three_points():
    x, w = W.jacobi_gl(0, 0, 2)
    assert np.allclose(x, [-1, 0, 1], atol=1e-15)
    assert np.allclose(w, [1 / 3, 4 / 3, 1 / 3], atol=1e-15)


@pytest.mark.parametrize("n", [2, 3, 4, 6])
def test_vandermonde_orthonormal_and_derivative(n):
    r, _ = W.jacobi_gl(0, 0, n - 1)
    V = W.vandermonde1d(n - 1, r)
    xg, wg = L.leggauss(n + 1)
    Vg = W.vandermonde1d(n - 1, xg)
    assert np.allclose(Vg.T @ (wg[:, None] * Vg), np.eye(n), atol=1e-13)
    D = W.dmatrix1d(n - 1, r, V)
    assert np.max(np.abs(D @ np.ones(n))) < 1e-12
    for k in range(1, n):
        assert np.allclose(D @ r ** k, k * r ** (k - 1), atol=1e-12)
    # SBP: M D + (M D)^T = diag(-1, 0, ..., 0, 1)
    invV = np.linalg.inv(V)
    M = invV.T @ invV
    S = M @ D
    B = np.zeros((n, n)); B[0, 0] = -1; B[-1, -1] = 1
    assert np.allclose(S + S.T, B, atol=1e-12)
    # M is the exact mass matrix of the nodal basis
    lag = np.linalg.solve(V.T, Vg.T).T   # Lagrange basis at Gauss points
    assert np.allclose(lag.T @ (wg[:, None] * lag), M, atol=1e-13)


def test_plane_wave_basis_and_inflow_closed_form():
    from tests._cases import build_case
    c = build_case(8, 2, 64, 0.0)
    p = c["p"]
    k = p["k"]
    assert np.allclose(k, -k[::-1], rtol=0, atol=1e-6 * np.abs(k).max())
    assert np.min(np.abs(k)) > 0.4 * p["dk"]                       # even Nk: no k = 0
    assert np.allclose(np.diff(k), -2 * np.pi / p["L_xi"])
    assert np.allclose(np.abs(p["phi"]) ** 2, 1.0 / p["L_xi"])
    # phi\(cos transform of f) equals sqrt(L) dk/(2 pi) f for an even f
    # (the cosine transform is phi applied to that vector)
    hb = W.H_VT / 2 / np.pi
    E = c["E"][0, 0, 0] + hb ** 2 * k ** 2 / 2 / c["m"][0] / W.Q_VT
    fL = 2.0 / (1.0 + np.exp(W.Q_VT * (E - c["efl"]) / (W.KB_VT * 300.0)))
    expect = np.sqrt(p["L_xi"]) * p["dk"] / (2 * np.pi) * fL
    assert np.allclose(p["verteilung_l"], expect, rtol=1e-9, atol=1e-12 * expect.max())
