"""Pins of the oracle's Poisson step, density projection and Tsu-Esaki inflow."""
import numpy as np
import pytest
import scipy.integrate as si
import scipy.sparse.linalg as spla

from oracle import poisson as P
from oracle import wigner1d as W
from tests._cases import build_case, oracle_system


@pytest.mark.parametrize("n", [2, 3, 4, 5])
def test_projection_exact_for_polynomials(n):
    ne, L = 7, 30e-9
    r, _ = W.jacobi_gl(0, 0, n - 1)
    x = np.linspace(0, L, ne * (n - 1) + 1)
    T, widths = P.projection_matrix(ne, r, x, L)
    vert = np.linspace(0, L, ne + 1)
    xs = np.concatenate([vert[e] + (r + 1) / 2 * (vert[e + 1] - vert[e]) for e in range(ne)])
    # a global polynomial of degree n-1 is represented exactly on every element
    c = np.random.default_rng(0).standard_normal(n)
    f = np.polynomial.polynomial.polyval(xs / L, c)
    got = T @ f
    edges = np.concatenate([[x[0]], (x[:-1] + x[1:]) / 2, [x[-1]]])
    F = np.polynomial.polynomial.polyint(c)
    exact = (np.polynomial.polynomial.polyval(edges[1:] / L, F) - np.polynomial.polynomial.polyval(edges[:-1] / L, F)) * L
    assert np.allclose(got, exact, rtol=1e-12, atol=1e-12 * np.abs(exact).max())
    # conservation of the total charge for a discontinuous field
    g = np.random.default_rng(1).standard_normal(ne * n)
    M = np.linalg.inv(W.vandermonde1d(n - 1, r) @ W.vandermonde1d(n - 1, r).T)
    total = sum(np.sum(M @ g[e * n:(e + 1) * n]) * (vert[e + 1] - vert[e]) / 2 for e in range(ne))
    assert abs(np.sum(T @ g) - total) <= 1e-12 * np.abs(T @ g).sum()
    assert np.isclose(widths.sum(), L)


def test_poisson_matrix_exact_for_quadratic_and_interfaces():
    nx, dx = 41, 0.5
    x = np.arange(nx) * dx * 1e-9
    M = P.poisson_matrix(np.full(nx, 12.9), dx)
    phi = 3e15 * x ** 2 - 2e6 * x            # phi'' = 6e15 V/m^2
    r = (M @ phi)[1:-1]
    assert np.allclose(r, P.EPS0 * 12.9 * 6e15, rtol=1e-9)
    # dielectric step half-way between nodes 19 and 20: the exact solution of
    # (eps phi')' = 0 is piecewise linear with eps*phi' continuous; its nodal
    # values satisfy the discrete system exactly (harmonic-mean face)
    er = np.where(np.arange(nx) <= 19, 12.9, 6.45)
    xm = (x[19] + x[20]) / 2
    s_l = 1e7
    s_r = s_l * 12.9 / 6.45
    phi2 = np.where(x < xm, s_l * x, s_l * xm + s_r * (x - xm))
    M2 = P.poisson_matrix(er, dx)
    res = (M2 @ phi2)[1:-1]
    flux = P.EPS0 * 12.9 * s_l / (dx * 1e-9)
    assert np.max(np.abs(res)) < 1e-9 * flux
    # an arithmetic-mean face would leave a residual at the interface nodes
    phi3 = np.where(x < x[20], s_l * x, s_l * x[20] + s_r * (x - x[20]))
    assert np.max(np.abs((M2 @ phi3)[1:-1])) > 1e-2 * flux


def test_newton_step_converges_quadratically_on_boltzmann_density():
    nx, dx, T = 61, 0.5, 300.0
    x = np.arange(nx) * dx * 1e-9
    M = P.poisson_matrix(np.full(nx, 12.9), dx)
    nd = np.where((np.arange(nx) < 20) | (np.arange(nx) > 40), 1e24, 1e21)
    n0 = 1e24
    bias = 0.05
    ref_phi = bias * np.arange(nx) / (nx - 1)
    def density(phi):          # Boltzmann electrons around a fixed quasi-Fermi level
        return n0 * np.exp(P.Q * (phi - ref_phi) / (P.KB * T))
    phi = ref_phi.copy()
    steps = []
    for _ in range(8):
        phi, info = P.poisson_step(M, density(phi), nd, phi, bias, T)
        steps.append(info["unmixed_step"])
    assert steps[-1] < 1e-12 and steps[3] < 1e-3 * steps[1]
    res = (M @ phi - P.Q * (density(phi) - nd))[1:-1]
    assert np.max(np.abs(res)) < 1e-6 * P.Q * 1e24


def test_tsu_esaki_flat_band_density_matches_3d_fermi_integral():
    # flat band, no barrier, zero bias: the uniform Wigner state's density is
    # the 3D Fermi-Dirac density 2/(2pi)^3 int FD d^3k of the contacts
    c = build_case(24, 2, 128, 0.0, fermi_model="TSU_ESAKI", barrier_ev=0.0, l_xi=60e-9)
    A, G, rhs = oracle_system(c)
    x = spla.spsolve((A + G).tocsc(), rhs)
    m = c["m"]
    n_nodes = np.full(24 * 3, m[0])
    n, j, _ = W.observables(c["p"], x, n_nodes)
    hb, q, kB, T = W.H_VT / 2 / np.pi, W.Q_VT, W.KB_VT, 300.0
    ef = c["efl"]
    integrand = lambda k: k * k / (1 + np.exp((hb ** 2 * k * k / 2 / m[0] - q * ef) / (kB * T)))
    n3d = si.quad(integrand, 0, 5e9)[0] / np.pi ** 2
    mid = n[len(n) // 2]
    assert abs(mid / n3d - 1) < 0.05, (mid, n3d)
    assert np.max(np.abs(j)) < 1e-3 * q * n3d * 1e5
