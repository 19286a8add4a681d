"""Pins of the oracle's operator: restatements agree, invariants hold."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import wigner1d as W
from tests._cases import build_case, oracle_system, rand_c


@pytest.fixture(scope="module")
def small():
    c = build_case(12, 2, 32, 0.1)
    c["A"], c["G"], c["rhs"] = oracle_system(c)
    return c


def test_literal_vectorised_drift_equals_loop_form(small):
    GL = W.get_drift_literal(small["mat"], small["p"], small["E"][0, :, 0])
    G = small["G"]
    assert abs(GL - G).max() <= 1e-13 * abs(G).max()
    assert GL.nnz == G.nnz


def test_matrix_free_equals_assembled(small):
    op = W.MatrixFreeOperator.build(small["mat"], small["p"], small["E"][0, :, 0])
    u = rand_c(small["A"].shape[0])
    y = (small["A"] + small["G"]) @ u
    assert np.linalg.norm(op.apply(u) - y) <= 1e-13 * np.linalg.norm(y)


def test_transport_upwind_structure(small):
    p = small["p"]
    ne, n, nk = p["N_chi"], p["N_K_chi"], p["Nk"]
    A = small["A"].tocoo()
    e_r, j_r = A.row // (n * nk), A.row % nk
    e_c, j_c = A.col // (n * nk), A.col % nk
    assert np.all(j_r == j_c)                     # diagonal in k
    k = p["k"][j_r]
    up = e_c == e_r - 1
    down = e_c == e_r + 1
    assert np.all(k[up] > 0) and np.all(k[down] < 0)
    assert np.all(np.abs(e_c - e_r) <= 1)


def test_constant_state_with_matching_inflow_is_stationary():
    # f_L = f_R: then A_diff (1 (x) f) == rhs exactly (interior fluxes cancel,
    # boundary fluxes equal the inflow data).
    c = build_case(10, 3, 32, 0.0)
    p = c["p"]
    p = dict(p)
    p["verteilung_r"] = p["verteilung_l"]
    A, rhs = W.get_diff(p)
    u = np.kron(np.ones(p["N_chi"] * p["N_K_chi"]), p["verteilung_l"])
    assert np.linalg.norm(A @ u - rhs) <= 1e-12 * np.linalg.norm(rhs)


def test_constant_potential_leaves_only_cap():
    c = build_case(8, 2, 32, 0.0)
    E0 = np.zeros_like(c["E"][0, :, 0])
    G0 = W.get_drift(c["mat"], c["p"], E0)
    G1 = W.get_drift(c["mat"], c["p"], E0 + 0.37)
    assert abs(G1 - G0).max() <= 1e-12 * abs(G0).max()


def test_drift_hermitian_part_is_cap_absorption(small):
    # Real potential part: each element block is anti-Hermitian; the CAP adds
    # a positive semidefinite Hermitian part (Schur product of SPD invM with
    # the PSD quadrature Gram matrix, times a PSD tridiagonal CAP stencil).
    p = small["p"]
    blk = p["N_K_chi"] * p["Nk"]
    G = small["G"].toarray()
    B = G[:blk, :blk]
    H = (B + B.conj().T) / 2
    ev = np.linalg.eigvalsh(H)
    assert ev.min() >= -1e-12 * ev.max()
    # without the CAP the block is anti-Hermitian
    import oracle.wigner1d as WW
    saved = WW.get_cap
    try:
        WW.get_cap = lambda q, L, d: np.zeros_like(np.asarray(q, float))
        Gn = WW.get_drift(small["mat"], p, small["E"][0, :, 0]).toarray()[:blk, :blk]
    finally:
        WW.get_cap = saved
    assert np.linalg.norm(Gn + Gn.conj().T) <= 1e-12 * np.linalg.norm(Gn)


def test_xi_zero_edge_and_antisymmetry():
    c = build_case(8, 2, 32, 0.15)
    B, xq, xe = W.drift_samples(c["mat"], c["p"], c["E"][0, :, 0])
    mid = len(xe) // 2
    assert xe[mid] == 0.0 and np.all(B[:, mid] == 0)
    # the real part is odd in xi for quadrature points whose x +- xi/2 stay inside (0, L)
    L = c["p"]["L_chi"]
    inside = (xq[:, None] + np.abs(xe[None, :]) / 2 < L) & (xq[:, None] - np.abs(xe[None, :]) / 2 >= 0)
    R = B.real
    odd = R + R[:, ::-1]
    assert np.max(np.abs(odd[inside & inside[:, ::-1]])) < 1e-12


def test_direct_solve_and_reflection_symmetry():
    # symmetric device, equal Fermi levels: rho(e,a,k) = rho(N-1-e, n-1-a, -k)
    c = build_case(16, 2, 64, 0.0)
    A, G, rhs = oracle_system(c)
    K = (A + G).tocsc()
    x = spla.spsolve(K, rhs)
    assert np.linalg.norm(K @ x - rhs) <= 1e-12 * np.linalg.norm(rhs)
    p = c["p"]
    X = x.reshape(p["N_chi"], p["N_K_chi"], p["Nk"])
    assert np.linalg.norm(X - X[::-1, ::-1, ::-1]) <= 1e-12 * np.linalg.norm(X)


def test_solve_driver_outputs():
    c = build_case(8, 2, 32, 0.1)
    DG = W.solve_transport_dg_2d(c["mat"], c["E"], c["Vs"], c["efl"], c["efr"])
    nodes = 8 * 3
    assert DG["n"].shape == (nodes, 1) and DG["j"].shape == (nodes,)
    assert DG["RHO"].shape == (1, 1, nodes, 32)
    assert np.all(np.isfinite(DG["n"])) and np.all(DG["n"] > 0)
