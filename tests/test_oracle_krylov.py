"""Pins of the oracle's line preconditioner and GMRES."""
import numpy as np
import scipy.sparse.linalg as spla

from oracle import krylov
from oracle import wigner1d as W
from tests._cases import build_case, line_operator, oracle_system, rand_c


def test_line_preconditioner_inverts_line_operator():
    c = build_case(10, 3, 32, 0.2)
    A, G, rhs = oracle_system(c)
    p = c["p"]
    op = W.MatrixFreeOperator.build(c["mat"], p, c["E"][0, :, 0])
    P = line_operator(A, G, p["N_chi"], p["N_K_chi"], p["Nk"])
    r = rand_c(P.shape[0], 7)
    z = krylov.LinePreconditioner(op).apply(r)
    assert np.linalg.norm(P @ z - r) <= 1e-11 * np.linalg.norm(r)


def test_gmres_matches_direct():
    c = build_case(16, 2, 64, 0.1)
    A, G, rhs = oracle_system(c)
    K = (A + G).tocsc()
    x_ref = spla.spsolve(K, rhs)
    op = W.MatrixFreeOperator.build(c["mat"], c["p"], c["E"][0, :, 0])
    Bbar = op.edge_coefficients()
    pre = krylov.LinePreconditioner(op, Bbar)
    x, info = krylov.gmres(lambda u: op.apply(u, Bbar), pre.apply, rhs, restart=30, tol=1e-12)
    assert info["flag"] == 0
    assert np.linalg.norm(x - x_ref) <= 1e-9 * np.linalg.norm(x_ref)
