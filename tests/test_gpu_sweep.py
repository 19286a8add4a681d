"""The batched bias sweep (wdg_solve_batch) against the oracle.

* Reduced sweep: 8 uneven instances drawn like the 64-instance sweep
  (devices.rtd_sweep: bias U[0, 0.4] V, barrier U[0.25, 0.35] eV, p = 3) on
  grids the oracle's direct sparse LU finishes in seconds (12-24 elements x
  32-96 k).  Every instance of one batched call is checked against its own
  direct solve (DG_code/solve_transport_DG_2D.m:101, oracle.wigner1d): the
  FP64 solution to 1e-9, density and current to 1e-9 (FP64) and 1e-7 (mixed).
* The 64-instance sweep at its real sizes (Nx in {128, 256, 384, 512}, Nk in
  {256, 512, 768, 1024}, seed 0; 55.7 M unknowns in one call): every instance
  converges in mixed precision, and its true residual, recomputed from the
  returned solution by the oracle's FP64 matrix-free operator
  (oracle.wigner1d.MatrixFreeOperator), is below the tolerance.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import wigner1d as W
from real_space_dg_amd.devices import rtd_sweep
from real_space_dg_amd.solver import SolveOptions, WignerOperator, solve_batch

pytestmark = pytest.mark.gpu
M0 = 9.109e-31


def _operator(case):
    mat, E, _, efl, efr = case
    op = WignerOperator.from_mat(mat)
    m = mat["me_x"][0, :, 0] * M0
    op.set_potential(mat["x"] * 1e-9, E[0, :, 0], efl, efr, m[0], m[-1], temp=mat["Temp"])
    return op


def _m_nodes(op):
    return np.full(op.n_chi * op.n_k_chi, 0.067 * M0)


@pytest.fixture(scope="module")
def reduced():
    cases = rtd_sweep(8, seed=0, n_elems=(12, 16, 24), nks=(32, 64, 96))
    print("\n[case] reduced sweep", [(c[0]["Nx"], c[0]["dg"]["params"]["N"]) for c in cases], flush=True)
    refs = []
    for mat, E, _, efl, efr in cases:
        m = mat["me_x"][0, :, 0] * M0
        p = W.init_params(mat, efl, efr, E[0, :, 0], m, 1.0)
        A, rhs = W.get_diff(p)
        K = (A + W.get_drift(mat, p, E[0, :, 0])).tocsc()
        x = spla.spsolve(K, rhs)
        mn = np.full(p["N_chi"] * p["N_K_chi"], 0.067 * M0)
        n, j, _ = W.observables(p, x, mn)
        refs.append((x, n, j))
    return cases, refs


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_reduced_sweep_matches_direct_solves(reduced, prec):
    cases, refs = reduced
    ops = [_operator(c) for c in cases]
    try:
        tol = 1e-12 if prec == "fp64" else 1e-10
        infos = solve_batch(ops, SolveOptions(tol=tol, precision=prec, restart=30, restart_max=64))
        bound = 1e-9 if prec == "fp64" else 1e-7
        for op, info, (x_ref, n_ref, j_ref) in zip(ops, infos, refs):
            assert info["flag"] == 0 and info["true_relres"] <= tol
            n, j, _ = op.observables(_m_nodes(op))
            en = np.max(np.abs(n - n_ref)) / np.max(np.abs(n_ref))
            ej = np.max(np.abs(j - j_ref)) / np.max(np.abs(j_ref))
            print(f"  {op.n_chi}x{op.n_k_chi}x{op.nk}: {info['iterations']} its, n err {en:.1e}, j err {ej:.1e}")
            assert en < bound and ej < bound
            if prec == "fp64":
                x = op.solution()
                assert np.linalg.norm(x - x_ref) <= 1e-9 * np.linalg.norm(x_ref)
    finally:
        for op in ops:
            op.close()


def test_full_sweep_converges_with_independent_residuals():
    cases = rtd_sweep(64, seed=0)
    print(f"\n[case] 64-instance sweep, {sum(c[0]['dg']['params']['N'] * (c[0]['Nx'] - 1) // 3 * 4 for c in cases)} "
          "unknowns", flush=True)
    ops = [_operator(c) for c in cases]
    try:
        tol = 1e-10
        infos = solve_batch(ops, SolveOptions(tol=tol, precision="mixed", restart=30, restart_max=64))
        assert all(i["flag"] == 0 for i in infos)
        print("  iterations: mean %.0f, max %d" % (np.mean([i["iterations"] for i in infos]),
                                                  max(i["iterations"] for i in infos)), flush=True)
        worst = 0.0
        for op, info, (mat, E, _, efl, efr) in zip(ops, infos, cases):
            m = mat["me_x"][0, :, 0] * M0
            p = W.init_params(mat, efl, efr, E[0, :, 0], m, 1.0)
            mf = W.MatrixFreeOperator.build(mat, p, E[0, :, 0])
            _, rhs = W.get_diff(p)
            x = op.solution()
            r = np.linalg.norm(rhs - mf.apply(x)) / np.linalg.norm(rhs)
            worst = max(worst, r)
            assert r <= 1.05 * tol, (op.n_chi, op.nk, r, info["true_relres"])
        print(f"  largest recomputed true residual {worst:.2e}", flush=True)
    finally:
        for op in ops:
            op.close()
