"""End-to-end GPU solves against the oracle's direct sparse solve."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import wigner1d as W
from real_space_dg_amd.solver import SolveOptions, solve_transport_DG_2D
from tests._cases import build_case, gpu_operator, oracle_system

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=[(64, 2, 128, 0.0), (32, 3, 128, 0.2)],
                ids=lambda c: "ne%d_p%d_nk%d_b%.1f" % c)
def solved(request):
    ne, deg, nk, bias = request.param
    print(f"\n[case] ne={ne} degree={deg} nk={nk} bias={bias}", flush=True)
    c = build_case(ne, deg, nk, bias)
    A, G, rhs = oracle_system(c)
    c["K"] = (A + G).tocsc()
    c["rhs"] = rhs
    c["x_ref"] = spla.spsolve(c["K"], rhs)
    c["op"] = gpu_operator(c)
    yield c
    c["op"].close()


def test_fp64_matches_direct(solved):
    x, info = solved["op"].solve(SolveOptions(tol=1e-13, restart=30, precision="fp64"))
    print(info)
    assert info["flag"] == 0 and info["true_relres"] <= 1e-13
    err = np.linalg.norm(x - solved["x_ref"]) / np.linalg.norm(solved["x_ref"])
    print("fp64 solution rel err vs direct", err)
    assert err < 1e-10
    # observables
    m_nodes = np.full(solved["p"]["N_chi"] * solved["p"]["N_K_chi"], 0.067 * 9.109e-31)
    n, j, _ = solved["op"].observables(m_nodes)
    n_ref, j_ref, _ = W.observables(solved["p"], solved["x_ref"], m_nodes)
    en = np.max(np.abs(n - n_ref)) / np.max(np.abs(n_ref))
    ej = np.max(np.abs(j - j_ref)) / max(np.max(np.abs(j_ref)), 1e-300)
    print("density err", en, "current err", ej)
    assert en < 1e-10
    assert ej < 1e-8


@pytest.mark.parametrize("method", ["fft-256", "fft-wide", "gemm"])
def test_mixed_reaches_tolerance(solved, method):
    """fft-wide runs the Gram-Schmidt multi-dot inside the Theta apply
    (theta_fused_dot), fft-256 and gemm as a kernel of its own."""
    solved["op"].set_theta_method(method.split("-")[0])
    if method.startswith("fft"):
        solved["op"].set_fft_block(method.split("-")[1])
    try:
        _mixed_reaches_tolerance(solved)
    finally:
        solved["op"].set_theta_method("auto")
        solved["op"].set_fft_block("auto")


def _mixed_reaches_tolerance(solved):
    x, info = solved["op"].solve(SolveOptions(tol=1e-10, restart=30, precision="mixed"))
    print(info)
    assert info["flag"] == 0 and info["true_relres"] <= 1e-10
    # refined to 1e-12, the mixed path lands on the FP64 oracle solution
    x, info = solved["op"].solve(SolveOptions(tol=1e-12, restart=30, precision="mixed"))
    assert info["flag"] == 0 and info["true_relres"] <= 1e-12
    err = np.linalg.norm(x - solved["x_ref"]) / np.linalg.norm(solved["x_ref"])
    print("mixed solution rel err vs direct", err)
    m_nodes = np.full(solved["p"]["N_chi"] * solved["p"]["N_K_chi"], 0.067 * 9.109e-31)
    n, j, _ = solved["op"].observables(m_nodes)
    n_ref, j_ref, _ = W.observables(solved["p"], solved["x_ref"], m_nodes)
    en = np.max(np.abs(n - n_ref)) / np.max(np.abs(n_ref))
    ej = np.max(np.abs(j - j_ref)) / max(np.max(np.abs(j_ref)), 1e-300)
    print("mixed density err", en, "current err", ej)
    assert en < 1e-9
    assert ej < 1e-7


@pytest.mark.parametrize("block", ["256", "wide"])
@pytest.mark.parametrize("basis", ["fp32", "bf16", "fp16"])
def test_mixed_krylov_basis_storage(solved, basis, block):
    """The inner basis may be stored narrower than FP32; the FP64 refinement
    still reaches the tolerance and lands on the oracle solution.  With the
    wide Theta block a 16-bit basis is dotted inside the apply."""
    solved["op"].set_fft_block(block)
    try:
        x, info = solved["op"].solve(SolveOptions(tol=1e-12, restart=30, precision="mixed", krylov_basis=basis))
    finally:
        solved["op"].set_fft_block("auto")
    print(basis, block, info)
    assert info["flag"] == 0 and info["true_relres"] <= 1e-12
    err = np.linalg.norm(x - solved["x_ref"]) / np.linalg.norm(solved["x_ref"])
    assert err < 1e-9


def test_warm_start_and_flag(solved):
    op = solved["op"]
    x, info = op.solve(SolveOptions(tol=1e-10))
    x2, info2 = op.solve(SolveOptions(tol=1e-10), x0=x)
    assert info2["iterations"] == 0 and info2["flag"] == 0
    _, info3 = op.solve(SolveOptions(tol=1e-10, max_iter=3))
    assert info3["flag"] == 1 and info3["status"] == "iterative-gmres-not-converged"


def test_solve_transport_dg_2d_mirror(solved):
    c = solved
    DG = solve_transport_DG_2D(c["mat"], c["E"], c["Vs"], c["efl"], c["efr"], SolveOptions(tol=1e-13))
    ref = W.solve_transport_dg_2d(c["mat"], c["E"], c["Vs"], c["efl"], c["efr"])
    en = np.max(np.abs(DG["n"] - ref["n"])) / np.max(np.abs(ref["n"]))
    er = np.max(np.abs(DG["RHO"] - ref["RHO"])) / np.max(np.abs(ref["RHO"]))
    print("mirror n err", en, "RHO err", er)
    assert en < 1e-10 and er < 1e-10
    assert DG["RHO"].shape == ref["RHO"].shape


def test_graph_replay_matches_eager(solved):
    """Captured cycle graphs replayed across restarts (with host syncs and
    eager work between replays) reproduce the eager launches exactly."""
    op = solved["op"]
    runs = []
    for use_graph in (False, True, True, False, True):
        for prec in ("fp64", "mixed"):
            x, info = op.solve(SolveOptions(tol=1e-10, precision=prec, use_graph=use_graph))
            runs.append((prec, use_graph, x, info["iterations"], info["true_relres"]))
    for prec in ("fp64", "mixed"):
        rs = [r for r in runs if r[0] == prec]
        x0 = rs[0][2]
        for _, g, x, its, rr in rs[1:]:
            assert its == rs[0][3]
            assert np.array_equal(x, x0), (prec, g, np.max(np.abs(x - x0)))


def test_residual_history_matches_the_solve_info():
    """wdg_solve_resvec: one value per Arnoldi step plus the initial
    residual; GMRES estimates never grow inside a restart cycle; the last
    value is the final estimate, below the tolerance, and each restart
    starts from the true residual the solver recomputes."""
    from tests._cases import build_case, gpu_operator
    c = build_case(8, 2, 64, 0.1)
    op = gpu_operator(c)
    try:
        for prec, restart in (("fp64", 12), ("mixed", 12)):
            x, info = op.solve(SolveOptions(tol=1e-10, restart=restart, precision=prec))
            r = info["resvec"]
            print(prec, info["iterations"], r[:3], r[-3:], flush=True)
            assert info["flag"] == 0 and r.size == info["iterations"] + 1
            assert r[0] == info["initial_true_relres"]
            assert np.isclose(r[-1], info["reported_relres"], rtol=1e-6, atol=0)
            assert np.all(np.isfinite(r)) and np.all(r >= 0)
            if prec == "fp64":
                # FP64 cycles run the full restart length: inside each, the
                # Givens estimates are non-increasing; the last one is below tol
                assert r[-1] <= 1e-10
                steps = np.diff(r[1:])            # steps[k]: step k+1 -> k+2
                inside = [k for k in range(steps.size) if (k + 1) % restart != 0]
                assert np.all(steps[inside] <= 1e-13 * r[0]), steps[inside].max()
            else:
                # inner FP32 cycles stop at a relative target; the estimates
                # still fall by the tolerance overall
                assert r[-1] <= 1e-8
        # a solve that converges before the first step records only the initial value
        x, info = op.solve(SolveOptions(tol=1e-10, precision="fp64"), x0=x)
        assert info["iterations"] == 0 and info["resvec"].size == 1
    finally:
        op.close()
