"""End-to-end regression against the reference's stored full-2D RTD sweep
(2026-10-28_08-47/Sim_Res.mat, three biases committed as fixtures): the
stationary solve at a stored self-consistent potential, with the run's own
settings (periodic Y, contact-mode reservoirs with neutrality Fermi levels,
upwind fluxes, FV-consistent potential, line-transverse preconditioner,
GMRES(40) to relres 5e-6), reproduces the stored density and current.  The
stored results carry the run's GMRES tolerance (relres ~5e-6), so agreement
is checked to 1e-3 in the density, not to 1e-10.  The stored potential is
also a fixed point of the transport + 2D Poisson map within the run's
self-consistency tolerances."""
import numpy as np
import pytest

from real_space_dg_amd import full2d as P
from real_space_dg_amd import poisson2d as PS
from tests._sim_res import load

pytestmark = pytest.mark.gpu

MAT, CASES = load()


@pytest.mark.parametrize("case", CASES, ids=lambda c: "bias%.1f" % c["bias"])
def test_stored_sweep_density_and_current(case):
    print(f"\n[case] Sim_Res bias {case['bias']} V", flush=True)
    r = P.solve_transport_DG_full2D(MAT, case["V"], case["EfL"], case["EfR"])
    info = r.info["solve"]
    print("solve", info, flush=True)
    assert info["flag"] == 0 and info["relres"] <= 5e-6 and info["preconditioner"] == "line-transverse"
    assert r.n.shape == case["n"].shape and r.j.shape == case["j"].shape
    en = np.max(np.abs(r.n - case["n"])) / np.max(np.abs(case["n"]))
    jm = abs(np.mean(r.j) - np.mean(case["j"])) / abs(np.mean(case["j"]))
    ej = np.max(np.abs(r.j - case["j"])) / np.max(np.abs(case["j"]))
    print(f"density {en:.3e}  mean current {jm:.3e}  current {ej:.3e}", flush=True)
    assert en < 1e-3
    if case["bias"] > 0:
        assert jm < 1e-3 and ej < 1e-2
    else:   # zero bias: a small residual current, matched in the mean
        assert jm < 1e-2
    # self-consistency: the stored potential is a fixed point of transport +
    # Poisson (RTD/solve2D_rtd_self_consistent.m:143-176, Newton-Raphson
    # Poisson as in the run) within the run's tolerances 5e-4 V and 8e-2
    M, types, periodic = PS.eval_poisson2d(MAT)
    n_pois, tinfo = PS.project_dg_density(np.real(r.nDG), r.setup, periodic)
    cand, pinfo = PS.poisson_step(M, types, n_pois, MAT, case["phi"], case["bias"], "newton-raphson")
    defect = float(np.max(np.abs(cand - case["phi"])))
    print(f"fixed point: defect {defect:.3e} V (run {case['last_potential_defect']:.3e}), Poisson relres "
          f"{pinfo['physicalRelresBefore']:.3e} (run {case['last_poisson_relres']:.3e}), charge transfer "
          f"{tinfo['relativeConservationError']:.1e}", flush=True)
    assert tinfo["relativeConservationError"] < 1e-12 and pinfo["converged"]
    assert defect <= 5e-4 and pinfo["physicalRelresBefore"] <= 8e-2
    # and it reproduces the run's own last-iteration diagnostics (the
    # transport states differ within the GMRES tolerance)
    assert abs(defect - case["last_potential_defect"]) <= 0.1 * case["last_potential_defect"]
    assert abs(pinfo["physicalRelresBefore"] - case["last_poisson_relres"]) <= 0.05 * case["last_poisson_relres"]


def test_self_consistent_loop_converges_at_the_stored_state(tmp_path):
    """The warm-started loop, started from the stored 0.2 V potential,
    stops by the run's criteria within three iterations and stays at the
    stored state; its result written as a Sim_Res.mat-layout file reads
    back with the stored file's sweep shapes and the loop's values."""
    from real_space_dg_amd.sim_res import load_sim_res, save_sim_res
    case = CASES[1]
    res = PS.solve2d_rtd_self_consistent(MAT, initial_phi=case["phi"], biases=[case["bias"]],
                                         poisson_method="newton-raphson", max_iterations=6)
    sw, hist = res["sweep"], res["histories"][0]
    for h in hist:
        print(h, flush=True)
    assert res["converged"] and sw["status"][0] == "fixed-point-converged" and len(hist) <= 3
    assert np.max(np.abs(sw["phi"][0] - case["phi"])) < 1e-4
    assert abs(np.mean(sw["current"][0]) - np.mean(case["j"])) <= 1e-3 * abs(np.mean(case["j"]))
    f = tmp_path / "Sim_Res.mat"
    save_sim_res(f, res, MAT, poisson_method="newton-raphson")
    d = load_sim_res(f)
    assert d["sweep"]["density"].shape == case["n"].shape + (1,)
    assert d["sweep"]["potential"].shape == case["V"].shape + (1,)
    assert np.array_equal(d["sweep"]["phi"][:, :, 0], sw["phi"][0])
    assert d["sweep"]["status"] == ["fixed-point-converged"]
    assert len(d["selfConsistency"]["history"][0]) == len(hist)


def test_self_consistent_transient_stays_at_the_stationary_state():
    """Crank-Nicolson steps with a 2D Poisson update every 10 steps, started
    from the stationary solution at the stored 0.2 V state: a self-consistent
    stationary state barely moves (potential updates at the fixed-point
    defect, density within 1e-3)."""
    from real_space_dg_amd.solver import SolveOptions
    case = CASES[1]
    setup = P.Full2DSetup(MAT, case["V"], case["EfL"], case["EfR"])
    with P.Full2DOperator(setup) as op:
        op.upload_rhs()
        op.set_line_transverse()
        _, info = op.solve(SolveOptions(tol=5e-6, restart=40, max_iter=6000))
        assert info.flag == 0
        n0, _, _ = op.observables()
        op.set_preconditioner("rowabs")
        phi, hist = PS.transient_cn_self_consistent(op, case["phi"], case["bias"], 1e-15, 20, poisson_every=10,
                                                    precision="mixed", tol=1e-8)
        for h in hist:
            print(h, flush=True)
        n1, _, _ = op.observables()
    assert len(hist) == 2 and all(h["flag"] == 0 and h["poissonConverged"] for h in hist)
    assert max(h["potentialUpdate"] for h in hist) < 1e-3
    assert np.max(np.abs(n1 - n0)) <= 1e-3 * np.max(np.abs(n0))
    assert np.max(np.abs(phi - case["phi"])) < 1e-3
