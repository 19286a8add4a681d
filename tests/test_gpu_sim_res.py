"""End-to-end regression against the reference's stored full-2D RTD sweep
(2026-10-28_08-47/Sim_Res.mat, three biases committed as fixtures): the
stationary solve at a stored self-consistent potential, with the run's own
settings (periodic Y, contact-mode reservoirs with neutrality Fermi levels,
upwind fluxes, FV-consistent potential, line-transverse preconditioner,
GMRES(40) to relres 5e-6), reproduces the stored density and current.  The
stored results carry the run's GMRES tolerance (relres ~5e-6), so agreement
is checked to 1e-3 in the density, not to 1e-10."""
import numpy as np
import pytest

from real_space_dg_amd import full2d as P
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
