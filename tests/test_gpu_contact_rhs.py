"""Contact-mode reservoirs and the inflow rhs built on the device
(wdg2_set_contact_rhs) against the host FP64 assembly of the same modes
(real_space_dg_amd.contact_modes, pinned by the reference's
contactModeReservoirFull2DTest.m assertions and the stored production sweep):
the box modes of the reference's own test setups (plain, masked contacts,
k_z-integrated), the periodic Fourier modes (Y-constant and Y-varying
contact potential), and the reference's production RTD grid at a stored
bias.  Bounds: 1e-12 relative (max norm) on each reservoir and on the rhs;
the stationary solve from the device rhs equals the one from the host rhs."""
import copy

import numpy as np
import pytest

from real_space_dg_amd import full2d as P
from real_space_dg_amd.solver import SolveOptions
from tests._sim_res import load
from tests.test_contact_modes import contact_mode_mat, masked_mat, periodic_contact_mat

pytestmark = pytest.mark.gpu


def kz_mat():
    m = contact_mode_mat()
    m["dg"]["params"].update({"full2D_contactModeIntegrateKz": True, "full2D_contactMode_Nkz": 101,
                              "full2D_contactModeFermiModel": "external"})
    return m


def sim_res_case():
    mat, cases = load()
    c = cases[1]
    return mat, c["V"], c["EfL"], c["EfR"]


CASES = {
    "box": lambda: (contact_mode_mat(), None, 0.1, 0.05),
    "box-masked": lambda: (masked_mat(), None, 0.1, 0.05),
    "box-kz": lambda: (kz_mat(), None, 0.1, 0.05),
    "periodic": lambda: (periodic_contact_mat(False), None, 0.25, 0.2),
    "periodic-y-profile": lambda: (periodic_contact_mat(True), None, 0.25, 0.2),
    "sim-res-0.2V": sim_res_case,
}


def rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


@pytest.mark.parametrize("name", list(CASES))
def test_device_reservoirs_and_rhs_match_host(name):
    mat, V, efl, efr = CASES[name]()
    V = mat["V"] if V is None else V
    print(f"\n[case] contact rhs {name}", flush=True)
    s = P.Full2DSetup(mat, V, efl, efr)
    host_rhs = s.rhs()
    with P.Full2DOperator(s) as op:
        res = op.upload_contact_rhs(want_reservoirs=True)
        for side in ("source", "drain"):
            ref = s.boundary[side]["rho"]
            e = rel(res[side], ref)
            print(f"  {side} reservoir rel err {e:.2e} (|rho| max {np.max(np.abs(ref)):.3e})", flush=True)
            assert e < 1e-12
        e = rel(op.device_rhs(), host_rhs)
        print(f"  rhs rel err {e:.2e}", flush=True)
        assert e < 1e-12
        assert abs(op.rhs_norm - np.linalg.norm(host_rhs)) <= 1e-12 * np.linalg.norm(host_rhs)


def test_solve_from_device_rhs_equals_host_rhs():
    mat = periodic_contact_mat(True)
    s = P.Full2DSetup(mat, mat["V"], 0.25, 0.2)
    opts = SolveOptions(tol=1e-11, restart=40, max_iter=20000)
    with P.Full2DOperator(s) as op:
        op.set_preconditioner("rowabs")
        op.upload_rhs()
        x_host, i1 = op.solve(opts)
        op.upload_contact_rhs()
        x_dev, i2 = op.solve(opts)
    assert i1.flag == 0 and i2.flag == 0
    assert np.linalg.norm(x_dev - x_host) <= 1e-9 * np.linalg.norm(x_host)
