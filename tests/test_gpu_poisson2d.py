"""The full-2D Poisson update on the device (wdg2_poisson_setup /
wdg2_poisson_step, poisson2d.DevicePoisson2D) against the host FP64 path of
the same loop step: the zero-separation density of the device solution
(get_Observables_full2D.m:24-60), its conservative projection onto the
Poisson dual volumes (RTD/project_DG_density_RTD.m) and one direct or
Newton-Raphson update (RTD/solve_Poisson2D_rtd.m:85-130) by SciPy's sparse
LU.  The self-consistent FET slab on a reduced grid (8 x 4 DG elements p = 1,
8 x 8 relative cells, n+ contacts, contact-mode reservoirs), with Dirichlet
X contacts and, separately, periodic Y (the reference's production setting),
one transport instance and the sum of two (transverse modes).  Bounds: the
banded LU and SuperLU differ by rounding only, 1e-11 relative on the
potential and the step, 1e-12 on the projected density."""
import copy

import numpy as np
import pytest

from real_space_dg_amd import full2d as P
from real_space_dg_amd import poisson2d as PS
from real_space_dg_amd.devices import fet_slab_self_consistent
from real_space_dg_amd.solver import SolveOptions

pytestmark = pytest.mark.gpu

EFL, EFR, BIAS = 0.1, 0.05, 0.05


@pytest.fixture(scope="module", params=["dirichlet", "periodic"])
def slab(request):
    mat = fet_slab_self_consistent(n_el_x=8, n_el_y=4, n_rho=8)
    if request.param == "periodic":
        mat = copy.deepcopy(mat)
        mat["dg"]["params"]["full2D_Y_boundary"] = "periodic"
    print(f"\n[case] FET slab 8x4x8x8 self-consistent, Y {request.param}", flush=True)
    x = np.asarray(mat["x"], float)
    # a perturbed potential so that the update is not trivial
    phi0 = np.repeat((BIAS * x / x[-1])[:, None], mat["Ny"], axis=1)
    phi0 = phi0 + 0.01 * np.sin(np.pi * x / x[-1])[:, None] * np.cos(np.linspace(0, np.pi, mat["Ny"]))[None, :]
    if request.param == "periodic":
        phi0[:, -1] = phi0[:, 0]
    Vc = np.asarray(mat["Vc"], float)
    ops = []
    for off in (0.0, 0.005):
        s = P.Full2DSetup(mat, Vc - phi0 + off, EFL, EFR)
        op = P.Full2DOperator(s)
        op.upload_contact_rhs()
        op.set_preconditioner("rowabs")
        _, info = op.solve(SolveOptions(tol=1e-11, precision="fp64", restart=40), return_x=False)
        assert info.flag == 0, (info.flag, info.true_relres)
        ops.append(op)
    yield dict(mat=mat, phi0=phi0, ops=ops)
    for op in ops:
        op.close()


def _host(slab, ops, method, alpha=1.0):
    mat = slab["mat"]
    M, types, periodic = PS.eval_poisson2d(mat)
    nDG = sum(np.real(op.observables()[0]) for op in ops)
    n, _ = PS.project_dg_density(nDG, ops[0].setup, periodic)
    phi, info = PS.poisson_step(M, types, n, mat, slab["phi0"], BIAS, method, alpha=alpha)
    return phi, info, n


@pytest.mark.parametrize("method", ["direct", "newton-raphson"])
@pytest.mark.parametrize("ninst", [1, 2])
def test_device_step_matches_host(slab, method, ninst):
    ops = slab["ops"][:ninst]
    dev = PS.DevicePoisson2D(ops[0])
    phi_d, info_d = dev.step(ops, slab["phi0"], BIAS, method)
    phi_h, info_h, n_h = _host(slab, ops, method)
    scale = max(1.0, np.max(np.abs(phi_h)))
    e_phi = np.max(np.abs(phi_d - phi_h)) / scale
    e_n = np.max(np.abs(info_d["density"] - n_h)) / np.max(np.abs(n_h))
    e_step = abs(info_d["unmixedStep"] - info_h["unmixedStep"]) / info_h["unmixedStep"]
    print(f"{method} x{ninst}: phi {e_phi:.2e} density {e_n:.2e} step {e_step:.2e} "
          f"(step {info_h['unmixedStep']:.3e} V, relres {info_h['physicalRelresBefore']:.3e}, "
          f"linear {info_d['linearRelres']:.1e})")
    assert info_h["unmixedStep"] > 1e-6   # a real update
    assert e_phi < 1e-11
    assert e_n < 1e-12
    assert e_step < 1e-11
    assert np.max(np.abs(info_d["candidateUnmixed"] - info_h["candidateUnmixed"])) / scale < 1e-11
    assert abs(info_d["physicalRelresBefore"] - info_h["physicalRelresBefore"]) <= \
        1e-10 * info_h["physicalRelresBefore"]
    assert abs(info_d["physicalRelresAfter"] - info_h["physicalRelresAfter"]) <= \
        1e-8 * info_h["physicalRelresBefore"]
    assert info_d["converged"] and info_d["linearRelres"] < 1e-12


def test_mixed_update_and_contacts(slab):
    """alpha < 1 mixes the update on free nodes only; the Dirichlet contacts
    hold 0 and the bias exactly."""
    op = slab["ops"][0]
    dev = PS.DevicePoisson2D(op)
    phi_d, info_d = dev.step([op], slab["phi0"], BIAS, "newton-raphson", alpha=0.3)
    phi_h, info_h, _ = _host(slab, [op], "newton-raphson", alpha=0.3)
    assert np.max(np.abs(phi_d - phi_h)) < 1e-11 * max(1.0, np.max(np.abs(phi_h)))
    assert np.all(phi_d[0, :] == 0.0) and np.all(phi_d[-1, :] == BIAS)
    if slab["mat"]["dg"]["params"].get("full2D_Y_boundary") == "periodic":
        assert np.array_equal(phi_d[:, -1], phi_d[:, 0])


def test_rejects_bad_input(slab):
    op = slab["ops"][0]
    dev = PS.DevicePoisson2D(op)
    with pytest.raises(ValueError):
        dev.step([op], slab["phi0"], BIAS, "bisection")
    bad = slab["phi0"].copy()
    bad[3, 2] = np.nan
    with pytest.raises(RuntimeError):
        dev.step([op], bad, BIAS, "direct")
