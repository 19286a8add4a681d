"""Full-2D operator on the GPU against the FP64 oracle (oracle/full2d.py):
apply, diagonal, row sums, preconditioned GMRES solve and observables."""
import numpy as np
import pytest

from oracle import full2d as F
from real_space_dg_amd import full2d as P
from real_space_dg_amd.solver import SolveOptions

from test_full2d_setup import CONFIGS
from test_oracle_full2d import small_mat

pytestmark = pytest.mark.gpu


def tiled_mat(nrx=64, nry=10, nk=2, flux="rusanov", ybnd="specular"):
    """Several rho_y tiles per element and a rho_x row that is not a wave multiple."""
    m = small_mat()
    m["dg"]["params"].update({"N_K_X": nk, "N_K_Y": nk, "N_rho_x": nrx, "N_rho_y": nry, "full2D_X_flux": flux,
                              "full2D_Y_flux": flux, "full2D_Y_boundary": ybnd})
    return m


ALL = dict(CONFIGS)
ALL["tiled-64x10"] = (tiled_mat(), 0.2, 0.1)
ALL["tiled-80x7-upwind-p2"] = (tiled_mat(80, 7, 3, "upwind", "zero-inflow"), 0.2, 0.1)


def tiled_periodic():
    """Periodic Y with several rho_y tiles per element (wrapped halo rows)."""
    from test_oracle_full2d import periodic_mat
    m = periodic_mat("upwind", "fv-consistent")
    m["dg"]["params"].update({"N_K_X": 2, "N_K_Y": 2, "N_el_Y": 4, "N_rho_x": 64, "N_rho_y": 10})
    return m


ALL["tiled-periodic-64x10"] = (tiled_periodic(), 0.2, 0.1)


def oracle_parts(m, EfL, EfR):
    p = F.init_params_full2d(m, m["V"], EfL, EfR)
    b = F.boundary_full2d(m, p, EfL, EfR, m["V"])
    parts, rparts, _ = F.get_diff_full2d(m, p, b)
    return p, parts, rparts, F.DriftFull2D(m, p, m["V"])


def rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


@pytest.mark.parametrize("dv", ["interp", "table"])
@pytest.mark.parametrize("name", list(ALL))
def test_apply_diagonal_rowabs(name, dv):
    """Both sources of Delta V (interpolated in the apply, or streamed from
    the stored table; wdg2_set_dv_mode) against the oracle."""
    m, EfL, EfR = ALL[name]
    print(f"full2d apply {name} dv={dv}", flush=True)
    p, parts, rparts, drift = oracle_parts(m, EfL, EfR)
    n = p["nTotal"]
    rng = np.random.default_rng(0)
    u = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    y_or = F.apply_parts(parts, u, p) + drift.apply(u)
    with P.Full2DOperator(P.Full2DSetup(m, m["V"], EfL, EfR)) as op:
        op.set_dv_mode(dv)
        y = op.apply(u)
        assert rel(y, y_or) < 1e-12, rel(y, y_or)
        ym = op.apply(u, precision="mixed")
        assert rel(ym, y_or) < 2e-5, rel(ym, y_or)
        d_or = F.kron_diagonal(parts, p) + drift.diagonal()
        assert rel(op.diagonal(), d_or) < 1e-12
        s_or = F.kron_row_abs_sum(parts, p) + drift.row_abs()
        assert rel(op.row_abs_sum(), s_or) < 1e-12


@pytest.mark.parametrize("name", list(ALL))
def test_observables_of_the_same_state(name):
    """wdg2_observables on a given state against the oracle's
    get_Observables_full2D restatement of the same state: the density and
    both current rows at zero separation, every configuration (periodic Y,
    several rho_y tiles, p = 1 and 2)."""
    m, EfL, EfR = ALL[name]
    p = F.init_params_full2d(m, m["V"], EfL, EfR)
    rng = np.random.default_rng(7)
    u = rng.standard_normal(p["nTotal"]) + 1j * rng.standard_normal(p["nTotal"])
    with P.Full2DOperator(P.Full2DSetup(m, m["V"], EfL, EfR)) as op:
        nDG, jxDG, jyDG = op.observables(rho=u)
    n_or, jx_or, jy_or = F.observables_full2d(u, p, m)
    print(f"{name}: n {rel(nDG, n_or):.1e} jx {rel(jxDG, jx_or):.1e} jy {rel(jyDG, jy_or):.1e}")
    assert rel(nDG, n_or) < 1e-12
    assert rel(jxDG, jx_or) < 1e-12 and rel(jyDG, jy_or) < 1e-12


@pytest.mark.parametrize("name", ["rusanov-specular-fv", "upwind-zero-inflow", "central-closed-faces-collocated",
                                  "p1-even-rho", "periodic-rusanov-fv"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_solve_matches_direct_and_observables(name, precision):
    m, EfL, EfR = ALL[name]
    print(f"full2d solve {name} {precision}", flush=True)
    p, A, rhs, rho_or = F.solve_direct_full2d(m, m["V"], EfL, EfR)
    with P.Full2DOperator(P.Full2DSetup(m, m["V"], EfL, EfR)) as op:
        assert rel(op.assemble_rhs(), rhs) < 1e-12
        # restarted GMRES on these small non-normal systems needs the longest
        # Krylov space and row scaling (Jacobi stagnates, as on the CPU)
        assert op.set_preconditioner("rowabs") == "rowabs"
        op.upload_rhs()
        x, info = op.solve(SolveOptions(tol=1e-10, restart=64, max_iter=60000, precision=precision))
        print(f"iterations {info.iterations} relres {info.true_relres:.2e}", flush=True)
        assert info.flag == 0 and info.true_relres <= 1e-10
        assert rel(x, rho_or) < 1e-6, rel(x, rho_or)
        nDG, jxDG, jyDG = op.observables()
    n_or, jx_or, jy_or = F.observables_full2d(rho_or, p, m)
    assert rel(nDG, n_or) < 1e-6
    assert rel(jxDG, jx_or) < 1e-5 and rel(jyDG, jy_or) < 1e-5


def test_preconditioners_and_rhs_roundtrip():
    m, EfL, EfR = ALL["rusanov-specular-fv"]
    p, parts, rparts, drift = oracle_parts(m, EfL, EfR)
    with P.Full2DOperator(P.Full2DSetup(m, m["V"], EfL, EfR)) as op:
        op.upload_rhs()
        rhs = np.empty(op.n, complex)
        from real_space_dg_amd import _lib
        _lib.check(op._lib.wdg_rhs(op.ctx, _lib.dptr(rhs)), op.ctx)
        assert rel(rhs, F.assemble_rhs(rparts, p)) < 1e-14
        r = np.random.default_rng(3).standard_normal(op.n) + 0j
        assert op.set_preconditioner("rowabs") == "rowabs"
        s = F.kron_row_abs_sum(parts, p) + drift.row_abs()
        assert rel(op.precond(r), r / s) < 1e-12
        assert op.set_preconditioner("none") == "none"
        assert rel(op.precond(r), r) == 0
        x, info = op.solve(SolveOptions(tol=1e-6, restart=64, max_iter=60000))
        assert info.flag == 0


def test_reference_signature_entry_point():
    m = small_mat()
    m["dg"]["params"].update({"full2D_solverTol": 1e-9, "full2D_preconditioner": "rowabs",
                              "full2D_gmresRestart": 64, "full2D_solverMaxIt": 3000})
    res = P.solve_transport_DG_full2D(m, m["V"], 0.2, 0.1)
    assert res.status == "solved-gmres" and res.info["solve"]["trueRelres"] <= 1e-9
    p, A, rhs, rho_or = F.solve_direct_full2d(m, m["V"], 0.2, 0.1)
    assert rel(res.rho, rho_or) < 1e-5
    # projected onto the material Y grid only (get_Observables_full2D.m:137-158)
    nxd = P.Full2DSetup(m, m["V"], 0.2, 0.1).X.ndof
    assert res.n.shape == (nxd, m["Ny"]) and res.j.shape == (nxd,)
    n_or, _, _ = F.observables_full2d(rho_or, p, m)
    assert rel(res.nDG, n_or) < 1e-5
    # info.solve.resvec as the reference reports it (times ||rhs||,
    # solve_transport_DG_full2D.m:324): ||rhs|| first (zero start), one value
    # per iteration, the last one the final estimate below tol * ||rhs||
    rv, rn = res.info["solve"]["resvec"], np.linalg.norm(rhs)
    assert rv.size == res.info["solve"]["iterations"] + 1
    assert abs(rv[0] - rn) <= 1e-10 * rn
    assert rv[-1] <= 1e-9 * rn and np.all(np.isfinite(rv))


@pytest.mark.parametrize("name", ["rusanov-specular-fv", "upwind-zero-inflow", "central-closed-faces-collocated",
                                  "p1-even-rho", "tiled-80x7-upwind-p2", "periodic-rusanov-fv",
                                  "periodic-upwind-collocated"])
def test_line_transverse_matches_oracle(name):
    """pre.apply(r) of get_LineTransversePreconditioner_full2D.m on the device
    (modes, X sweeps, transverse inverses) against the oracle's LU form."""
    m, EfL, EfR = ALL[name]
    print(f"full2d line-transverse {name}", flush=True)
    p, parts, rparts, drift = oracle_parts(m, EfL, EfR)
    b = F.boundary_full2d(m, p, EfL, EfR, m["V"])
    _, _, info = F.get_diff_full2d(m, p, b)
    lt = F.LineTransverse(m, p, m["V"], info, drift)
    r = np.random.default_rng(5).standard_normal(p["nTotal"]) + 1j * np.random.default_rng(6).standard_normal(p["nTotal"])
    z_or = lt.apply(r)
    with P.Full2DOperator(P.Full2DSetup(m, m["V"], EfL, EfR)) as op:
        summary = op.set_line_transverse()
        assert summary["sigma"] == pytest.approx(lt.sigma, rel=1e-12)
        z = op.precond(r)
        assert rel(z, z_or) < 1e-10, rel(z, z_or)
        assert op.set_preconditioner("line-transverse") == "line-transverse"


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_line_transverse_solve(precision):
    """GMRES(40) right-preconditioned by the line-transverse factors with the
    production shift 0.03 (RTD/run_RTD_full2D.m:39) reaches the direct
    solution in about 600 iterations (scipy GMRES with the oracle's apply:
    586; row scaling needs about 8000 with restart 64)."""
    import copy
    m, EfL, EfR = ALL["rusanov-specular-fv"]
    m = copy.deepcopy(m)
    m["dg"]["params"]["full2D_lineRelativeShift"] = 0.03
    p, A, rhs, rho_or = F.solve_direct_full2d(m, m["V"], EfL, EfR)
    with P.Full2DOperator(P.Full2DSetup(m, m["V"], EfL, EfR)) as op:
        op.upload_rhs()
        op.set_line_transverse()
        x, info = op.solve(SolveOptions(tol=1e-10, restart=40, max_iter=20000, precision=precision))
        print(f"line-transverse {precision}: iterations {info.iterations} relres {info.true_relres:.2e}", flush=True)
        assert info.flag == 0 and rel(x, rho_or) < 1e-6
        assert info.iterations < 1000


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_crank_nicolson_full2d_matches_oracle(precision):
    """Crank-Nicolson steps of the full-2D operator (rowabs rebuilt for the
    shifted operator) against the oracle's directly solved steps; the state
    moves towards the stationary solution."""
    from oracle import transient as TR
    m, EfL, EfR = ALL["rusanov-specular-fv"]
    p, A, rhs, rho_inf = F.solve_direct_full2d(m, m["V"], EfL, EfR)
    lam = abs(A).sum(axis=1).max()
    dt = 20.0 / lam
    rho0 = np.zeros(p["nTotal"], complex)
    ref = TR.crank_nicolson(A, rhs, rho0, dt, 8)
    with P.Full2DOperator(P.Full2DSetup(m, m["V"], EfL, EfR)) as op:
        op.upload_rhs()
        assert op.set_preconditioner("rowabs") == "rowabs"
        tol = 1e-12 if precision == "fp64" else 1e-10
        rho, info = op.transient_cn(dt, 8, tol=tol, restart=40, precision=precision, rho0=rho0)
        print(f"CN full2D {precision}: {info}", flush=True)
        assert info["flag"] == 0 and info["steps"] == 8
        assert rel(rho, ref) < 1e-8, rel(rho, ref)
        assert np.linalg.norm(rho - rho_inf) < np.linalg.norm(rho0 - rho_inf)
        nDG, _, _ = op.observables()
        n_ref, _, _ = F.observables_full2d(ref, p, m)
        assert rel(nDG, n_ref) < 1e-7
        # the stationary solve still works after the transient (preconditioner restored)
        x, sinfo = op.solve(SolveOptions(tol=1e-8, restart=64, max_iter=60000))
        assert sinfo.flag == 0


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_crank_nicolson_batch_matches_single_runs(precision):
    """wdg_transient_cn_batch over three full-2D instances of different
    fluxes and boundaries gives each instance the state its own
    wdg_transient_cn run gives, and the oracle's directly solved steps
    (small grids: the oracle's sparse LU of the larger tiled ones takes
    minutes)."""
    from oracle import transient as TR
    from real_space_dg_amd.solver import transient_cn_batch
    names = ["rusanov-specular-fv", "upwind-zero-inflow", "periodic-rusanov-fv"]
    tol = 1e-12 if precision == "fp64" else 1e-10
    ops, refs, rho0s, dts = [], [], [], []
    try:
        for k, name in enumerate(names):
            m, EfL, EfR = ALL[name]
            p, A, rhs, rho_inf = F.solve_direct_full2d(m, m["V"], EfL, EfR)
            dt = 20.0 / abs(A).sum(axis=1).max()
            rng = np.random.default_rng(5 + k)
            rho0 = 0.1 * np.max(np.abs(rho_inf)) * (rng.standard_normal(p["nTotal"])
                                                     + 1j * rng.standard_normal(p["nTotal"]))
            op = P.Full2DOperator(P.Full2DSetup(m, m["V"], EfL, EfR))
            op.upload_rhs()
            assert op.set_preconditioner("rowabs") == "rowabs"
            ops.append(op)
            print(f"CN batch instance {name}: N = {p['nTotal']}", flush=True)
            rho0s.append(rho0)
            dts.append(dt)
            refs.append((A, rhs))
        # one time step for all: the smallest of the three stability scales
        dt = min(dts)
        single = [op.transient_cn(dt, 5, tol=tol, restart=40, precision=precision, rho0=r0)
                  for op, r0 in zip(ops, rho0s)]
        outs, infos = transient_cn_batch(ops, dt, 5, tol=tol, restart=40, precision=precision, rho0=rho0s)
        print(f"CN batch {precision}: {infos}", flush=True)
        for (rho_s, info_s), rho_b, info_b, (A, rhs), r0 in zip(single, outs, infos, refs, rho0s):
            assert info_b["flag"] == 0 and info_b["steps"] == 5
            assert info_b["iterations"] == info_s["iterations"]
            assert rel(rho_b, rho_s) < 1e-12
            ref = TR.crank_nicolson(A, rhs, r0, dt, 5)
            assert rel(rho_b, ref) < 1e-8, rel(rho_b, ref)
        # the batch leaves each instance's final state on the device
        for op, rho_b, name in zip(ops, outs, names):
            nDG, _, _ = op.observables()
            n_ref, _, _ = op.observables(rho_b)
            assert rel(nDG, n_ref) < 1e-12
        # an empty batch is a no-op, a repeated instance an error
        assert transient_cn_batch([], dt, 1) == ([], [])
        from real_space_dg_amd._lib import WdgError
        with pytest.raises(WdgError, match="twice"):
            transient_cn_batch([ops[0], ops[0]], dt, 1)
    finally:
        for op in ops:
            op.close()
