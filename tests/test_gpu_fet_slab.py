"""The full-2D FET-slab transient (Crank-Nicolson, dt = 1 fs) on a reduced
grid against the oracle: 8 x 4 DG elements p = 1, 8 x 8 relative cells
(N = 8,192; the benchmark runs 64 x 32 x 64 x 64 = 33.5 M), Gaussian gate
barrier 0.2 eV with sigma 5 nm, rusanov fluxes, row-sum preconditioner.

The oracle applies the full-2D operator (oracle/full2d.py: transport parts
of DG_code/core/get_Diff_full2D.m and the FV-consistent drift of
get_Drift_full2D.m) and steps it with a direct sparse LU per step
(oracle.transient.crank_nicolson).  Single instance in both precisions, and
the eight transverse-mode instances (gate potential + 5 meV per mode) of one
batched launch (wdg_transient_cn_batch).
"""
import numpy as np
import pytest

from oracle import full2d as F
from oracle import transient as TR
from real_space_dg_amd import full2d as P
from real_space_dg_amd.devices import fet_slab

pytestmark = pytest.mark.gpu

DT = 1e-15
EFL, EFR = 0.1, 0.05


def _system(m, V):
    """The oracle's matrix-free full-2D operator: apply, row-abs sums, rhs."""
    p = F.init_params_full2d(m, V, EFL, EFR)
    bnd = F.boundary_full2d(m, p, EFL, EFR, V)
    parts, rparts, _ = F.get_diff_full2d(m, p, bnd)
    drift = F.DriftFull2D(m, p, V)
    apply_a = lambda u: F.apply_parts(parts, u, p) + drift.apply(u)  # noqa: E731
    rowabs = F.kron_row_abs_sum(parts, p) + drift.row_abs()
    return p, apply_a, rowabs, F.assemble_rhs(rparts, p)


def _cn(system, nsteps):
    p, apply_a, rowabs, rhs = system
    return TR.crank_nicolson_mf(apply_a, rhs, np.zeros(p["nTotal"], complex), DT, nsteps, tol=1e-13,
                                row_scale=rowabs)


def rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


@pytest.fixture(scope="module")
def slab():
    m, V = fet_slab(n_el_x=8, n_el_y=4, n_rho=8)
    print(f"\n[case] FET slab {m['Nx']}x{m['Ny']} material points, N_rho 8x8", flush=True)
    return m, V, _system(m, V)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_fet_slab_crank_nicolson_matches_oracle(slab, precision):
    m, V, system = slab
    p = system[0]
    n = p["nTotal"]
    ref = _cn(system, 3)
    with P.Full2DOperator(P.Full2DSetup(m, V, EFL, EFR)) as op:
        assert op.n == n
        op.upload_rhs()
        assert op.set_preconditioner("rowabs") == "rowabs"
        tol = 1e-12 if precision == "fp64" else 1e-10
        rho, info = op.transient_cn(DT, 3, tol=tol, restart=30, precision=precision, rho0=np.zeros(n, complex))
        print(f"FET slab CN {precision}: {info}", flush=True)
        assert info["flag"] == 0 and info["steps"] == 3
        assert rel(rho, ref) < 1e-8, rel(rho, ref)
        nDG, jx, _ = op.observables()
        n_ref, jx_ref, _ = F.observables_full2d(ref, p, m)
        assert rel(nDG, n_ref) < 1e-8
        assert rel(jx, jx_ref) < 1e-7


@pytest.mark.parametrize("basis", ["fp16", "bf16"])
@pytest.mark.parametrize("precond", ["rowabs", "jacobi"])
def test_mixed_basis_read_in_place(slab, precond, basis):
    """The FP32 solver's scaling preconditioners read the 16-bit basis vector
    in place (f2_scale_basis_kernel: real row-sum scale, complex Jacobi scale;
    FP16 and BF16 packets): the same three steps against the oracle."""
    m, V, system = slab
    p = system[0]
    n = p["nTotal"]
    ref = _cn(system, 3)
    with P.Full2DOperator(P.Full2DSetup(m, V, EFL, EFR)) as op:
        op.upload_rhs()
        used = op.set_preconditioner(precond)
        print(f"preconditioner {precond} -> {used}, basis {basis}", flush=True)
        rho, info = op.transient_cn(DT, 3, tol=1e-10, restart=30, precision="mixed", rho0=np.zeros(n, complex),
                                    krylov_basis=basis)
        print(f"FET slab CN mixed {used} {basis}: {info}", flush=True)
        assert info["flag"] == 0 and info["steps"] == 3
        assert rel(rho, ref) < 1e-8, rel(rho, ref)


def test_fet_slab_eight_mode_batch_matches_oracle(slab):
    from real_space_dg_amd.solver import transient_cn_batch
    m, V, system0 = slab
    p = system0[0]
    n = p["nTotal"]
    ops, refs = [], []
    try:
        for k in range(8):
            Vk = V + 0.005 * k
            refs.append(_cn(system0 if k == 0 else _system(m, Vk), 2))
            op = P.Full2DOperator(P.Full2DSetup(m, Vk, EFL, EFR))
            op.upload_rhs()
            op.set_preconditioner("rowabs")
            ops.append(op)
        outs, infos = transient_cn_batch(ops, DT, 2, tol=1e-10, restart=30, precision="mixed",
                                         rho0=[np.zeros(n, complex)] * 8)
        print(f"FET slab 8-mode batch: {[i['iterations'] for i in infos]}", flush=True)
        for rho, info, ref in zip(outs, infos, refs):
            assert info["flag"] == 0 and info["steps"] == 2
            assert rel(rho, ref) < 1e-8, rel(rho, ref)
        # the modes differ (the potential offset moves the state)
        assert rel(outs[7], outs[0]) > 1e-6
    finally:
        for op in ops:
            op.close()


def test_self_consistent_batch_matches_oracle_blocks():
    """The self-consistent FET slab (contact-mode reservoirs built on the
    device, n+ contacts, 0.05 V) with two transverse-mode instances (offsets
    0 and 5 meV) stepped together, a Poisson update every 3 steps from their
    summed density (poisson2d.transient_cn_batch_self_consistent), against
    the oracle block by block: each instance's Crank-Nicolson steps by the
    oracle's matrix-free operator at the potential Vc - phi + offset the
    block ran with (reservoirs as the setup built them), and the next phi
    from the oracle states through the same projection and Poisson step."""
    import copy

    from real_space_dg_amd import poisson2d as PS
    from real_space_dg_amd.devices import fet_slab_self_consistent

    mat = fet_slab_self_consistent(n_el_x=8, n_el_y=4, n_rho=8)
    bias, offsets = 0.05, [0.0, 0.005]
    x = np.asarray(mat["x"], float)
    phi0 = np.repeat((bias * x / x[-1])[:, None], mat["Ny"], axis=1)
    Vc = np.asarray(mat["Vc"], float)
    ops = []
    try:
        for off in offsets:
            s = P.Full2DSetup(mat, Vc - phi0 + off, EFL, EFR)
            op = P.Full2DOperator(s)
            op.upload_contact_rhs()
            op.set_preconditioner("rowabs")
            ops.append(op)
        n = ops[0].n
        seen = []
        phi, hist = PS.transient_cn_batch_self_consistent(ops, offsets, phi0, bias, DT, 6, poisson_every=3,
                                                          precision="fp64", tol=1e-12,
                                                          rho0=[np.zeros(n, complex)] * 2, on_block=seen.append)
        print("SC blocks:", [(h["steps"], h["gmresIterations"], round(h["potentialUpdate"], 6)) for h in hist],
              flush=True)
        assert len(hist) == 2 and all(h["flag"] == 0 for h in hist)
        assert [h["steps"] for h in seen] == [3, 6] and all(a is b for a, b in zip(seen, hist))
        # the oracle, block by block, with the reservoirs of each instance's setup
        M, types, periodic = PS.eval_poisson2d(mat)
        states = [np.zeros(n, complex)] * 2
        phi_ref = phi0
        for b, h in enumerate(hist):
            assert np.max(np.abs(h["phi"] - phi_ref)) <= 1e-9 * max(1.0, np.max(np.abs(phi_ref)))
            nsum = 0
            for i, (op, off) in enumerate(zip(ops, offsets)):
                mi = copy.deepcopy(mat)
                mi["dg"]["params"]["full2D_reservoirModel"] = "zero"
                host = P.Full2DSetup(mat, Vc - phi0 + off, EFL, EFR)
                mi["dg"]["params"]["full2D_sourceRho"] = host.boundary["source"]["rho"]
                mi["dg"]["params"]["full2D_drainRho"] = host.boundary["drain"]["rho"]
                V = Vc - phi_ref + off
                p, apply_a, rowabs, rhs = _system(mi, V)
                states[i] = TR.crank_nicolson_mf(apply_a, rhs, states[i], DT, 3, tol=1e-13, row_scale=rowabs)
                nsum = nsum + np.real(F.observables_full2d(states[i], p, mi)[0])
            npois, _ = PS.project_dg_density(nsum, ops[0].setup, periodic)
            phi_ref, _ = PS.poisson_step(M, types, npois, mat, phi_ref, bias, "newton-raphson")
        for op, st in zip(ops, states):
            assert rel(_device_state(op), st) < 1e-8
        assert np.max(np.abs(phi - phi_ref)) <= 1e-8 * max(1.0, np.max(np.abs(phi_ref)))
    finally:
        for op in ops:
            op.close()


def _device_state(op):
    from real_space_dg_amd._lib import check
    from real_space_dg_amd._lib import dptr
    x = np.empty(op.n, np.complex128)
    check(op._lib.wdg_solution(op.ctx, dptr(x)), op.ctx)
    return x
