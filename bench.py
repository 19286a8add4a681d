#!/usr/bin/env python3
"""Bench: stationary Wigner solves/s to true relative residual 1e-10 on one MI355X.

Workload (RTD with bias): 60 nm double-barrier RTD (0.3 eV x 3 nm barriers
around a 5 nm well), 0.2 V linear drop, 512 DG elements of degree 3 in x,
1024 plane waves / FV cells in k (N = 2,097,152 complex unknowns), GMRES(30)
with the right line preconditioner, stopped at true relres 1e-11 (below the
1e-10 target, so that density and current are within 1e-8 of the FP64 oracle).

A step = one stationary solve of that instance from the zero start: build the
potential tables, inflow RHS and preconditioner data on the device
(wdg_set_potential), GMRES to true relres <= 1e-10 (wdg_solve) and the
observables n, j (wdg_observables).  All solver state stays in HBM; only
n and j return to the host.  Timed: the mixed-precision solver (FP32 inner
GMRES cycles with an FP16-stored Krylov basis, Theta[V] by FP32 chirp-z
transforms, FP64 residual refinement) and, for comparison, the plain FP64
solver.  The timed solves replay captured HIP graphs of the restart cycles;
kernel durations are HIP event pairs on the context stream from a profiled
pass over the same solves (kernel-by-kernel launches, since event pairs do
not sit inside a replayed graph).  Every library call synchronises its
stream before returning, so the wall clock brackets finished GPU work.

Beside the metric the line reports the batched 64-instance bias sweep, the
16-bias self-consistent (Gummel) workload and the full-2D FET-slab transient
(Crank-Nicolson steps of 33.5M unknowns), each in both precisions.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "GMRES solves/s to 1e-10 (mixed vs FP64, 1 GPU); Wigner apply % HBM roofline"
HBM_PEAK_GBS = 8000.0
MFMA_PEAK_TFS = {"bf16": 2500.0, "fp64": 78.6, "fp32": 157.3}
M0 = 9.109e-31

WORKLOAD = dict(n_elem=512, degree=3, nk=1024, bias_v=0.2)
MFMA_KERNELS = {"theta_gemm1", "theta_gemm2"}   # the GEMM formulation of Theta (not the default)
NO_GRAPHS = False   # --no-graphs: solves and transients launch their restart cycles kernel by kernel


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--fp64-steps", type=int, default=2)
    ap.add_argument("--profile-steps", type=int, default=2,
                    help="solves per precision in the profiled pass that times each kernel")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-iters", type=int, default=30,
                    help="GMRES iterations of the CPU baseline's sample at the headline workload (one restart cycle)")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the batched-sweep and self-consistent workloads reported beside the metric")
    ap.add_argument("--only", choices=["batched_sweep", "self_consistent", "transient_full2d", "transient_full2d_batch",
                                       "transient_full2d_sc"],
                    default=None,
                    help="run just this side workload and print its JSON (profiling runs)")
    ap.add_argument("--full2d-steps", type=int, default=2, help="Crank-Nicolson steps of the full-2D transient")
    ap.add_argument("--full2d-instances", type=int, default=8,
                    help="transverse-mode instances of the batched full-2D transient (--only transient_full2d_batch)")
    ap.add_argument("--full2d-precisions", default="mixed,fp64", help="precisions of the full-2D transient (--only)")
    ap.add_argument("--full2d-dv", choices=["interp", "table"], default="table",
                    help="Delta V source of the full-2D apply: a stored table (the library default), or "
                         "interpolated in the kernel")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch the solves' and transients' restart cycles kernel by kernel instead of "
                         "replaying captured graphs (same kernels; for kernel tracers)")
    ap.add_argument("--theta", choices=["auto", "fft", "gemm"], default="auto",
                    help="formulation of Theta[V] in the headline solves (wdg_set_theta_method)")
    ap.add_argument("--calibrate-copy", type=float, default=0.0, metavar="BYTES",
                    help="first run one plain device copy of BYTES bytes (HBM-counter calibration)")
    return ap.parse_args()


def chirp_length(nk: int) -> int:
    """Transform length of the chirp-z Theta: the smallest 64 * 2^a 3^b >= 2 nk - 1."""
    P = 2 * nk - 1
    while True:
        m, r = divmod(P, 64)
        if r == 0:
            while m % 2 == 0:
                m //= 2
            while m % 3 == 0:
                m //= 3
            if m == 1:
                return P
        P += 1


def kernel_roofline(prof: dict, name: str, peak_kind: str, precision: str = "mixed",
                    t_launch_ms: float | None = None) -> dict | None:
    """Roofline of one kernel from the profiled pass; t_launch_ms (when given)
    replaces the pass's per-launch event timing as the launch duration."""
    st = prof[name]
    if not st["launches"]:
        return None
    t_ms = st["avg_ms"] if t_launch_ms is None else t_launch_ms
    per_launch = st["units"] / max(st["launches"], 1)
    if name in MFMA_KERNELS:
        ach = per_launch / (t_ms * 1e-3) / 1e12
        peak = MFMA_PEAK_TFS[peak_kind]
        out = {"bound": "mfma", "achieved": ach, "peak": peak, "unit": "TFLOP/s", "frac": ach / peak}
    else:
        if name == "theta_fft":
            # the whole operator apply: read u once, write y once (2 s N bytes;
            # the potential's edge table counts as generated on the fly)
            s = 16 if precision == "fp64" else 8
            N = WORKLOAD["n_elem"] * (WORKLOAD["degree"] + 1) * WORKLOAD["nk"]
            per_launch = 2.0 * s * N
        ach = per_launch / (t_ms * 1e-3) / 1e9
        out = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS}
    out.update({"kernel": name, "avg_launch_ms": t_ms, "launches": st["launches"],
                "algorithmic_per_launch": per_launch, "traffic": traffic_for(name, precision)})
    if t_launch_ms is not None:
        out["launch_timing"] = ("back-to-back applies between one HIP event pair on the context stream "
                                "(wdg_time_apply, 50 launches)")
        out["avg_launch_ms_event_pairs_in_solves"] = st["avg_ms"]
    busy = valu_busy_for(name, precision)
    if busy is not None:
        out["valu_busy"] = busy
    if name == "theta_fft":
        # VALU-bound: four length-P transforms per DG node at 5 P log2 P flops each
        P = chirp_length(WORKLOAD["nk"])
        nodes = WORKLOAD["n_elem"] * (WORKLOAD["degree"] + 1)
        flops = 4 * 5 * P * np.log2(P) * nodes
        out["fft_flops_per_launch"] = flops
        out["fft_tflops"] = flops / (t_ms * 1e-3) / 1e12
    return out


def valu_busy_for(name: str, precision: str = "mixed"):
    """Fraction of the kernel's duration its SIMDs spend issuing VALU
    instructions (SQ_INSTS_VALU x 4 cycles / (SIMDs x cycles)), from the
    rocprofv3 SQ counter pass recorded in profiles/traffic.json; None if not
    recorded."""
    f = ROOT / "profiles" / "traffic.json"
    if not f.exists():
        return None
    try:
        d = json.loads(f.read_text())
        v = d.get("kernels_fp64" if precision == "fp64" else "kernels", {}).get(name)
        return None if v is None else v.get("valu_busy")
    except (OSError, ValueError):
        return None


def traffic_for(name: str, precision: str = "mixed"):
    """HBM bytes per launch of the bench workload's kernel, measured with two
    rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) and recorded in
    profiles/traffic.json; None if not recorded."""
    f = ROOT / "profiles" / "traffic.json"
    if not f.exists():
        return None
    try:
        d = json.loads(f.read_text())
        v = d.get("kernels_fp64" if precision == "fp64" else "kernels", {}).get(name)
        return None if v is None else v.get("bytes_per_launch")
    except (OSError, ValueError):
        return None


def cpu_baseline(mat, E, efl, efr, its_per_solve: int, sample_iters: int) -> dict:
    """The oracle's FP64 matrix-free operator + line preconditioner + GMRES on
    the host: setup of one instance plus a bounded number of GMRES iterations,
    scaled to the iteration count of a full solve."""
    from oracle import krylov as K
    from oracle import wigner1d as W
    try:
        from threadpoolctl import threadpool_info
        threads = max([i.get("num_threads", 1) for i in threadpool_info()] or [1])
    except Exception:
        threads = os.cpu_count() or 1
    m = mat["me_x"][0, :, 0] * M0
    t0 = time.perf_counter()
    p = W.init_params(mat, efl, efr, E[0, :, 0], m, 1.0)
    op = W.MatrixFreeOperator.build(mat, p, E[0, :, 0])
    Bbar = op.edge_coefficients()
    pre = K.LinePreconditioner(op, Bbar)
    _, rhs = W.get_diff(p)
    t_setup = time.perf_counter() - t0
    t1 = time.perf_counter()
    _, info = K.gmres(lambda u: op.apply(u, Bbar), pre.apply, rhs, restart=30, tol=1e-10,
                      maxit=sample_iters)
    t_iter = (time.perf_counter() - t1) / max(info["iterations"], 1)
    t_solve = t_setup + t_iter * its_per_solve
    return {"value": 1.0 / t_solve, "unit": "solves/s", "cores": int(threads), "kind": "port",
            "sample": (f"oracle FP64 restatement (numpy, {threads} BLAS threads): instance setup "
                       f"{t_setup:.2f} s + {info['iterations']} GMRES(30) iterations at "
                       f"{t_iter:.3f} s each on the same workload, scaled to {its_per_solve} "
                       f"iterations per solve (the FP64 GPU solver's count)"),
            "seconds_per_iteration": t_iter, "setup_seconds": t_setup}


def c1_timed_solves(threads: int) -> dict:
    """The plain-method workload (64 elements p=2 x 128 k, zero bias, GMRES(30)
    + line preconditioner to 1e-10) solved completely on the host by the
    oracle and on the GPU in both precisions: wall times of whole solves, no
    extrapolation."""
    from oracle import krylov as K
    from oracle import wigner1d as W
    from real_space_dg_amd.devices import rtd_1d
    from real_space_dg_amd.solver import SolveOptions, WignerOperator
    mat, E, _, efl, efr = rtd_1d(64, 2, 128, bias_v=0.0)
    m = mat["me_x"][0, :, 0] * M0
    t0 = time.perf_counter()
    p = W.init_params(mat, efl, efr, E[0, :, 0], m, 1.0)
    op = W.MatrixFreeOperator.build(mat, p, E[0, :, 0])
    Bbar = op.edge_coefficients()
    pre = K.LinePreconditioner(op, Bbar)
    _, rhs = W.get_diff(p)
    _, cinfo = K.gmres(lambda u: op.apply(u, Bbar), pre.apply, rhs, restart=30, tol=1e-10, maxit=20000)
    t_cpu = time.perf_counter() - t0
    out = {"workload": "RTD 60 nm, zero bias, 64 DG elements p=2 x 128 k, GMRES(30) + line preconditioner "
                       "to true relres 1e-10 (setup included on both sides)",
           "cpu": {"seconds": t_cpu, "iterations": int(cinfo["iterations"]), "threads": threads,
                   "true_relres": float(cinfo["true_relres"]) if "true_relres" in cinfo else None,
                   "kind": "oracle FP64 restatement (numpy), one complete solve"}}
    g = WignerOperator.from_mat(mat)
    for prec in ("fp64", "mixed"):
        def solve():
            g.set_potential(mat["x"] * 1e-9, E[0, :, 0], efl, efr, m[0], m[-1], temp=mat["Temp"])
            return g.solve(SolveOptions(tol=1e-10, restart=30, precision=prec), return_x=False)[1]
        solve()
        reps = 10
        t0 = time.perf_counter()
        infos = [solve() for _ in range(reps)]
        dt = (time.perf_counter() - t0) / reps
        out["gpu_" + prec] = {"seconds": dt, "iterations": int(infos[-1]["iterations"]),
                              "true_relres": float(infos[-1]["true_relres"])}
    g.close()
    out["gpu_fp64_over_cpu"] = out["cpu"]["seconds"] / out["gpu_fp64"]["seconds"]
    print(f"[bench] C1 complete solves: CPU {t_cpu:.2f} s, GPU FP64 {1e3 * out['gpu_fp64']['seconds']:.2f} ms",
          file=sys.stderr, flush=True)
    return out


def copy_calibration(nbytes: float):
    """One plain device-to-device copy of a known size (counter calibration)."""
    import torch
    n = int(nbytes) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.mul(a, 1.0, out=b)
    torch.cuda.synchronize()
    del a, b


def batched_sweep(n_instances: int = 64, seed: int = 0) -> dict:
    """64-instance bias sweep (uneven grids from ``seed``) solved by one
    batched call per precision; solves/s over the whole batch."""
    from real_space_dg_amd.devices import rtd_sweep
    from real_space_dg_amd.solver import SolveOptions, WignerOperator, solve_batch
    cases = rtd_sweep(n_instances, seed=seed)
    ops = []
    for mat, E, _, efl, efr in cases:
        op = WignerOperator.from_mat(mat)
        m = mat["me_x"][0, :, 0] * M0
        op.set_potential(mat["x"] * 1e-9, E[0, :, 0], efl, efr, m[0], m[-1], temp=mat["Temp"])
        ops.append(op)
    out = {"workload": f"{n_instances} RTD instances, seed {seed}: bias U[0,0.4] V, elements in "
                       "{128,256,384,512} (p=3), k in {256,512,768,1024}, barrier U[0.25,0.35] eV; "
                       "one wdg_solve_batch call, tol 1e-10",
           "unknowns_total": int(sum(op.N for op in ops))}
    for prec in ("mixed", "fp64"):
        opts = SolveOptions(tol=1e-10, precision=prec, restart=30, restart_max=64)
        solve_batch(ops[:2], opts)   # graph/tile warmup
        t0 = time.perf_counter()
        infos = solve_batch(ops, opts)
        dt = time.perf_counter() - t0
        its = [i["iterations"] for i in infos]
        # total work in unknown-iterations: the batch time over it is the
        # batch's cost per unknown and GMRES iteration, comparable with one
        # large solve's (the batch is work-bound when the two agree)
        work = float(sum(it * op.N for it, op in zip(its, ops)))
        out[prec] = {"solves_per_s": n_instances / dt, "seconds": dt, "converged": int(sum(i["flag"] == 0 for i in infos)),
                     "iterations_mean": float(np.mean(its)), "iterations_max": int(max(its)),
                     "true_relres_max": float(max(i["true_relres"] for i in infos)),
                     "unknown_iterations": work, "ps_per_unknown_iteration": 1e12 * dt / work,
                     "slowest_instance": {"iterations": int(max(its)), "unknowns": int(ops[int(np.argmax(its))].N)}}
        print(f"[bench] batched sweep {prec}: {dt:.2f} s", file=sys.stderr, flush=True)
    out["mixed_over_fp64"] = out["mixed"]["solves_per_s"] / out["fp64"]["solves_per_s"]
    for op in ops:
        op.close()
    return out


def self_consistent(n_bias: int = 16) -> dict:
    """Gummel loop of 16 bias points (256 elements p=3, 512 k), FP64 Poisson
    Newton updates on the device, Anderson mixing, stop at 1e-6 V."""
    from real_space_dg_amd.devices import rtd_self_consistent
    from real_space_dg_amd.gummel import GummelOptions, solve_rtd_self_consistent
    from real_space_dg_amd.solver import SolveOptions
    mat = rtd_self_consistent(n_elem=256, degree=3, nk=512, n_bias=n_bias, bias_max=0.3)
    out = {"workload": f"self-consistent RTD, {n_bias} biases in [0,0.3] V, 256 elements p=3 x 512 k, "
                       "n+ 1e24 m^-3 in 20 nm contacts, 1e21 m^-3 elsewhere, eps_r 12.9, Gummel tol 1e-6 V"}
    res = {}
    for prec in ("mixed", "fp64"):
        opt = GummelOptions(max_iterations=60, potential_tolerance=1e-6, acceleration="anderson", mixing=0.5)
        opt.transport = SolveOptions(tol=1e-10, precision=prec, restart=30, restart_max=64)
        r = solve_rtd_self_consistent(mat, opt)
        res[prec] = r
        out[prec] = {"seconds": r["seconds"], "bias_points_per_s": n_bias / r["seconds"],
                     "converged": int(np.sum(r["converged"])), "gummel_iterations_max": int(np.max(r["iterations"])),
                     "gmres_iterations_total": int(np.sum(r["gmres_iterations"]))}
        print(f"[bench] self-consistent {prec}: {r['seconds']:.2f} s", file=sys.stderr, flush=True)
    a, b = res["mixed"], res["fp64"]
    out["mixed_vs_fp64"] = {
        "phi_max_abs_diff_V": float(max(np.max(np.abs(x - y)) for x, y in zip(a["phi"], b["phi"]))),
        "density_max_rel_diff": float(max(np.max(np.abs(x - y)) / np.max(np.abs(y))
                                          for x, y in zip(a["density"], b["density"]))),
        "current_max_rel_diff": float(np.max(np.abs(a["current"] - b["current"])) / np.max(np.abs(b["current"]))),
    }
    return out


def transient_full2d(steps: int = 2, dt: float = 1e-15, precisions=("mixed", "fp64"), dv: str = "table",
                     compare_interp: bool = True) -> dict:
    """Full-2D FET slab of the transient workload: 40x20 nm, 64x32 DG elements
    p=1, 64x64 relative cells (N = 33,554,432), Gaussian 0.2 eV gate barrier
    (sigma 5 nm); Crank-Nicolson steps of dt from the zero state with one
    row-scaled GMRES(30) solve to 1e-8 per step, in FP64 and mixed precision.
    Reports steps/s, iterations per step, the density difference of the two
    precisions and the full-2D apply kernel against the HBM roofline; with
    compare_interp the mixed steps run once more with Delta V interpolated
    in the apply instead of streamed from the table."""
    from real_space_dg_amd import full2d as F2
    from real_space_dg_amd.devices import fet_slab
    from real_space_dg_amd.solver import PROFILE_BITS
    mat, V = fet_slab()
    setup = F2.Full2DSetup(mat, V, 0.1, 0.05)
    out = {"workload": "FET slab 40x20 nm, 64x32 DG elements p=1 x 64x64 relative cells, Gaussian gate "
                       f"0.2 eV sigma 5 nm; Crank-Nicolson dt {dt:g} s x {steps} steps from zero, GMRES(30) + "
                       "row scaling to 1e-8 per step", "unknowns": setup.n_total}
    dens = {}
    out["delta_v"] = dv
    with F2.Full2DOperator(setup) as op:
        op.set_dv_mode(dv)
        op.upload_rhs()
        op.set_preconditioner("rowabs")
        op.transient_cn(dt, 1, tol=1e-8, precision="mixed", rho0=np.zeros(setup.n_total, complex),
                        return_rho=False, use_graph=not NO_GRAPHS)   # graphs / tables warm
        mask = (1 << PROFILE_BITS["f2_apply"]) | (1 << PROFILE_BITS["f2_dense"])
        for prec in precisions:
            # the library launches FP64 full-2D Crank-Nicolson cycles kernel by
            # kernel (graph replay of them is off, wdg_cn_opts)
            use_graph = prec != "fp64" and not NO_GRAPHS
            op.set_profiling(0)
            t0 = time.perf_counter()
            _, info = op.transient_cn(dt, steps, tol=1e-8, precision=prec, rho0=np.zeros(setup.n_total, complex),
                                      return_rho=False, use_graph=use_graph)
            wall = time.perf_counter() - t0
            nDG, _, _ = op.observables()
            dens[prec] = nDG
            op.set_profiling(mask)      # kernel durations: one more step, launched kernel by kernel
            op.transient_cn(dt, 1, tol=1e-8, precision=prec, rho0=np.zeros(setup.n_total, complex),
                            return_rho=False)
            prof = op.profile()
            op.set_profiling(0)
            out[prec] = {"steps_per_s": steps / wall, "seconds": wall, "graph_replay": use_graph,
                         "gmres_iterations_per_step":
                         info["iterations"] / max(info["steps"], 1), "max_step_relres": info["max_relres"],
                         "flag": info["flag"]}
            k = prof["f2_apply"]
            if k["launches"]:
                ach = k["units"] / k["launches"] / (k["avg_ms"] * 1e-3) / 1e9
                out[prec]["f2_apply"] = {"avg_ms": k["avg_ms"], "launches": k["launches"], "bound": "hbm",
                                         "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                         "frac": ach / HBM_PEAK_GBS}
            d = prof["f2_dense"]
            if d["launches"]:
                out[prec]["f2_dense_avg_ms"] = d["avg_ms"]
            print(f"[bench] transient full2D {prec}: {wall:.2f} s", file=sys.stderr, flush=True)
        if compare_interp and dv == "table" and "mixed" in precisions:
            op.set_dv_mode("interp")
            op.transient_cn(dt, 1, tol=1e-8, precision="mixed", rho0=np.zeros(setup.n_total, complex),
                            return_rho=False, use_graph=not NO_GRAPHS)
            t0 = time.perf_counter()
            op.transient_cn(dt, steps, tol=1e-8, precision="mixed", rho0=np.zeros(setup.n_total, complex),
                            return_rho=False, use_graph=not NO_GRAPHS)
            wall = time.perf_counter() - t0
            op.set_profiling(mask)
            op.transient_cn(dt, 1, tol=1e-8, precision="mixed", rho0=np.zeros(setup.n_total, complex),
                            return_rho=False)
            k = op.profile()["f2_apply"]
            op.set_profiling(0)
            out["mixed_delta_v_interp"] = {"steps_per_s": steps / wall, "f2_apply_avg_ms": k["avg_ms"]}
            print(f"[bench] transient full2D mixed, Delta V interpolated: {wall:.2f} s", file=sys.stderr, flush=True)
    if "mixed" in out and "fp64" in out:
        out["mixed_over_fp64"] = out["mixed"]["steps_per_s"] / out["fp64"]["steps_per_s"]
        out["mixed_vs_fp64_density_rel_diff"] = float(np.max(np.abs(dens["mixed"] - dens["fp64"])) /
                                                      np.max(np.abs(dens["fp64"])))
    return out


def transient_full2d_sc(instances: int = 8, steps: int = 20, poisson_every: int = 10, dt: float = 1e-15,
                        bias: float = 0.05, precisions=("mixed",)) -> dict:
    """The self-consistent FET-slab transient: the slab with n+ contacts,
    contact-mode reservoirs (built on the device) and `instances`
    transverse-mode instances (gate potential + 5 meV per mode) stepped
    together by Crank-Nicolson (dt 1 fs), a 2D Poisson update of the shared
    potential every `poisson_every` steps from the summed density
    (poisson2d.transient_cn_batch_self_consistent).  Reports instance-steps/s
    per precision, GMRES iterations and the Poisson updates."""
    from real_space_dg_amd import full2d as F2
    from real_space_dg_amd import poisson2d as PS
    from real_space_dg_amd.devices import fet_slab_self_consistent
    mat = fet_slab_self_consistent()
    x = np.asarray(mat["x"], float)
    phi0 = np.repeat((bias * x / x[-1])[:, None], mat["Ny"], axis=1)
    Vc = np.asarray(mat["Vc"], float)
    offsets = [0.005 * i for i in range(instances)]
    out = {"workload": f"FET slab 40x20 nm, 64x32 DG elements p=1 x 64x64 relative cells, n+ 1e24 m^-3 contacts, "
                       f"contact-mode reservoirs, {bias} V; {instances} transverse-mode instances (+5 meV each) in "
                       f"one batched Crank-Nicolson call, dt {dt:g} s x {steps} steps from zero, 2D Poisson every "
                       f"{poisson_every} steps, GMRES(30) + row scaling to 1e-8 per step",
           "instances": instances, "steps": steps}
    for prec in precisions:
        ops = []
        try:
            for off in offsets:
                op = F2.Full2DOperator(F2.Full2DSetup(mat, Vc - phi0 + off, 0.1, 0.1, host_reservoirs=False))
                op.upload_contact_rhs()
                op.set_preconditioner("rowabs")
                ops.append(op)
                print(f"[bench] self-consistent FET slab {prec}: instance {len(ops)} set up", file=sys.stderr,
                      flush=True)
            n = ops[0].n
            t0 = time.perf_counter()
            phi, hist = PS.transient_cn_batch_self_consistent(
                ops, offsets, phi0, bias, dt, steps, poisson_every=poisson_every, precision=prec, tol=1e-8,
                rho0=[np.zeros(n, complex)] * instances, use_graph=prec != "fp64",
                on_block=lambda h: print(f"[bench] self-consistent FET slab {prec}: {h['steps']} steps, "
                                         f"{time.perf_counter() - t0:.1f} s, Poisson update "
                                         f"{h['potentialUpdate']:.2e} V", file=sys.stderr, flush=True))
            wall = time.perf_counter() - t0
        finally:
            for op in ops:
                op.close()
        its = [sum(h["gmresIterations"]) for h in hist]
        out[prec] = {"instance_steps_per_s": instances * steps / wall, "seconds": wall,
                     "gmres_iterations_per_instance_step": sum(its) / (instances * steps),
                     "flag": max(h["flag"] for h in hist), "max_step_relres": max(h["maxStepRelres"] for h in hist),
                     "poisson_updates_V": [h["potentialUpdate"] for h in hist],
                     "poisson_relres": [h["poissonPhysicalRelres"] for h in hist]}
        print(f"[bench] self-consistent FET slab {prec}: {wall:.1f} s", file=sys.stderr, flush=True)
    return out


def transient_full2d_batch(instances: int = 8, steps: int = 1, dt: float = 1e-15, precision: str = "mixed") -> dict:
    """The full-2D FET slab transient with several transverse-mode instances
    (subband offsets of 5 meV each on the gate potential) stepped together by
    wdg_transient_cn_batch, against the same instances stepped one after
    another.  Reports instance-steps/s of both."""
    from real_space_dg_amd import full2d as F2
    from real_space_dg_amd.devices import fet_slab
    from real_space_dg_amd.solver import transient_cn_batch
    mat, V = fet_slab()
    setups = [F2.Full2DSetup(mat, V + 0.005 * i, 0.1, 0.05) for i in range(instances)]
    n = setups[0].n_total
    out = {"workload": f"FET slab 40x20 nm, 64x32 DG elements p=1 x 64x64 relative cells, {instances} transverse-mode "
                       f"instances (gate potential + 5 meV x mode); Crank-Nicolson dt {dt:g} s x {steps} steps from "
                       "zero, GMRES(30) + row scaling to 1e-8 per step", "unknowns_per_instance": n,
           "instances": instances, "precision": precision}
    ops = []
    try:
        for s in setups:
            op = F2.Full2DOperator(s)
            op.upload_rhs()
            op.set_preconditioner("rowabs")
            ops.append(op)
        zeros = [np.zeros(n, complex)] * instances
        transient_cn_batch(ops, dt, 1, tol=1e-8, precision=precision, rho0=zeros, return_rho=False)   # warm
        t0 = time.perf_counter()
        _, infos = transient_cn_batch(ops, dt, steps, tol=1e-8, precision=precision, rho0=zeros, return_rho=False)
        batched = time.perf_counter() - t0
        dens_b = [op.observables()[0] for op in ops]
        print(f"[bench] transient full2D batch of {instances}: {batched:.2f} s", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        its = []
        for op in ops:
            _, info = op.transient_cn(dt, steps, tol=1e-8, precision=precision, rho0=zeros[0], return_rho=False)
            its.append(info["iterations"])
        serial = time.perf_counter() - t0
        print(f"[bench] transient full2D serial {instances}: {serial:.2f} s", file=sys.stderr, flush=True)
        diff = max(float(np.max(np.abs(op.observables()[0] - d)) / np.max(np.abs(d))) for op, d in zip(ops, dens_b))
    finally:
        for op in ops:
            op.close()
    out["batched"] = {"instance_steps_per_s": instances * steps / batched, "seconds": batched,
                      "gmres_iterations_per_step": [i["iterations"] / max(i["steps"], 1) for i in infos],
                      "flags": [i["flag"] for i in infos]}
    out["serial"] = {"instance_steps_per_s": instances * steps / serial, "seconds": serial,
                     "gmres_iterations_per_step": [i / max(steps, 1) for i in its]}
    out["batched_over_serial"] = serial / batched
    out["batched_vs_serial_density_rel_diff"] = diff
    return out


def main():
    args = parse()
    global NO_GRAPHS
    NO_GRAPHS = args.no_graphs
    if args.gpus != 1:
        print(f"bench.py: runs on one GPU only (--gpus {args.gpus} requested); "
              "multi-GPU sharding is design-only.", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy > 0:
        copy_calibration(args.calibrate_copy)
    if args.only:
        fn = {"batched_sweep": batched_sweep, "self_consistent": self_consistent,
              "transient_full2d": lambda: transient_full2d(steps=args.full2d_steps,
                                                           precisions=tuple(args.full2d_precisions.split(",")),
                                                           dv=args.full2d_dv),
              "transient_full2d_batch": lambda: transient_full2d_batch(instances=args.full2d_instances,
                                                                       steps=args.full2d_steps),
              "transient_full2d_sc": lambda: transient_full2d_sc(instances=args.full2d_instances,
                                                                 steps=args.full2d_steps,
                                                                 precisions=tuple(args.full2d_precisions.split(",")))
              }[args.only]
        print(json.dumps({args.only: fn()}))
        return
    from real_space_dg_amd.devices import rtd_1d
    from real_space_dg_amd.solver import PROFILE_BITS, RTD_BENCH_TOL, SolveOptions, WignerOperator

    mat, E, Vs, efl, efr = rtd_1d(WORKLOAD["n_elem"], WORKLOAD["degree"], WORKLOAD["nk"],
                                  bias_v=WORKLOAD["bias_v"])
    m = mat["me_x"][0, :, 0] * M0
    m_nodes = np.concatenate([m[e * WORKLOAD["degree"]: e * WORKLOAD["degree"] + WORKLOAD["degree"] + 1]
                              for e in range(WORKLOAD["n_elem"])])
    op = WignerOperator.from_mat(mat)
    op.set_theta_method(args.theta)

    def step(prec: str, tol: float = RTD_BENCH_TOL):
        op.set_potential(mat["x"] * 1e-9, E[0, :, 0], efl, efr, m[0], m[-1], temp=mat["Temp"])
        _, info = op.solve(SolveOptions(tol=tol, restart=30, precision=prec, max_iter=20000,
                                        use_graph=not NO_GRAPHS), return_x=False)
        n, j, _ = op.observables(m_nodes)
        return info, n, j

    print(f"[bench] N={op.N} reference FP64 solve (tol 1e-12)", file=sys.stderr, flush=True)
    ref_info, n_ref, j_ref = step("fp64", 1e-12)
    for _ in range(args.warmup):
        step("mixed")
        step("fp64")
    all_mask = sum(1 << b for b in PROFILE_BITS.values())

    # Timed steps replay the captured restart-cycle graphs.  Per-kernel event
    # pairs cannot sit inside a replayed graph, so the kernel durations come
    # from a second, profiled pass over the same solves (launched one kernel
    # at a time with an event pair around each).
    op.set_profiling(0)
    t0 = time.perf_counter()
    mixed = [step("mixed") for _ in range(args.steps)]
    t_mixed = time.perf_counter() - t0
    print(f"[bench] mixed: {args.steps} solves in {t_mixed:.3f} s", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    fp64 = [step("fp64") for _ in range(args.fp64_steps)]
    t_fp64 = time.perf_counter() - t0
    print(f"[bench] fp64: {args.fp64_steps} solves in {t_fp64:.3f} s", file=sys.stderr, flush=True)

    op.set_profiling(all_mask)
    t0 = time.perf_counter()
    for _ in range(args.profile_steps):
        step("mixed")
    t_prof_mixed = (time.perf_counter() - t0) / args.profile_steps
    prof_mixed = op.profile()
    op.set_profiling(all_mask)
    t0 = time.perf_counter()
    for _ in range(args.profile_steps):
        step("fp64")
    t_prof_fp64 = (time.perf_counter() - t0) / args.profile_steps
    prof_fp64 = op.profile()
    op.set_profiling(0)
    print(f"[bench] profiled pass: {args.profile_steps} solves per precision", file=sys.stderr, flush=True)
    # the launch time of the operator apply (in the FFT form one theta_fft
    # launch) from back-to-back applies: the profiled pass brackets every
    # kernel with its own event pair while the host submits kernel by kernel,
    # which also times the host's submission gaps
    fft_form = prof_mixed["theta_fft"]["launches"] > 0 and prof_fp64["theta_fft"]["launches"] > 0
    t_apply = {p: op.time_apply(p, 50) for p in ("mixed", "fp64")} if fft_form else {"mixed": None, "fp64": None}

    v_mixed = args.steps / t_mixed
    v_fp64 = args.fp64_steps / t_fp64
    info_m, n_m, j_m = mixed[-1]
    info_d, n_d, j_d = fp64[-1]
    scale_n = np.max(np.abs(n_ref))
    scale_j = np.max(np.abs(j_ref))
    err = {
        "mixed_density_rel_err_vs_fp64_1e-12": float(np.max(np.abs(n_m - n_ref)) / scale_n),
        "mixed_current_rel_err_vs_fp64_1e-12": float(np.max(np.abs(j_m - j_ref)) / scale_j),
        "fp64_density_rel_err_vs_fp64_1e-12": float(np.max(np.abs(n_d - n_ref)) / scale_n),
        "fp64_current_rel_err_vs_fp64_1e-12": float(np.max(np.abs(j_d - j_ref)) / scale_j),
    }
    # against the oracle's FP64 solution of the same workload (GMRES to 1e-13
    # on the CPU, tests/golden/make_rtd_c2_observables.py): max-norm errors
    # relative to the largest |n| and |j|
    gold_path = ROOT / "tests" / "golden" / "rtd_c2_observables.npz"
    if gold_path.exists():
        gold = np.load(gold_path, allow_pickle=False)
        gn, gj = gold["n"], gold["j"]
        for name, (nn, jj) in {"mixed": (n_m, j_m), "fp64": (n_d, j_d), "fp64_1e-12": (n_ref, j_ref)}.items():
            err[f"{name}_density_rel_err_vs_oracle"] = float(np.max(np.abs(nn - gn)) / np.max(np.abs(gn)))
            err[f"{name}_current_rel_err_vs_oracle"] = float(np.max(np.abs(jj - gj)) / np.max(np.abs(gj)))
        err["oracle"] = f"FP64 numpy oracle, GMRES(30) + line preconditioner, true relres {float(gold['true_relres']):.1e}"
    # the transport apply alone (the HBM-bound part of the Wigner apply:
    # transport + CAP diagonal, 2 s N bytes), through wdg_apply_parts
    u_probe = np.random.default_rng(0).standard_normal(op.N) + 1j * np.random.default_rng(1).standard_normal(op.N)
    transport_roof = {}
    for prec in ("mixed", "fp64"):
        op.set_profiling(1 << PROFILE_BITS["transport"])
        for _ in range(20):
            op.apply(u_probe, prec, parts="transport")
        transport_roof[prec] = kernel_roofline(op.profile(), "transport", "fp32", prec)
        op.set_profiling(0)
    tot = {k: v["avg_ms"] * v["launches"] for k, v in prof_mixed.items() if v["launches"]}
    dominant = max(tot, key=tot.get)
    roof = kernel_roofline(prof_mixed, dominant, "bf16",
                           t_launch_ms=t_apply["mixed"] if dominant == "theta_fft" else None)
    kernels = {k: {"avg_ms": v["avg_ms"], "launches": v["launches"],
                   "share": tot.get(k, 0.0) / max(sum(tot.values()), 1e-30)} for k, v in prof_mixed.items()}
    line = {
        "metric": METRIC,
        "value": v_mixed,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_mixed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "mixed: fp32 inner GMRES (fp16-stored Krylov basis), fp64 residual/refinement",
        "data": "synthetic RTD potential (deterministic), no learned weights",
        "config": {"workload": "RTD 60 nm, 0.2 V bias, 512 DG elements p=3 x 1024 k, GMRES(30)+line "
                               f"preconditioner to true relres {RTD_BENCH_TOL:g} (<= 1e-10, tight enough "
                               "that n and j are within 1e-8 of the FP64 oracle)",
                   "unknowns": op.N, "restart": 30, "tol": RTD_BENCH_TOL},
        "fp64_solves_per_s": v_fp64,
        "fp64_ms_per_step": 1e3 * t_fp64 / args.fp64_steps,
        "mixed_over_fp64": v_mixed / v_fp64,
        "iterations": {"mixed": info_m["iterations"], "mixed_refinements": info_m["outer_iterations"],
                       "fp64": info_d["iterations"], "fp64_ref_1e-12": ref_info["iterations"]},
        "true_relres": {"mixed": info_m["true_relres"], "fp64": info_d["true_relres"]},
        "errors": err,
        "roofline": roof,
        # the chirp-z kernel is the whole operator apply (transport rows fused in)
        "wigner_apply_roofline": kernel_roofline(prof_mixed, "theta_fft", "fp32", t_launch_ms=t_apply["mixed"]),
        "wigner_apply_roofline_fp64": kernel_roofline(prof_fp64, "theta_fft", "fp64", "fp64",
                                                      t_launch_ms=t_apply["fp64"]),
        # the transport apply alone, 2 s N bytes per launch (wdg_apply_parts, transport part)
        "transport_apply_roofline": transport_roof,
        "kernels_mixed": kernels,
        "theta_block_fp32": op.fft_block,   # threads per workgroup of the FP32 chirp-z kernel (autotuned)
        "kernel_timing": (f"HIP event pairs on the context stream around every kernel, over a profiled pass of "
                          f"{args.profile_steps} further solves per precision; that pass launches kernel by "
                          f"kernel (no graph replay): {1e3 * t_prof_mixed:.2f} ms/solve mixed, "
                          f"{1e3 * t_prof_fp64:.2f} ms/solve FP64"),
    }
    op.close()
    if not args.no_extras:
        line["batched_sweep"] = batched_sweep()
        line["self_consistent"] = self_consistent()
        line["transient_full2d"] = transient_full2d(steps=args.full2d_steps)
    if not args.no_cpu_baseline:
        print("[bench] CPU baseline sample", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(mat, E, efl, efr, info_d["iterations"], args.cpu_sample_iters)
        line["c1_timed_solves"] = c1_timed_solves(line["cpu_baseline"]["cores"])
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line))


if __name__ == "__main__":
    main()
