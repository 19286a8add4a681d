"""Self-consistent RTD: Gummel loop of the Wigner transport and FP64 Poisson.

Host-side mirror of RTD/solve2D_rtd_self_consistent.m:88-205 for the 1D
(Ny = 1, Dirichlet-X) case.  Per bias point and Gummel iteration:

  V = Vc - phi                                   (:108)
  EfL = Vc(1) + 0.05, EfR = Vc(end) + 0.05 - bias (:109-110)
  transport solve, warm-started from the previous iterate (:116, :185)
  candidate = Poisson step on the projected transport density (:145)
  converged if max|candidate - phi| <= potentialTolerance, the relative
  density change <= densityTolerance and the Poisson residual at the
  current potential <= poissonResidualTolerance (:175-186)
  otherwise phi += mixing * (candidate - phi) (:186)

``acceleration='anderson'`` replaces the damped update by Anderson mixing
of depth ``anderson_depth`` on the same fixed-point map phi -> candidate
(the convergence test is unchanged: it still uses the unmixed defect).

Unlike the reference, which sweeps the biases one after another with
continuation, all bias points of ``mat['Vdd']`` iterate together: every
Gummel iteration solves the transport of all unconverged biases in one
batched call (wdg_solve_batch) and runs their FP64 Poisson updates on the
device (wdg_poisson_step).  The transport stays on the device between
iterations (warm start from the device-resident solution).
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from .solver import SolveOptions, WignerOperator, _node_values, solve_batch

M0 = 9.109e-31


@dataclass
class GummelOptions:
    """Defaults of RTD/solve2D_rtd_self_consistent.m:228-231."""
    max_iterations: int = 100
    mixing: float = 0.1
    potential_tolerance: float = 5e-4          # [V], unmixed defect
    density_tolerance: float = 5e-4            # relative density change
    poisson_residual_tolerance: float = 8e-2   # relative Poisson residual
    poisson_method: str = "newton-raphson"
    acceleration: str = "none"                 # 'none' (damped) or 'anderson'
    anderson_depth: int = 3
    transport: SolveOptions = field(default_factory=lambda: SolveOptions(
        tol=1e-9, precision="mixed", restart=30, restart_max=64))


class _Anderson:
    """Type-II Anderson mixing on the residual f = G(phi) - phi."""

    def __init__(self, depth: int, beta: float):
        self.depth, self.beta = depth, beta
        self.xs, self.fs = [], []

    def update(self, phi, f):
        self.xs.append(phi.copy())
        self.fs.append(f.copy())
        self.xs = self.xs[-(self.depth + 1):]
        self.fs = self.fs[-(self.depth + 1):]
        if len(self.fs) == 1:
            return phi + self.beta * f
        dF = np.stack([b - a for a, b in zip(self.fs[:-1], self.fs[1:])], axis=1)
        dX = np.stack([b - a for a, b in zip(self.xs[:-1], self.xs[1:])], axis=1)
        gamma, *_ = np.linalg.lstsq(dF, f, rcond=None)
        return phi + f - (dX + dF) @ gamma


def solve_rtd_self_consistent(mat: dict, options: GummelOptions | None = None, verbose: bool = False):
    """Batched Gummel loop over the bias points ``mat['Vdd']``.

    Returns a dict with per-bias ``current`` (mean j over the DG nodes,
    A/m^2), ``density`` (projected on the material grid, m^-3), ``phi`` [V],
    ``potential`` (Vc - phi, eV), ``iterations``, ``converged``, ``status``,
    ``history`` (per-iteration diagnostics) and the wall time.
    """
    opt = options or GummelOptions()
    if opt.acceleration not in ("none", "anderson"):
        raise ValueError("acceleration must be 'none' or 'anderson'")
    if not (0 < opt.mixing <= 1):
        raise ValueError("mixing must lie in (0, 1]")
    t0 = time.perf_counter()
    biases = np.asarray(mat["Vdd"], float)
    nb = biases.size
    x_nm = np.asarray(mat["x"], float)
    x_m = x_nm * 1e-9
    vc = np.asarray(mat["Vc"], float)
    L = x_nm[-1]
    ramp_len = L - mat["Lc1"] - mat["Lc2"]
    if ramp_len <= 0:
        raise ValueError("RTD:ContactGeometry: contact regions must leave a positive central length")
    ramp = np.clip((x_nm - mat["Lc1"]) / ramp_len, 0.0, 1.0)           # :49
    off = float(mat.get("fermi_offset", 0.05))
    me = np.asarray(mat["me_x"], float)[0, :, 0] * M0
    temp = float(mat["Temp"])
    ops = []
    for _ in range(nb):
        op = WignerOperator.from_mat(mat)
        op.poisson_setup(x_m, mat["eps"], np.asarray(mat["Nd"], float) - np.asarray(mat.get("Na", 0.0 * vc)),
                         temp=temp, method=opt.poisson_method)
        ops.append(op)
    m_nodes = _node_values(ops[0].n_chi, ops[0].n_k_chi, me)
    phi = [b * ramp for b in biases]                                    # :91
    active = list(range(nb))
    iters = np.zeros(nb, int)
    converged = np.zeros(nb, bool)
    status = ["iteration-limit"] * nb
    history = [[] for _ in range(nb)]
    prev_n = [None] * nb
    current = np.zeros(nb)
    dens = [None] * nb
    j_last = [None] * nb
    gmres_its = np.zeros(nb, int)
    accel = [_Anderson(opt.anderson_depth, opt.mixing) for _ in range(nb)]
    for it in range(1, opt.max_iterations + 1):
        if not active:
            break
        for ib in active:
            V = vc - phi[ib]
            ops[ib].set_potential(x_m, V, vc[0] + off, vc[-1] + off - biases[ib], me[0], me[-1], temp=temp,
                                  fermi_model="TSU_ESAKI")
        topts = SolveOptions(**{**opt.transport.__dict__, "warm_device": it > 1})
        infos = solve_batch([ops[ib] for ib in active], topts)
        still = []
        for ib, info in zip(active, infos):
            iters[ib] = it
            gmres_its[ib] += info["iterations"]
            h = {"iteration": it, "transportFlag": info["flag"], "transportRelres": info["true_relres"],
                 "gmresIterations": info["iterations"], "warmStartUsed": it > 1}
            if info["flag"] != 0:                                       # :123-126
                status[ib] = "transport-not-converged"
                history[ib].append(h)
                continue
            n_dg, j_dg, _ = ops[ib].observables(m_nodes)
            if prev_n[ib] is not None:
                h["densityChange"] = float(np.linalg.norm(n_dg - prev_n[ib]) /
                                           max(np.linalg.norm(n_dg), np.linalg.norm(prev_n[ib]), 1e-300))
            else:
                h["densityChange"] = np.inf
            cand, n_cv, pinfo = ops[ib].poisson_step(phi[ib], biases[ib])
            h["potentialDefect"] = float(np.max(np.abs(cand - phi[ib])))
            h["poissonPhysicalRelres"] = pinfo["physical_relres_before"]
            h["chargeConservation"] = abs(pinfo["charge_integral"] - pinfo["dg_integral"]) / max(
                abs(pinfo["dg_integral"]), 1e-300)
            history[ib].append(h)
            dens[ib] = n_cv
            j_last[ib] = j_dg
            current[ib] = float(np.mean(j_dg))
            if verbose:
                print(f"[gummel] bias {biases[ib]:.3f} it {it} gmres {info['iterations']} "
                      f"defect {h['potentialDefect']:.2e} dn {h['densityChange']:.2e} "
                      f"res {h['poissonPhysicalRelres']:.2e}", flush=True)
            if (h["potentialDefect"] <= opt.potential_tolerance and h["densityChange"] <= opt.density_tolerance
                    and h["poissonPhysicalRelres"] <= opt.poisson_residual_tolerance):
                converged[ib] = True
                status[ib] = "fixed-point-converged"
                continue
            prev_n[ib] = n_dg
            if opt.acceleration == "anderson":
                phi[ib] = accel[ib].update(phi[ib], cand - phi[ib])
                phi[ib][0], phi[ib][-1] = 0.0, biases[ib]
            else:
                phi[ib] = phi[ib] + opt.mixing * (cand - phi[ib])       # :186
            still.append(ib)
        active = still
    for op in ops:
        op.close()
    return {"bias": biases, "current": current, "density": dens, "phi": phi,
            "potential": [vc - p for p in phi], "j": j_last, "iterations": iters, "converged": converged,
            "status": status, "history": history, "gmres_iterations": gmres_its,
            "seconds": time.perf_counter() - t0}
