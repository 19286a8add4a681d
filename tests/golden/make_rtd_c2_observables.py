"""Write tests/golden/rtd_c2_observables.npz: density n(x) and current j(x) of
the 512-element (p = 3) x 1024-k RTD at 0.2 V bias, solved by the FP64 oracle.

Workload: 60 nm double-barrier RTD (0.3 eV x 3 nm barriers around a 5 nm
well), m = 0.067 m0, T = 300 K, E_F = 0.1 eV, linear 0.2 V drop; N =
512*4*1024 = 2,097,152 unknowns.  Solver: the oracle's matrix-free operator
(oracle.wigner1d.MatrixFreeOperator: transport of DG_code/core/get_Diff.m and
Theta[V] of get_Drift_accel_trial_permutedShell.m as two dense phi products)
with the exact per-k line preconditioner and right-preconditioned GMRES(30)
(oracle.krylov) to a true relative residual of 1e-13; observables as
DG_code/solve_transport_DG_2D.m:114-156 (oracle.wigner1d.observables).

Runs in this container in about ten minutes (numpy, all cores):
    python tests/golden/make_rtd_c2_observables.py
"""
from __future__ import annotations

import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import krylov as K  # noqa: E402
from oracle import wigner1d as W  # noqa: E402
from real_space_dg_amd.devices import rtd_1d  # noqa: E402

M0 = 9.109e-31
N_ELEM, DEGREE, NK, BIAS = 512, 3, 1024, 0.2
TOL = 1e-13


def main():
    mat, E, _, efl, efr = rtd_1d(N_ELEM, DEGREE, NK, bias_v=BIAS)
    m = mat["me_x"][0, :, 0] * M0
    t0 = time.perf_counter()
    p = W.init_params(mat, efl, efr, E[0, :, 0], m, 1.0)
    op = W.MatrixFreeOperator.build(mat, p, E[0, :, 0])
    bbar = op.edge_coefficients()
    pre = K.LinePreconditioner(op, bbar)
    _, rhs = W.get_diff(p)
    print(f"setup {time.perf_counter() - t0:.1f} s, N = {rhs.size}", flush=True)

    def progress(its):
        if its % 10 == 0:
            print(f"  iteration {its} ({time.perf_counter() - t0:.0f} s)", flush=True)

    x, info = K.gmres(lambda u: op.apply(u, bbar), pre.apply, rhs, restart=30, tol=TOL, maxit=3000,
                      callback=progress)
    res = np.linalg.norm(rhs - op.apply(x, bbar)) / np.linalg.norm(rhs)
    print(f"GMRES: {info['iterations']} iterations, true relres {res:.3e}", flush=True)
    m_nodes = np.concatenate([m[e * DEGREE: e * DEGREE + DEGREE + 1] for e in range(N_ELEM)])
    n, j, _ = W.observables(p, x, m_nodes)
    out = Path(__file__).with_name("rtd_c2_observables.npz")
    np.savez_compressed(out, n=n, j=j, true_relres=res, iterations=info["iterations"],
                        workload=np.array(f"rtd_1d({N_ELEM}, {DEGREE}, {NK}, bias_v={BIAS}); "
                                          f"oracle GMRES(30) + line preconditioner, tol {TOL}"))
    print(f"wrote {out}", flush=True)


if __name__ == "__main__":
    main()
