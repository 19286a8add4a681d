"""FP64 CPU restatement of the reference's explicit RK4 transient (1D path).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

  DG_code/core/transient.m:12-16 and
  DG_code/solve_transport_DG_2D_transient_SC_new.m:145-158   -> rk4

    k1 = rhs - A rho;  k2 = rhs - A(rho + dt/2 k1);  k3 = rhs - A(rho + dt/2 k2)
    k4 = rhs - A(rho + dt k3);  rho += dt/6 (k1 + 2 k2 + 2 k3 + k4)

with A prescaled by a_scale (2e-10 in the reference, :148) and rhs by
b_scale (1 there).  Observables after each saved step as
solve_transport_DG_2D_transient_SC_new.m:172-178 (n_xi even branch).

Also the Crank-Nicolson (trapezoidal) step of the same ODE,
(I + dt/2 a A) rho+ = (I - dt/2 a A) rho + dt b rhs            -> crank_nicolson
which has no counterpart in the reference (its full-2D transient entry points
are stubs, DG_code/solve_transport_DG_full2D_transient_flatband.m:8-11), and
the same step for operators too large to assemble, each step solved by the
oracle's GMRES to a tight tolerance                     -> crank_nicolson_mf

Pinning (tests/test_oracle_transient.py): against the exact solution of the
linear ODE, rho(t) = rho_inf + expm(-a A t)(rho0 - rho_inf), with the global
error falling as dt^4 (RK4) and dt^2 (Crank-Nicolson); the stationary state
is a fixed point of both; on a diagonal A each Crank-Nicolson step multiplies
a mode by (1 - z/2)/(1 + z/2), z = dt a lambda; crank_nicolson_mf equals
crank_nicolson on an assembled operator.
"""
from __future__ import annotations

import numpy as np

from . import wigner1d as W


def rk4(A, rhs, rho0, dt, nsteps, a_scale=2e-10, b_scale=1.0, save_every=1, p=None, m_nodes=None):
    """Returns (rho, n_t, j_t); n_t, j_t are (nodes, nsteps // save_every)
    when p and m_nodes are given, else None."""
    As = A * a_scale
    b = rhs * b_scale
    rho = np.asarray(rho0, complex).copy()
    n_t, j_t = [], []
    for step in range(1, nsteps + 1):
        k1 = b - As @ rho
        k2 = b - As @ (rho + dt / 2 * k1)
        k3 = b - As @ (rho + dt / 2 * k2)
        k4 = b - As @ (rho + dt * k3)
        rho = rho + dt / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
        if p is not None and step % save_every == 0:
            n, j, _ = W.observables(p, rho, m_nodes)
            n_t.append(n)
            j_t.append(j)
    if p is None:
        return rho, None, None
    return rho, np.array(n_t).T, np.array(j_t).T


def crank_nicolson(A, rhs, rho0, dt, nsteps, a_scale=1.0, b_scale=1.0):
    """nsteps trapezoidal steps of d rho/dt = b_scale rhs - a_scale A rho,
    each solved directly (sparse LU of I + dt/2 a A, factored once)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    n = A.shape[0]
    As = sp.csc_matrix(A) * a_scale
    I = sp.identity(n, format="csc")
    lu = spla.splu((I + 0.5 * dt * As).tocsc())
    minus = (I - 0.5 * dt * As).tocsr()
    g = dt * b_scale * np.asarray(rhs, complex)
    rho = np.asarray(rho0, complex).copy()
    for _ in range(nsteps):
        rho = lu.solve(minus @ rho + g)
    return rho



def crank_nicolson_mf(apply_a, rhs, rho0, dt, nsteps, a_scale=1.0, b_scale=1.0, tol=1e-13, restart=60,
                      row_scale=None):
    """crank_nicolson with A given by its apply: each step's system
    (I + dt/2 a A) x = (I - dt/2 a A) rho + dt b rhs solved by the oracle's
    right-preconditioned GMRES to true relative residual tol, with the row
    scaling 1 / (1 + dt/2 a row_scale) when row_scale (|A| row sums) is given."""
    from .krylov import gmres
    h = 0.5 * dt * a_scale
    g = dt * b_scale * np.asarray(rhs, complex)
    rho = np.asarray(rho0, complex).copy()
    if row_scale is None:
        apply_m = lambda v: v  # noqa: E731
    else:
        inv = 1.0 / (1.0 + h * np.asarray(row_scale, float))
        apply_m = lambda v: inv * v  # noqa: E731
    shifted = lambda v: v + h * apply_a(v)  # noqa: E731
    for _ in range(nsteps):
        b = rho - h * apply_a(rho) + g
        rho, info = gmres(shifted, apply_m, b, restart=restart, tol=tol, maxit=200000, x0=rho)
        if info["flag"] != 0:
            raise RuntimeError(f"crank_nicolson_mf: step solve stopped at relres {info['true_relres']:.2e}")
    return rho
