"""FP64 CPU restatement of the RTD Poisson step and DG-density projection.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates, for the 1D (Ny = 1, Dirichlet-X) case:
  RTD/eval_Poisson2D_rtd.m:1-200      -> poisson_matrix   (harmonic-mean eps, identity contact rows)
  RTD/solve_Poisson2D_rtd.m:85-130    -> poisson_step     ('direct' / 'newton-raphson')
  RTD/project_DG_density_RTD.m:1-69   -> projection_matrix, project_density

Pinning (tests/test_oracle_poisson.py): the projection integrates each
element's Lagrange polynomial exactly (checked on polynomial densities and
by conservation of the total charge); the Poisson matrix reproduces the
exact solution of eps phi'' = const on a uniform grid (quadratic phi) and
the harmonic-mean flux continuity across a dielectric step (the exact
two-layer solution satisfies the discrete system; an interface off the face
does not); iterated on a Boltzmann density the Newton step converges
quadratically and leaves a vanishing nonlinear Poisson residual.  The
Gummel iteration itself (bias loop, mixing, tolerances) is restated in
real_space_dg_amd/gummel.py and checked on the GPU against this step.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

EPS0 = 8.8541878128e-12      # RTD/eval_Poisson2D_rtd.m:25
Q = 1.602176634e-19          # RTD/solve_Poisson2D_rtd.m:24
KB = 1.380649e-23            # RTD/solve_Poisson2D_rtd.m:25


def poisson_matrix(eps_r, dx_nm):
    """1D div(eps0 eps_r grad) with Dirichlet contact rows (identity).

    RTD/eval_Poisson2D_rtd.m:107-186 with Ny = 1: interior row k couples
    k-1, k+1 with cx * harmonic_mean(eps_k, eps_nb), cx = eps0*1e18/dx^2.
    """
    er = np.asarray(eps_r, float)
    nx = er.size
    cx = EPS0 * 1e18 / dx_nm ** 2
    rows, cols, vals = [], [], []
    for ix in range(nx):
        if ix == 0 or ix == nx - 1:
            rows.append(ix); cols.append(ix); vals.append(1.0)
            continue
        diag = 0.0
        for nb in (ix - 1, ix + 1):
            c = cx * (2 * er[ix] * er[nb] / (er[ix] + er[nb]))
            rows.append(ix); cols.append(nb); vals.append(c)
            diag -= c
        rows.append(ix); cols.append(ix); vals.append(diag)
    return sp.csr_matrix((vals, (rows, cols)), shape=(nx, nx))


def projection_matrix(n_chi, r_nodes, x_mat_m, l_chi):
    """T (n_mat x n_chi*n) and dual-cell widths: exact integrals of each
    element's Lagrange basis over the Poisson control volumes
    [mid(x_{i-1},x_i), mid(x_i,x_{i+1})].  RTD/project_DG_density_RTD.m:39-69.
    """
    r = np.asarray(r_nodes, float)
    n = r.size
    nodes = np.asarray(x_mat_m, float)
    edges = np.concatenate([[nodes[0]], (nodes[:-1] + nodes[1:]) / 2, [nodes[-1]]])
    widths = np.diff(edges)
    vert = np.linspace(0.0, l_chi, n_chi + 1)
    jac = (vert[1:] - vert[:-1]) / 2
    nq = int(np.ceil(n / 2))
    pts, wts = np.polynomial.legendre.leggauss(nq)
    T = np.zeros((nodes.size, n_chi * n))
    for e in range(n_chi):
        left, right = vert[e], vert[e + 1]
        rows = np.nonzero((edges[:-1] < right) & (edges[1:] > left))[0]
        for row in rows:
            a = max(edges[row], left)
            b = min(edges[row + 1], right)
            sample = ((a + b) / 2 + (b - a) / 2 * pts - (left + right) / 2) / jac[e]
            basis = np.ones((nq, n))
            for j in range(n):
                for k in range(n):
                    if k != j:
                        basis[:, j] *= (sample - r[k]) / (r[j] - r[k])
            T[row, e * n:(e + 1) * n] = (b - a) / 2 * (wts @ basis)
    return T, widths


def project_density(T, widths, n_dg):
    return (T @ np.asarray(n_dg, float)) / widths


def poisson_step(M, n, net_doping, phi0, bias, temp, method="newton-raphson", alpha=1.0):
    """One electrostatic update.  RTD/solve_Poisson2D_rtd.m:85-130 (Ny = 1).

    M phi = q (n - Nd + Na); Dirichlet phi = 0 (source) and bias (drain).
    'direct': delta = -Aff^-1 F; 'newton-raphson': J = Aff - diag(q^2 n/(kB T)).
    Returns (phi_new, info).
    """
    nx = M.shape[0]
    bc = np.zeros(nx)
    bc[-1] = bias
    fixed = np.array([0, nx - 1])
    free = np.arange(1, nx - 1)
    charge = Q * (np.asarray(n, float) - np.asarray(net_doping, float))
    v = np.asarray(phi0, float).copy()
    v[fixed] = bc[fixed]
    F = M @ v - charge
    Aff = M[free][:, free]
    if method == "direct":
        J = Aff
    else:
        J = Aff - sp.diags(Q ** 2 * np.asarray(n, float)[free] / (KB * temp))
    delta = -spla.spsolve(J.tocsc(), F[free])
    out = v.copy()
    out[free] = v[free] + alpha * delta
    scale = max(np.max(np.abs(Q * np.asarray(n, float)[free])), np.max(np.abs(Q * np.asarray(net_doping, float)[free])),
                np.max(np.abs((M @ v)[free])), Q)
    info = {"physical_relres_before": float(np.max(np.abs(F[free])) / scale),
            "unmixed_step": float(np.max(np.abs(delta)))}
    return out, info
