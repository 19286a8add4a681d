"""FP64 CPU restatement of the Krylov solve: line preconditioner + GMRES.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Follows DG_code/solve_transport_DG_full2D.m:276-334 and :462-484 (RHS
normalisation, right preconditioning, restarted GMRES, true-residual check)
and the x-line factor of DG_code/core/get_LineTransversePreconditioner_full2D.m
:174-220 restated in the k basis (the rho_x eigen-transform is the identity
there because the velocity is already diagonal): per wave number k an exact
upwind block-bidiagonal solve along x of A_diff + diag_k(G).

The diagonal-in-k part of G is computed here straight from its definition
G[(e,a,j),(e,b,j)] = i * sum_{l,l'} conj(phi_lj) Cbar_ab(l,l') phi_l'j, with the
actual phi matrix.

Pinning: the preconditioner is checked to invert the sparse operator
A_diff + diag_k(G) assembled from oracle.wigner1d (tests/test_oracle_krylov.py);
GMRES is checked against the direct sparse solve.  The GMRES iteration
trajectory (MATLAB's gmres internals) is parity unpinned; only converged
solutions are compared.
"""
from __future__ import annotations

import numpy as np

from .wigner1d import MatrixFreeOperator


class LinePreconditioner:
    def __init__(self, op: MatrixFreeOperator, Bbar=None):
        p = op.p
        self.p = p
        n, Ne, Nk = p["N_K_chi"], p["N_chi"], p["Nk"]
        self.n, self.Ne, self.Nk = n, Ne, Nk
        if Bbar is None:
            Bbar = op.edge_coefficients()
        phi = p["phi"]
        dg = Bbar[..., :-1] + Bbar[..., 1:]                    # (e,a,b,l) diagonal of Cbar
        off = Bbar[..., 1:-1]                                  # (e,a,b,l) off diagonal
        ph2 = np.abs(phi) ** 2
        cross = np.conj(phi[:-1, :]) * phi[1:, :]
        lead = dg.shape[:3]
        Gd = 1j * ((dg.reshape(-1, dg.shape[-1]) @ ph2) +
                   (off.reshape(-1, off.shape[-1]) @ cross) +
                   (off.reshape(-1, off.shape[-1]) @ np.conj(cross))).reshape(lead + (Nk,))
        b = op.blocks
        k = p["k"]
        Q = p["Q_diff"] / (p["delta_chi"] / 2)
        D = Q * (k[None, None, :] * b["K1"][:, :, None] + np.abs(k)[None, None, :] * (b["K2L"] + b["K2R"])[:, :, None])
        blocks = np.transpose(D[None] + Gd, (0, 3, 1, 2))      # (e, j, a, b)
        self.blocks = blocks
        self.CL = Q * (k[:, None, None] * b["K3"][None] + np.abs(k)[:, None, None] * b["K4"][None])  # from e-1
        self.CR = Q * (k[:, None, None] * b["K5"][None] + np.abs(k)[:, None, None] * b["K6"][None])  # from e+1
        self.pos = k > 0

    def apply(self, r):
        n, Ne, Nk = self.n, self.Ne, self.Nk
        R = np.asarray(r).reshape(Ne, n, Nk)
        Z = np.zeros_like(R, dtype=complex)
        pos, neg = self.pos, ~self.pos
        for e in range(Ne):                     # forward sweep, k > 0
            rr = R[e][:, pos].T.copy()          # (k, a)
            if e > 0:
                rr -= np.einsum("jab,jb->ja", self.CL[pos], Z[e - 1][:, pos].T)
            Z[e][:, pos] = np.linalg.solve(self.blocks[e][pos], rr[..., None])[..., 0].T
        for e in range(Ne - 1, -1, -1):         # backward sweep, k < 0
            rr = R[e][:, neg].T.copy()
            if e < Ne - 1:
                rr -= np.einsum("jab,jb->ja", self.CR[neg], Z[e + 1][:, neg].T)
            Z[e][:, neg] = np.linalg.solve(self.blocks[e][neg], rr[..., None])[..., 0].T
        return Z.reshape(-1)


def gmres(apply_a, apply_m, b, restart=30, tol=1e-10, maxit=10000, x0=None, callback=None):
    """Right-preconditioned restarted GMRES (CGS2 Arnoldi, complex Givens).

    Stops when the true residual ||b - A x|| <= tol ||b|| (checked at every
    restart, as the reference overrides the flag with the true residual,
    DG_code/solve_transport_DG_full2D.m:314-320).  Returns (x, info).
    """
    b = np.asarray(b, complex)
    bn = np.linalg.norm(b)
    x = np.zeros_like(b) if x0 is None else np.array(x0, complex)
    its = 0
    resvec = []
    while True:
        r = b - apply_a(x)
        beta = np.linalg.norm(r)
        resvec.append(beta / bn)
        if beta <= tol * bn or its >= maxit:
            break
        m = restart
        V = np.zeros((m + 1, b.size), complex)
        H = np.zeros((m + 1, m), complex)
        cs = np.zeros(m, complex)
        sn = np.zeros(m, complex)
        g = np.zeros(m + 1, complex)
        g[0] = beta
        V[0] = r / beta
        jl = 0
        for j in range(m):
            w = apply_a(apply_m(V[j]))
            for _ in range(2):
                h = V[:j + 1].conj() @ w
                w = w - h @ V[:j + 1]
                H[:j + 1, j] += h
            H[j + 1, j] = np.linalg.norm(w)
            if H[j + 1, j] > 0:
                V[j + 1] = w / H[j + 1, j]
            for i in range(j):
                t = cs[i] * H[i, j] + sn[i] * H[i + 1, j]
                H[i + 1, j] = -np.conj(sn[i]) * H[i, j] + np.conj(cs[i]) * H[i + 1, j]
                H[i, j] = t
            a, bb = H[j, j], H[j + 1, j]
            rr = np.sqrt(abs(a) ** 2 + abs(bb) ** 2)
            cs[j] = np.conj(a) / rr if rr else 1.0
            sn[j] = np.conj(bb) / rr if rr else 0.0
            H[j, j] = rr
            H[j + 1, j] = 0
            g[j + 1] = -np.conj(sn[j]) * g[j]
            g[j] = cs[j] * g[j]
            its += 1
            jl = j + 1
            if callback:
                callback(its)
            if abs(g[j + 1]) <= tol * bn or its >= maxit or rr == 0:
                break
        y = np.linalg.solve(np.triu(H[:jl, :jl]), g[:jl])
        x = x + apply_m(y @ V[:jl])
    return x, {"iterations": its, "true_relres": resvec[-1], "flag": int(resvec[-1] > tol), "resvec": resvec}
