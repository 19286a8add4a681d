"""FP64 CPU restatement of the reference 1D-x x 1D-k stationary Wigner solve.

TEST INFRASTRUCTURE ONLY -- the checker, never the product.  Only tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg may import it.

Restates (FV-in-xi / plane-wave-k branch, ``permute_shell = true``):

  DG_code/core/JacobiGQ.m, JacobiGL.m, JacobiP.m, Vandermonde1D.m,
  GradJacobiP.m, GradVandermonde1D.m, Dmatrix1D.m    -> jacobi_* / vandermonde1d / dmatrix1d
  DG_code/initParams.m:58-221                        -> init_params
  DG_code/core/Basisfunktionen_Q_m.m:3-33            -> (inside init_params)
  DG_code/core/verteilung.m:77-170                   -> verteilung
  DG_code/core/get_Diff.m:1-83                       -> get_diff
  DG_code/core/V_transient.m, get_CAP.m              -> v_transient / get_cap
  DG_code/core/get_Drift_accel_trial_permutedShell.m -> drift_samples, get_drift,
                                                        get_drift_literal
  DG_code/core/get_Drift.m:128-153 (loop form)       -> get_drift
  DG_code/core/get_SysM.m:1-16                       -> get_sysm
  DG_code/solve_transport_DG_2D.m:1-225              -> solve_transport_dg_2d

Pinning (what checks each function beyond retyping; see tests/test_oracle_*.py):
  jacobi_gl / jacobi_gq       closed forms: GLL nodes are +-1 and roots of P'_N,
                              weights 2/(N(N+1)P_N^2), sum w = 2, exactness on
                              polynomials of degree 2N-1; n=3 nodes {-1,0,1}.
  vandermonde1d / dmatrix1d   orthonormality under exact quadrature; D*1 = 0;
                              D differentiates polynomials of degree n-1 exactly;
                              SBP identity M*D + (M*D)^T = diag(-1,0..0,1).
  plane-wave basis            closed form e^{ik xi}/sqrt(L); k grid symmetric,
                              no k=0 for even Nk.
  verteilung                  DG_MSA_3D closed form; cosine transform checked by
                              brute force; phi\\f checked as a residual.
                              TSU_ESAKI (transverse-integrated contacts, not
                              in the reference): flat-band zero-bias
                              density equals the 3D Fermi-Dirac integral.
  get_diff                    upwind structure (k>0 couples only e-1, k<0 only
                              e+1); constant state with matching inflow is a
                              stationary solution of the transport part; a
                              closed-form 1-element case.
  get_drift                   two independent restatements agree: literal port of
                              the vectorised index code (n=3) vs the loop-form
                              formula; constant V => only the CAP part survives;
                              xi=0 edge carries B=0; antisymmetry of the real
                              potential part in xi.
  solve_transport_dg_2d       direct sparse LU residual; zero-bias equilibrium
                              gives zero current in the flat-band limit
                              (no potential, symmetric contacts).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla
from numpy.polynomial import legendre as npleg


# ---------------------------------------------------------------------------
# DG primitives (Hesthaven-Warburton style), DG_code/core/Jacobi*.m
# ---------------------------------------------------------------------------

def jacobi_gq(alpha: float, beta: float, N: int):
    """Gauss quadrature of Jacobi type (alpha,beta).  DG_code/core/JacobiGQ.m:1-30."""
    if N == 0:
        return np.array([-(alpha - beta) / (alpha + beta + 2.0)]), np.array([2.0])
    h1 = 2.0 * np.arange(N + 1) + alpha + beta
    with np.errstate(divide="ignore", invalid="ignore"):
        d = -0.5 * (alpha ** 2 - beta ** 2) / (h1 + 2.0) / h1
    i = np.arange(1, N + 1, dtype=float)
    off = 2.0 / (h1[:N] + 2.0) * np.sqrt(
        i * (i + alpha + beta) * (i + alpha) * (i + beta) / (h1[:N] + 1.0) / (h1[:N] + 3.0))
    J = np.diag(d) + np.diag(off, 1)
    if alpha + beta < 10 * np.finfo(float).eps:
        J[0, 0] = 0.0
    J = J + J.T
    lam, V = np.linalg.eigh(J)
    w = V[0, :] ** 2 * 2 ** (alpha + beta + 1) / (alpha + beta + 1) * math.gamma(alpha + 1) * \
        math.gamma(beta + 1) / math.gamma(alpha + beta + 1)
    return lam, w


def jacobi_gl(alpha: float, beta: float, N: int):
    """Gauss-Lobatto nodes and weights.  DG_code/core/JacobiGL.m:1-32.

    The weights use the Legendre polynomial P_N (``legendreP`` of the symbolic
    toolbox, JacobiGL.m:48), i.e. w = 2/(N(N+1) P_N(x)^2).
    """
    if N == 0:
        return np.array([0.0]), np.array([2.0])
    if N == 1:
        return np.array([-1.0, 1.0]), np.array([1.0, 1.0])
    xint, _ = jacobi_gq(alpha + 1, beta + 1, N - 2)
    x = np.concatenate([[-1.0], xint, [1.0]])
    n = N + 1
    w = np.zeros(n)
    c = np.zeros(n)
    c[n - 1] = 1.0
    for i in range((n + 1) // 2):
        w[i] = 2.0 / (n * (n - 1)) / npleg.legval(x[i], c) ** 2
        w[n - 1 - i] = w[i]
    return x, w


def jacobi_p(x, alpha: float, beta: float, N: int):
    """Orthonormal Jacobi polynomial of order N at x.  DG_code/core/JacobiP.m:1-38."""
    xp = np.asarray(x, dtype=float).ravel()
    PL = np.zeros((N + 1, xp.size))
    gamma0 = 2 ** (alpha + beta + 1) / (alpha + beta + 1) * math.gamma(alpha + 1) * \
        math.gamma(beta + 1) / math.gamma(alpha + beta + 1)
    PL[0, :] = 1.0 / math.sqrt(gamma0)
    if N == 0:
        return PL[0]
    gamma1 = (alpha + 1) * (beta + 1) / (alpha + beta + 3) * gamma0
    PL[1, :] = ((alpha + beta + 2) * xp / 2 + (alpha - beta) / 2) / math.sqrt(gamma1)
    if N == 1:
        return PL[1]
    aold = 2 / (2 + alpha + beta) * math.sqrt((alpha + 1) * (beta + 1) / (alpha + beta + 3))
    for i in range(1, N):
        h1 = 2 * i + alpha + beta
        anew = 2 / (h1 + 2) * math.sqrt((i + 1) * (i + 1 + alpha + beta) * (i + 1 + alpha) *
                                        (i + 1 + beta) / (h1 + 1) / (h1 + 3))
        bnew = -(alpha ** 2 - beta ** 2) / h1 / (h1 + 2)
        PL[i + 1, :] = 1 / anew * (-aold * PL[i - 1, :] + (xp - bnew) * PL[i, :])
        aold = anew
    return PL[N]


def vandermonde1d(N: int, r):
    """V_ij = P_j(r_i).  DG_code/core/Vandermonde1D.m:1-8."""
    r = np.asarray(r, dtype=float).ravel()
    V = np.zeros((r.size, N + 1))
    for j in range(N + 1):
        V[:, j] = jacobi_p(r, 0, 0, j)
    return V


def grad_jacobi_p(r, alpha, beta, N):
    """DG_code/core/GradJacobiP.m:1-12."""
    r = np.asarray(r, dtype=float).ravel()
    if N == 0:
        return np.zeros(r.size)
    return math.sqrt(N * (N + alpha + beta + 1)) * jacobi_p(r, alpha + 1, beta + 1, N - 1)


def grad_vandermonde1d(N, r):
    """DG_code/core/GradVandermonde1D.m:1-11."""
    r = np.asarray(r, dtype=float).ravel()
    DV = np.zeros((r.size, N + 1))
    for i in range(N + 1):
        DV[:, i] = grad_jacobi_p(r, 0, 0, i)
    return DV


def dmatrix1d(N, r, V):
    """Dr = Vr / V.  DG_code/core/Dmatrix1D.m:1-9."""
    Vr = grad_vandermonde1d(N, r)
    return np.linalg.solve(V.T, Vr.T).T


# ---------------------------------------------------------------------------
# Physical constants as the reference files spell them (quirk: they differ
# per file, DG_code/initParams.m:101-104 vs verteilung.m:6-10 vs
# solve_transport_DG_2D.m:6-8).
# ---------------------------------------------------------------------------
HBAR_IP = 1.054571817e-34            # initParams.m:101
M_ELEKTRON_IP = 0.041 * 9.1094e-31   # initParams.m:102 (hard-coded 0.041 m0)
Q_IP = 1.602176634e-19               # initParams.m:103
KB_IP = 1.38064852e-23               # initParams.m:104
H_VT = 6.626e-34                     # verteilung.m:6, solve_transport_DG_2D.m:6
M0_VT = 9.109e-31                    # verteilung.m:7, solve_transport_DG_2D.m:7
KB_VT = 1.38e-23                     # verteilung.m:9
Q_VT = 1.602e-19                     # verteilung.m:10
NQX = 10                             # initParams.m:61


def _linspace(a, b, n):
    return np.linspace(a, b, n)


# ---------------------------------------------------------------------------
# initParams (FV branch)
# ---------------------------------------------------------------------------

def init_params(mat: dict, EfL: float, EfR: float, V, m, deg_factor: float) -> dict:
    """Geometry, basis and inflow data.  DG_code/initParams.m:58-221.

    ``V`` is the subband energy on the material grid (eV) and ``m`` the
    x-mass on the material grid (kg), as solve_transport_DG_2D.m:44-49
    passes them.  ``mat['fermi_model']`` selects the verteilung branch
    ('DG_MSA_3D' or 'DG_MSA_2D'); the reference derives it from mat.type
    (initParams.m:11-55) and routes RTD types to a hard-coded branch.
    """
    dgp = mat["dg"]["params"]
    if str(dgp.get("Xi_solver", "FV")).upper() != "FV":
        raise NotImplementedError("only the FV xi solver is restated")
    p: dict = {}
    p["permute_shell"] = bool(dgp.get("permute_shell", True))
    p["deg_factor_temp"] = deg_factor
    p["choose_fermi"] = mat.get("fermi_model", "DG_MSA_3D")
    n = int(dgp["N_K_chi"])
    p["N_K_chi"] = n
    Nx = int(mat["Nx"])
    if (Nx - 1) % (n - 1):
        raise ValueError("(mat.Nx-1) must be a multiple of N_K_chi-1")
    p["N_chi"] = (Nx - 1) // (n - 1)                       # initParams.m:59
    p["N_xi"] = int(dgp["N_xi"])                           # :60
    p["Nqx"] = NQX                                         # :61
    p["Nk"] = int(dgp["N"])                                # :62
    p["hquer"] = HBAR_IP
    p["m_elektron"] = M_ELEKTRON_IP
    p["q"] = Q_IP
    p["kB"] = KB_IP
    p["T"] = 300.0
    p["Q_diff"] = p["hquer"] / p["m_elektron"]             # :106
    p["Q_drift"] = p["q"] / p["hquer"]                     # :107
    p["U"] = 0.0                                           # :109
    p["L_xi"] = float(dgp["Ly"])                           # :110
    p["L_chi"] = float(mat["L_x"]) * 1e-9                  # :111
    N_chi = p["N_chi"]
    p["delta_chi"] = p["L_chi"] / N_chi                    # :155
    p["disk_Rechengebiet1"] = _linspace(0.0, p["L_chi"], N_chi + 1)
    chi = np.zeros(N_chi * n)
    for a in range(N_chi):                                 # :157-159 (equispaced!)
        chi[n * a:n * (a + 1)] = _linspace(p["disk_Rechengebiet1"][a], p["disk_Rechengebiet1"][a + 1], n)
    p["chi_Koord"] = chi
    N_xi = p["N_xi"]
    p["xi_Rechengebiet"] = _linspace(-p["L_xi"] / 2, p["L_xi"] / 2, N_xi)   # :179
    p["delta_xi"] = p["xi_Rechengebiet"][1] - p["xi_Rechengebiet"][0]       # :180-181
    p["xi_Rechengebiet_Elem"] = _linspace(-(p["L_xi"] - p["delta_xi"]) / 2,
                                          (p["L_xi"] - p["delta_xi"]) / 2, N_xi)  # :182
    p["x"], p["w"] = jacobi_gl(0, 0, NQX - 1)              # :186
    p["r"], _ = jacobi_gl(0, 0, n - 1)                     # :187
    p["V_tilde"] = vandermonde1d(n - 1, p["x"])            # :188
    p["V"] = vandermonde1d(n - 1, p["r"])                  # :189
    p["invV"] = np.linalg.inv(p["V"])                      # :190
    p["M"] = p["invV"].T @ p["invV"]                       # :191
    # Basisfunktionen_Q_m.m:22-32 (FV branch)
    Nk = p["Nk"]
    y = p["xi_Rechengebiet_Elem"]
    k = -2 * np.pi / p["L_xi"] * (np.arange(1, Nk + 1) - 0.5 * (Nk + 1))
    p["k"] = k
    p["dk"] = abs(k[1] - k[0])
    p["phi"] = (1.0 / np.sqrt(p["L_xi"])) * np.exp(1j * np.outer(y, k))
    p["verteilung_l"], p["verteilung_r"] = verteilung(mat, p, N_xi, p["phi"], EfL, EfR,
                                                      np.asarray(V, float), np.asarray(m, float))
    return p


def verteilung(mat, p, N_xi, phi, EfL, EfR, V, m):
    """Inflow distributions in the k basis.  DG_code/core/verteilung.m:77-170."""
    fL, fR, k_fermi, dk = inflow_k(mat, p, EfL, EfR, V, m)
    y = p["xi_Rechengebiet_Elem"]
    # cosine transform to xi (:109-112 / :158-161), then phi \ f (:113-114 / :167-168)
    Cm = np.cos(np.outer(y[:N_xi], k_fermi))
    fLn = Cm @ fL * dk / 2 / np.pi
    fRn = Cm @ fR * dk / 2 / np.pi
    fnl = np.linalg.solve(phi, fLn.astype(complex))
    fnr = np.linalg.solve(phi, fRn.astype(complex))
    return fnl, fnr


def inflow_k(mat, p, EfL, EfR, V, m):
    """Contact occupations f_L(k), f_R(k) on the k grid before the transform
    to xi (DG_code/core/verteilung.m:77-99 pure 1D Fermi-Dirac, :118-153
    k_z-integrated).  Returns (fL, fR, k, dk)."""
    hb = H_VT / 2 / np.pi
    q, kB, Temp = Q_VT, KB_VT, float(mat["Temp"])
    Nx = int(mat["Nx"])
    Nk = p["Nk"]
    k_fermi = -2 * np.pi / p["L_xi"] * (np.arange(1, Nk + 1) - 0.5 * (Nk + 1))
    dk = abs(k_fermi[1] - k_fermi[0])
    EL = V[0] + hb ** 2 * k_fermi ** 2 / 2 / m[0] / q
    ER = V[Nx - 1] + hb ** 2 * k_fermi ** 2 / 2 / m[Nx - 1] / q
    if p["choose_fermi"] == "DG_MSA_3D":                                   # :77-114
        with np.errstate(over="ignore"):      # exp -> inf far above EF: occupation 0, as in MATLAB
            fL = 2 * p["deg_factor_temp"] * (1.0 / (1.0 + np.exp(q * (EL - EfL) / (kB * Temp))))
            fR = 2 * p["deg_factor_temp"] * (1.0 / (1.0 + np.exp(q * (ER - EfR) / (kB * Temp))))
    elif p["choose_fermi"] == "TSU_ESAKI":
        # transverse-integrated contacts: the form of verteilung.m:64-65 ('LS')
        # m kT/(pi hbar^2) ln(1+exp(q(EF-E)/kT)) with the device's EF, band
        # edge and mass (the reference branch hard-codes EF and m)
        xl = q * (EfL - EL) / (kB * Temp)
        xr = q * (EfR - ER) / (kB * Temp)
        fL = m[0] * kB * Temp / (np.pi * hb ** 2) * np.logaddexp(0.0, xl)
        fR = m[Nx - 1] * kB * Temp / (np.pi * hb ** 2) * np.logaddexp(0.0, xr)
    elif p["choose_fermi"] == "DG_MSA_2D":                                 # :118-152
        me_z = float(mat["me_z_ch"])
        kz_max = np.sqrt(2 * 2 * M0_VT * me_z * q * (np.max(V + 0.3))) / hb
        kz = _linspace(0.0, kz_max, 1001)
        dkz = kz[1] - kz[0]
        Ekz = hb ** 2 * kz ** 2 / 2 / M0_VT / me_z / q
        fL = 2 / np.pi * dkz * np.sum(1.0 / (1.0 + np.exp(q * (EL[:, None] + Ekz[None, :] - EfL) / (kB * Temp))), axis=1)
        fR = 2 / np.pi * dkz * np.sum(1.0 / (1.0 + np.exp(q * (ER[:, None] + Ekz[None, :] - EfR) / (kB * Temp))), axis=1)
    else:
        raise NotImplementedError(p["choose_fermi"])
    return fL, fR, k_fermi, dk


# ---------------------------------------------------------------------------
# Transport (DG in chi, upwind, diag(k))
# ---------------------------------------------------------------------------

def transport_blocks(p: dict) -> dict:
    """Element matrices K1..K6 of get_Diff.m:198-221 (with the xi-scaled mass)."""
    n = p["N_K_chi"]
    Dr = p["M"] @ dmatrix1d(n - 1, p["r"], p["V"])          # get_Diff.m:198
    Dr[np.abs(Dr) < 1e-15] = 0.0                           # :203-204
    Mx = p["delta_xi"] / 2 * p["M"]                        # :200 (xi spacing quirk)
    invM = np.linalg.inv(Mx)                               # :202
    P1 = np.zeros((n, n)); P1[0, 0] = 1
    P2 = np.zeros((n, n)); P2[0, n - 1] = 1
    P3 = np.zeros((n, n)); P3[n - 1, n - 1] = 1
    P4 = np.zeros((n, n)); P4[n - 1, 0] = 1
    K1 = invM @ (Dr + 0.5 * P1 - 0.5 * P3)                 # :214
    K2L = invM @ (0.5 * P1)                                # :216
    K2R = invM @ (0.5 * P3)                                # :217
    K3 = -0.5 * invM @ P2                                  # :218
    K4 = K3                                                # :219
    K5 = 0.5 * invM @ P4                                   # :220
    K6 = -K5                                               # :221
    return dict(Dr=Dr, invM=invM, K1=K1, K2L=K2L, K2R=K2R, K3=K3, K4=K4, K5=K5, K6=K6)


def get_diff(p: dict):
    """Sparse transport operator and inflow RHS.  DG_code/core/get_Diff.m:1-83
    (permute_shell branch :242-259, full upwind alpha = 1, :185-191)."""
    n, N_chi, Nk = p["N_K_chi"], p["N_chi"], p["Nk"]
    b = transport_blocks(p)
    detJ = np.full(N_chi, p["delta_chi"] / 2)
    alpha = np.ones(N_chi - 1)
    aL = sp.diags(np.concatenate([[1.0], alpha]))
    aR = sp.diags(np.concatenate([alpha, [1.0]]))
    Sp1 = sp.diags(1.0 / detJ[1:], 1, shape=(N_chi, N_chi))
    Sm1 = sp.diags(1.0 / detJ[1:], -1, shape=(N_chi, N_chi))
    S0 = sp.diags(1.0 / detJ)
    H1 = sp.kron(Sp1, b["K5"]) + sp.kron(Sm1, b["K3"]) + sp.kron(S0, b["K1"])
    H2 = sp.kron(Sp1 @ aL, b["K6"]) + sp.kron(Sm1 @ aR, b["K4"]) + \
        sp.kron(S0 @ aL, b["K2L"]) + sp.kron(S0 @ aR, b["K2R"])
    D = sp.diags(p["k"])
    A = p["Q_diff"] * (sp.kron(H1, D) + sp.kron(H2, abs(D)))
    Q1 = np.zeros(n * N_chi); Q1[:n] = b["invM"][:, 0] / detJ[0]
    Q2 = np.zeros(n * N_chi); Q2[n * (N_chi - 1):] = b["invM"][:, n - 1] / detJ[-1]
    k = p["k"]
    rhs = np.kron(Q1, (k > 0) * (np.abs(k) * p["verteilung_l"])) + \
        np.kron(Q2, (k < 0) * (np.abs(k) * p["verteilung_r"]))
    return A.tocsr(), p["Q_diff"] * rhs


# ---------------------------------------------------------------------------
# Potential operator Theta[V] + CAP
# ---------------------------------------------------------------------------

def v_transient(x, xg, Ug, U, Lx):
    """Clamped linear interpolation.  DG_code/core/V_transient.m:1-8."""
    x = np.asarray(x, float)
    Vp = np.zeros_like(x)
    inside = x < Lx
    Vp[inside] = np.interp(x[inside], xg, Ug)
    Vp[x >= Lx] = np.min(Ug)
    Vp[x < 0] = Ug[0]
    return Vp


def get_cap(q, L_xi, delta_xi):
    """Polynomial complex absorbing potential.  DG_code/core/get_CAP.m:1-12."""
    W0, nn = 1.0, 1.0
    L0 = L_xi / 2 - delta_xi
    weite = L_xi * 0.40
    q = np.asarray(q, float)
    return (W0 / weite * (q - L0 + weite)) ** (2 * nn) * (q >= L0 - weite) + \
        (W0 / weite * (q + L0 - weite)) ** (2 * nn) * (q <= -L0 + weite)


def drift_samples(mat: dict, p: dict, EM):
    """B_{e,q}(xi_l) = V(x+xi/2) - V(x-xi/2) - i CAP(xi) on the N_xi+1 xi edges.

    get_Drift_accel_trial_permutedShell.m:15-66.  Returns (Nqx*N_chi, N_xi+1) complex.
    """
    x = p["x"]
    chi_anteil = np.diff(x) / 2 * p["delta_chi"]                     # :16-18
    chi_GL = np.concatenate([[0.0], np.cumsum(chi_anteil)])          # :47-49
    xq = (p["disk_Rechengebiet1"][:p["N_chi"], None] + chi_GL[None, :]).ravel()  # :52-55
    xe = _linspace(-p["L_xi"] / 2, p["L_xi"] / 2, p["N_xi"] + 1)      # :28
    xg = np.asarray(mat["x"], float) * 1e-9
    EM = np.asarray(EM, float)
    X = xq[:, None]
    Y = xe[None, :]
    B = v_transient(X + 0.5 * Y, xg, EM, p["U"], p["L_chi"]) - \
        v_transient(X - 0.5 * Y, xg, EM, p["U"], p["L_chi"]) - \
        1j * get_cap(np.broadcast_to(Y, (X.shape[0], Y.shape[1])), p["L_xi"], p["delta_xi"])
    return B, xq, xe


def gamma_pages(p: dict, B):
    """G_{e,q} = i phi^H C_{e,q} phi with C tridiagonal (delta_xi/4)(B_l+B_{l+1}; B_{l+1}).

    get_Drift_accel_trial_permutedShell.m:95-147 (and loop form get_Drift.m:81-87).
    Returns (Nqx*N_chi, Nk, Nk) complex.
    """
    phi = p["phi"]
    dg = B[:, :-1] + B[:, 1:]                        # main diagonal (:103)
    off = B[:, 1:-1]                                 # both off diagonals (:104)
    Cphi = dg[:, :, None] * phi[None, :, :]
    Cphi[:, :-1, :] += off[:, :, None] * phi[None, 1:, :]
    Cphi[:, 1:, :] += off[:, :, None] * phi[None, :-1, :]
    Cphi *= p["delta_xi"] / 4                        # :135
    return 1j * np.matmul(phi.conj().T[None, :, :], Cphi)   # :143-147


def drift_element_weights(p: dict):
    """S[q,a,b] = w_q W_qa W_qb and Q_drift*invM (xi-scaled) used by the DG projection."""
    W = p["V_tilde"] @ np.linalg.inv(p["V"])         # :183  W = V_tilde / V
    invM = np.linalg.inv(p["delta_xi"] / 2 * p["M"]) # :12-13
    S = p["w"][:, None, None] * W[:, :, None] * W[:, None, :]
    return W, invM, S


def get_drift(mat: dict, p: dict, EM):
    """Sparse Theta[V]+CAP block-diagonal operator (general N_K_chi).

    Loop form DG_code/core/get_Drift.m:128-153:
      G[(e,a,j),(e,b,m)] = Q_drift * invM_ab * sum_q w_q W_qa W_qb Gamma_{e,q}[j,m]
    (elementwise product with invM, get_Drift.m:141), rows in the permuted
    index ((e*n + a)*Nk + j).
    """
    n, N_chi, Nk = p["N_K_chi"], p["N_chi"], p["Nk"]
    B, _, _ = drift_samples(mat, p, EM)
    G = gamma_pages(p, B).reshape(N_chi, NQX, Nk, Nk)
    _, invM, S = drift_element_weights(p)
    coef = p["Q_drift"] * invM[None, :, :] * S            # (q, a, b)
    blocks = np.einsum("qab,eqjm->eajbm", coef, G)        # (e, a, j, b, m)
    rows = (np.arange(N_chi)[:, None, None, None, None] * n + np.arange(n)[None, :, None, None, None]) * Nk + \
        np.arange(Nk)[None, None, :, None, None]
    cols = (np.arange(N_chi)[:, None, None, None, None] * n + np.arange(n)[None, None, None, :, None]) * Nk + \
        np.arange(Nk)[None, None, None, None, :]
    rows = np.broadcast_to(rows, blocks.shape).ravel()
    cols = np.broadcast_to(cols, blocks.shape).ravel()
    dim = N_chi * n * Nk
    return sp.csr_matrix((blocks.ravel(), (rows, cols)), shape=(dim, dim))


def get_drift_literal(mat: dict, p: dict, EM):
    """Literal restatement of the vectorised sparse build (N_K_chi = 3 only).

    Follows get_Drift_accel_trial_permutedShell.m:183-333 statement by
    statement with MATLAB column-major reshapes, so that an indexing slip in
    either restatement shows up as a mismatch against ``get_drift``.
    """
    n, N_chi, Ny = p["N_K_chi"], p["N_chi"], p["Nk"]
    if n != 3:
        raise ValueError("the vectorised reference code hard-codes N_K_chi = 3")
    F = "F"
    B, _, _ = drift_samples(mat, p, EM)
    Cpages = gamma_pages(p, B)                                   # (num_iter, Nk, Nk) after permute [3,1,2]
    W = p["V_tilde"] @ np.linalg.inv(p["V"])                     # :183
    invM = np.linalg.inv(p["delta_xi"] / 2 * p["M"])             # :12-13
    w = p["w"].reshape(-1, 1)
    C = Cpages * np.tile(w, (N_chi, 1))[:, :, None]              # :200
    W_rep = np.tile(W, (N_chi, 1))                               # :202
    W_rep = np.tile(W_rep, (Ny * Ny, 1))                         # :203
    C = C.reshape(-1, 1, order=F)                                # :204
    C = W_rep * C                                                # :209
    C = C.T.reshape((3, 10, -1), order=F)                        # :212
    W_rep = W_rep.T.reshape((3, 10, -1), order=F)                # :213
    W_rep = np.transpose(W_rep, (1, 0, 2))                       # :214
    C = np.einsum("aqp,qbp->abp", C, W_rep)                      # :215 pagemtimes
    C = np.transpose(C, (1, 2, 0)).reshape((n * N_chi * Ny * Ny, n), order=F)   # :216
    C = np.tile(p["Q_drift"] * invM, (1, N_chi * Ny * Ny)).T * C               # :219
    # row index (:254-281)
    vspace = np.arange(0, (N_chi - 1) * n * Ny + 1, n * Ny, dtype=np.int64)
    Blockx = np.ones((N_chi, n, n), dtype=np.int64) * vspace[:, None, None]
    XX = np.array([[1, 1, 1], [Ny + 1] * 3, [2 * Ny + 1] * 3], dtype=np.int64)
    XX = np.transpose(np.repeat(XX[:, :, None], N_chi, axis=2), (2, 0, 1))
    T123 = Blockx + XX
    NN = np.tile(np.arange(Ny, dtype=np.int64).reshape(-1, 1), (1, N_chi))
    NN = NN.T.reshape(-1, order=F)
    T123 = np.tile(T123, (Ny, 1, 1))
    Offset = np.ones((N_chi * Ny, n, n), dtype=np.int64) * NN[:, None, None]
    Line_idx = T123 + Offset
    Line_idx = np.transpose(Line_idx, (1, 0, 2)).reshape((-1, n), order=F)
    Line_idx = np.tile(Line_idx, (Ny, 1))
    # column index (:286-317)
    XC = np.array([[1, Ny + 1, 2 * Ny + 1]] * 3, dtype=np.int64)
    XC = np.transpose(np.repeat(XC[:, :, None], N_chi, axis=2), (2, 0, 1))
    Blockx_c = np.ones((N_chi, n, n), dtype=np.int64) * vspace[:, None, None]
    Test_c = np.tile(Blockx_c + XC, (Ny, 1, 1))
    Column_idx = np.transpose(Test_c, (1, 0, 2)).reshape((-1, n), order=F)
    Column_idx = np.tile(Column_idx, (Ny, 1))
    spacer = np.tile(np.ones((n, n), dtype=np.int64), (N_chi * Ny * Ny, 1))
    spacer_offset = np.tile(np.arange(Ny, dtype=np.int64).reshape(-1, 1), (1, n * N_chi * Ny))
    spacer_offset = spacer_offset.T.reshape(-1, order=F)
    spacer = spacer * spacer_offset[:, None]
    Column_idx = Column_idx + spacer
    zl = Line_idx.reshape(-1, order=F) - 1
    spc = Column_idx.reshape(-1, order=F) - 1
    vals = C.reshape(-1, order=F)
    dim = N_chi * n * Ny
    return sp.csr_matrix((vals, (zl, spc)), shape=(dim, dim))


def get_sysm(mat: dict, p: dict, EM):
    """A = A_diff + G, rhs.  DG_code/core/get_SysM.m:1-16 (FV: no conjugate, :7)."""
    A, rhs = get_diff(p)
    G = get_drift(mat, p, EM)
    return (A + G).tocsr(), rhs


# ---------------------------------------------------------------------------
# Observables and the solve driver
# ---------------------------------------------------------------------------

def node_values(p: dict, field_on_material):
    """Material grid -> DG nodes (equispaced sharing), solve_transport_DG_2D.m:81-87."""
    n, N_chi = p["N_K_chi"], p["N_chi"]
    f = np.asarray(field_on_material)
    out = np.zeros((N_chi * n,) + f.shape[1:], dtype=f.dtype)
    for e in range(N_chi):
        out[e * n:(e + 1) * n] = f[e * (n - 1):e * (n - 1) + n]
    return out


def observables(p: dict, rho, m_nodes):
    """rho_u = phi * rho (per node); n, j at xi = 0.  solve_transport_DG_2D.m:114-156."""
    n, N_chi, Nk, N_xi = p["N_K_chi"], p["N_chi"], p["Nk"], p["N_xi"]
    R = np.asarray(rho).reshape(N_chi * n, Nk).T
    rho_u = p["phi"] @ R                                           # :118
    hb = H_VT / 2 / np.pi
    if N_xi % 2 == 1:
        c = (N_xi - 1) // 2
        n_DG = np.real(rho_u[c, :])
        j = p["q"] * hb / m_nodes * np.imag(rho_u[c + 1, :] - rho_u[c - 1, :]) / 2 / p["delta_xi"]
    else:
        c = N_xi // 2
        n_DG = 0.5 * (np.real(rho_u[c - 1, :]) + np.real(rho_u[c, :]))          # :154
        j = p["q"] * hb / m_nodes * np.imag(rho_u[c, :] - rho_u[c - 1, :]) / p["delta_xi"]   # :155
    return n_DG, j, rho_u


def solve_transport_dg_2d(mat: dict, Esub, Vsub, EfL: float, EfR: float, return_system=False):
    """Direct FP64 solve per valley/mode.  DG_code/solve_transport_DG_2D.m:1-225.

    Esub: (n_of_valleys, Nx, n_of_modes) subband energies [eV];
    Vsub: (n_of_valleys, Nx, n_of_modes, Ny) transverse wavefunctions.
    """
    Esub = np.asarray(Esub, float)
    Vsub = np.asarray(Vsub)
    nv, nm = int(mat["n_of_valleys"]), int(mat["n_of_modes"])
    deg = np.atleast_1d(np.asarray(mat["deg_factor"], float))
    me_x = np.asarray(mat["me_x"], float)
    Ny = int(mat["Ny"])
    mid = int(math.ceil(Ny / 2)) - 1
    dgp = mat["dg"]["params"]
    MPC = float(dgp.get("MPC", 1.0))
    out_n = None
    out_j = None
    RHO = None
    systems = []
    p = None
    for iv in range(nv):
        for im in range(nm):
            m = me_x[iv, :, mid] * M0_VT
            p = init_params(mat, EfL, EfR, Esub[iv, :, im], m, deg[iv])
            m_nodes = node_values(p, m)
            Vsub_nodes = node_values(p, Vsub[iv, :, im, :])
            A, rhs = get_sysm(mat, p, Esub[iv, :, im])
            if MPC != 1:
                A = MPC * A
                rhs = MPC * rhs
            rho = spla.spsolve(A.tocsc(), rhs)                 # :101
            n_DG, j, rho_u = observables(p, rho, m_nodes)
            n_mode = n_DG[:, None] * np.abs(Vsub_nodes) ** 2 / float(mat["dy"]) / 1e-9   # :203-205
            out_n = n_mode if out_n is None else out_n + n_mode
            out_j = j if out_j is None else out_j + j
            if RHO is None:
                RHO = np.zeros((nv, nm, rho_u.shape[1], rho_u.shape[0]), complex)
            RHO[iv, im] = rho_u.T
            if return_system:
                systems.append((A, rhs, rho))
    DG = dict(p=p, n=out_n, j=out_j, RHO=RHO, chi=p["chi_Koord"], xi=p["xi_Rechengebiet_Elem"])
    if return_system:
        DG["systems"] = systems
    return DG


# ---------------------------------------------------------------------------
# Matrix-free FP64 operator (same operator, factored form) for sizes where
# the sparse assembly does not fit: y = A_diff u + G u with
#   G u |_(e,a) = i Q_drift phi^H sum_b invM_ab Cbar_{e,ab} (phi u_(e,b)),
#   Cbar_{e,ab} = sum_q w_q W_qa W_qb C_{e,q}   (tridiagonal in xi).
# ---------------------------------------------------------------------------

@dataclass
class MatrixFreeOperator:
    p: dict
    B: np.ndarray                  # (N_chi*Nqx, N_xi+1) edge samples
    blocks: dict = field(default_factory=dict)

    @classmethod
    def build(cls, mat, p, EM):
        B, _, _ = drift_samples(mat, p, EM)
        return cls(p=p, B=B, blocks=transport_blocks(p))

    def edge_coefficients(self):
        """Bbar[e,a,b,l] = Q_drift*invM_ab*(delta_xi/4)*sum_q w_q W_qa W_qb B_{e,q}(l)."""
        p = self.p
        n, N_chi = p["N_K_chi"], p["N_chi"]
        _, invM, S = drift_element_weights(p)
        Bq = self.B.reshape(N_chi, NQX, -1)
        Bbar = np.einsum("qab,eql->eabl", S, Bq)
        return Bbar * (p["Q_drift"] * invM * p["delta_xi"] / 4)[None, :, :, None]

    def apply_transport(self, u):
        p, b = self.p, self.blocks
        n, N_chi, Nk = p["N_K_chi"], p["N_chi"], p["Nk"]
        U = u.reshape(N_chi, n, Nk)
        k = p["k"]
        ak = np.abs(k)
        J = p["delta_chi"] / 2
        y = np.einsum("ab,ebj->eaj", b["K1"], U) * k + \
            np.einsum("ab,ebj->eaj", b["K2L"] + b["K2R"], U) * ak
        # neighbour couplings: from e+1 (K5 / K6), from e-1 (K3 / K4)
        y[:-1] += np.einsum("ab,ebj->eaj", b["K5"], U[1:]) * k + np.einsum("ab,ebj->eaj", b["K6"], U[1:]) * ak
        y[1:] += np.einsum("ab,ebj->eaj", b["K3"], U[:-1]) * k + np.einsum("ab,ebj->eaj", b["K4"], U[:-1]) * ak
        return (p["Q_diff"] / J * y).reshape(-1)

    def apply_drift(self, u, Bbar=None):
        p = self.p
        n, N_chi, Nk = p["N_K_chi"], p["N_chi"], p["Nk"]
        if Bbar is None:
            Bbar = self.edge_coefficients()
        phi = p["phi"]
        U = u.reshape(N_chi * n, Nk).T                      # (Nk, cols)
        T = (phi @ U).T.reshape(N_chi, n, -1)               # (e, b, l)
        Z = np.zeros_like(T)
        diag = Bbar[..., :-1] + Bbar[..., 1:]               # (e,a,b,l)
        off = Bbar[..., 1:-1]
        Z += np.einsum("eabl,ebl->eal", diag, T)
        Z[:, :, :-1] += np.einsum("eabl,ebl->eal", off, T[:, :, 1:])
        Z[:, :, 1:] += np.einsum("eabl,ebl->eal", off, T[:, :, :-1])
        Y = 1j * (phi.conj().T @ Z.reshape(N_chi * n, -1).T)    # (Nk, cols)
        return Y.T.reshape(-1)

    def apply(self, u, Bbar=None):
        return self.apply_transport(u) + self.apply_drift(u, Bbar)
