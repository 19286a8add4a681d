"""Full-2D Wigner operator F(rho_x, rho_y, X, Y) on one MI355X.

Host side of the solve behind ``solve_transport_DG_full2D(mat, Vxy, EfL, EfR,
rho0)`` (DG_code/solve_transport_DG_full2D.m:1).  It reads ``mat`` and
``mat['dg']['params']`` with the reference's parameter names and defaults
(DG_code/initParams_full2D.m, core/get_Diff_full2D.m, core/get_Drift_full2D.m,
core/get_Boundary_full2D.m, core/get_Observables_full2D.m) and reduces the
operator to structured pieces:

* per-element DG blocks of each centre axis (H1, interior H2) and the face
  lifts of its two ends;
* the relative-axis operators: central difference -i D, identity (Rusanov),
  dense characteristic matrices |A|, A+ R, A- R (periodic rho_y: circulant);
* the potential on the material grid with the quadrature/projection factors
  of the FV-consistent (or collocated) discretisation and the CAP profiles;
* the inflow right-hand side and the sparse zero-separation weights of the
  observables.

libwdg then applies, preconditions, solves and evaluates observables with HIP
kernels (``wdg2_*`` in include/wigner_dg.h).  Nothing here evaluates the
operator on the CPU; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import math
from dataclasses import dataclass, field

import numpy as np
from numpy.polynomial import legendre as npleg

from . import _lib
from .contact_modes import contact_mode_data, contact_mode_reservoir
from ._lib import check, dptr
from .solver import SolveOptions, _as_c128

HBAR = 1.054571817e-34      # get_Diff_full2D.m:810-813 physicalConstants
M0 = 9.1093837015e-31
QE = 1.602176634e-19
KB = 1.38064852e-23         # get_DefaultReservoirRho_full2D.m

PRECONDITIONERS = {"none": _lib.WDG2_PRECOND_NONE, "jacobi": _lib.WDG2_PRECOND_JACOBI,
                   "rowabs": _lib.WDG2_PRECOND_ROWABS, "line-transverse": _lib.WDG2_PRECOND_LINE}


# ---------------------------------------------------------------------------
# parameters (readParam: params.name, else params.full2D.name)
# ---------------------------------------------------------------------------

def _empty(v):
    return v is None or (isinstance(v, (list, tuple, np.ndarray)) and np.size(v) == 0)


def read_param(params: dict, name: str, default=None):
    """DG_code/solve_transport_DG_full2D.m:1339-1349 (readParam)."""
    if isinstance(params, dict):
        v = params.get(name)
        if not _empty(v):
            return v
        sub = params.get("full2D")
        if isinstance(sub, dict) and not _empty(sub.get(name)):
            return sub[name]
    return default


def dg_params(mat: dict) -> dict:
    return (mat.get("dg") or {}).get("params") or {}


def material_axis(mat: dict, name: str, dname: str, nname: str) -> np.ndarray:
    """getCoordinateVector: mat.x if present, else (0:N-1)*dx (nm)."""
    if not _empty(mat.get(name)):
        return np.asarray(mat[name], float).ravel()
    return np.arange(int(mat[nname])) * float(mat[dname])


def mass_kg(m) -> np.ndarray:
    """massToKg (get_Diff_full2D.m:757-768): relative masses are scaled by m0;
    non-positive masses fall back to 0.041 m0."""
    m = np.asarray(m, float)
    kg = np.where(np.abs(m) > 1e-25, m * M0, m)
    return np.where(kg <= 0, 0.041 * M0, kg)


# ---------------------------------------------------------------------------
# DG reference element (Legendre-Gauss-Lobatto), independent of the MATLAB
# JacobiGL/Vandermonde1D code path: nodes from the roots of P'_{n-1}.
# ---------------------------------------------------------------------------

def lgl(n: int):
    """n Legendre-Gauss-Lobatto nodes and weights on [-1, 1]."""
    if n < 2:
        raise ValueError("at least two DG nodes per element are needed")
    N = n - 1
    cN = np.zeros(N + 1); cN[N] = 1.0
    inner = npleg.legroots(npleg.legder(cN)) if N > 1 else np.array([])
    r = np.concatenate([[-1.0], np.sort(inner.real), [1.0]])
    # Newton polish on (1 - r^2) P'_N(r) = 0 for the interior nodes
    for _ in range(3):
        if N > 1:
            d1 = npleg.legval(r[1:-1], npleg.legder(cN))
            d2 = npleg.legval(r[1:-1], npleg.legder(cN, 2))
            r[1:-1] -= d1 / d2
    w = 2.0 / (N * (N + 1) * npleg.legval(r, cN) ** 2)
    return r, w


def onb_vandermonde(n: int, r) -> tuple[np.ndarray, np.ndarray]:
    """Orthonormal Legendre basis P_j sqrt((2j+1)/2), j < n, and its derivative at r."""
    r = np.asarray(r, float)
    V = np.zeros((r.size, n)); Vr = np.zeros((r.size, n))
    for j in range(n):
        c = np.zeros(j + 1); c[j] = math.sqrt((2 * j + 1) / 2)
        V[:, j] = npleg.legval(r, c)
        Vr[:, j] = npleg.legval(r, npleg.legder(c)) if j else 0.0
    return V, Vr


@dataclass
class DGAxis:
    """Rectangular DG axis (initParams_full2D.m:143-197)."""
    nl: int
    ne: int
    vertices: np.ndarray
    r: np.ndarray
    V: np.ndarray
    M: np.ndarray
    S: np.ndarray

    @property
    def ndof(self):
        return self.nl * self.ne

    @property
    def jac(self):
        return np.diff(self.vertices) / 2

    @property
    def nodes(self):
        mid = 0.5 * (self.vertices[:-1] + self.vertices[1:])
        return (mid[:, None] + self.jac[:, None] * self.r[None, :]).ravel()


def dg_axis(coord: np.ndarray, nl: int, n_el=None, label="X") -> DGAxis:
    n_int = coord.size - 1
    if nl < 2:
        raise ValueError("N_K must be at least 2")
    if n_el is None:
        if n_int % (nl - 1):
            raise ValueError(f"{label} grid is not compatible with {nl} local DG nodes")
        n_el = n_int // (nl - 1)
    r, _ = lgl(nl)
    V, Vr = onb_vandermonde(nl, r)
    Vinv = np.linalg.inv(V)
    M = Vinv.T @ Vinv                       # M = inv(V V^T)
    Dr = Vr @ Vinv
    return DGAxis(nl, int(n_el), np.linspace(coord[0], coord[-1], int(n_el) + 1), r, V, M, M @ Dr)


@dataclass
class AxisBlocks:
    """Block-tridiagonal centre operator along one axis: diagonal n x n blocks
    per element plus the single column that couples to the neighbour's face
    node (lower: last node of e-1, upper: first node of e+1)."""
    diag: np.ndarray    # (ne, nl, nl)
    lower: np.ndarray   # (ne, nl)
    upper: np.ndarray   # (ne, nl)


def axis_operators(ax: DGAxis, periodic: bool = False) -> dict:
    """H1 (central), interior H2 and the face lifts of get_Diff_full2D.m:206-266.

    K1 = M^-1 (S + P1/2 - P3/2) on the element, the central-flux coupling
    -M^-1 e1 e_n^T / 2 to the left neighbour's last node and
    +M^-1 e_n e1^T / 2 to the right neighbour's first node, each scaled by 1/J;
    H2 = M^-1 (P1 + P3)/2 with the opposite-signed couplings, minus the
    physical end faces, which are lifts M^-1 e_face / J of the trace.
    """
    n, ne = ax.nl, ax.ne
    Mi = np.linalg.inv(ax.M)
    J = ax.jac
    e0 = np.zeros(n); e0[0] = 1
    en = np.zeros(n); en[-1] = 1
    K1 = Mi @ (ax.S + 0.5 * np.outer(e0, e0) - 0.5 * np.outer(en, en))
    K2 = 0.5 * Mi @ (np.outer(e0, e0) + np.outer(en, en))
    from_left = -0.5 * Mi[:, 0]           # column n-1 of -M^-1 P2 / 2
    from_right = 0.5 * Mi[:, n - 1]       # column 0 of M^-1 P4 / 2
    # a periodic axis closes the last/first faces into one interior interface
    # (close_PeriodicDGAxis_full2D.m): the end elements couple across the seam
    # with the interior fluxes, scaled by their own jacobian
    def up_j(e):
        return J[e + 1] if e < ne - 1 else J[e]
    lower_on = [e > 0 or periodic for e in range(ne)]
    upper_on = [e < ne - 1 or periodic for e in range(ne)]
    h1 = AxisBlocks(np.stack([K1 / J[e] for e in range(ne)]),
                    np.stack([from_left / J[e] if lower_on[e] else np.zeros(n) for e in range(ne)]),
                    np.stack([from_right / up_j(e) if upper_on[e] else np.zeros(n) for e in range(ne)]))
    d2 = []
    for e in range(ne):
        blk = K2.copy()
        if e == 0 and not periodic:
            blk -= 0.5 * np.outer(Mi[:, 0], e0)
        if e == ne - 1 and not periodic:
            blk -= 0.5 * np.outer(Mi[:, n - 1], en)
        d2.append(blk / J[e])
    h2 = AxisBlocks(np.stack(d2), np.stack([from_left / J[e] if lower_on[e] else np.zeros(n) for e in range(ne)]),
                    np.stack([-from_right / up_j(e) if upper_on[e] else np.zeros(n) for e in range(ne)]))
    return {"h1": h1, "h2i": h2, "lift_low": Mi[:, 0] / J[0], "lift_high": Mi[:, n - 1] / J[-1]}


# ---------------------------------------------------------------------------
# Relative FV axes and characteristic matrices
# ---------------------------------------------------------------------------

@dataclass
class RelAxis:
    """FV relative axis (initParams_full2D.m:199-214)."""
    n: int
    length: float

    @property
    def edges(self):
        return np.linspace(-self.length / 2, self.length / 2, self.n + 1)

    @property
    def cells(self):
        e = self.edges
        return 0.5 * (e[:-1] + e[1:])

    @property
    def delta(self):
        e = self.edges
        return e[1] - e[0]


def rho_modes(n: int, delta: float):
    """Eigenpairs of the Hermitian Toeplitz A = -i D, D = tridiag(-1, 0, 1)/(2 delta):
    lambda_k = cos(k pi/(n+1))/delta with orthonormal eigenvector
    sqrt(2/(n+1)) i^j sin(j k pi/(n+1)), j, k = 1..n (columns)."""
    j = np.arange(1, n + 1)
    Phi = np.sqrt(2.0 / (n + 1)) * (1j ** (j % 4))[:, None] * np.sin(np.pi * np.outer(j, j) / (n + 1))
    return Phi, np.cos(np.pi * j / (n + 1)) / delta


def rho_modes_periodic(n: int, delta: float):
    """Eigenpairs of the Hermitian circulant -i D with the periodic central
    difference (initParams_full2D.m:82-87): lambda_k = sin(2 pi k/n)/delta,
    eigenvectors exp(2 pi i j k/n)/sqrt(n)."""
    j = np.arange(n)
    Phi = np.exp(2j * np.pi * np.outer(j, j) / n) / np.sqrt(n)
    return Phi, np.sin(2 * np.pi * j / n) / delta


def characteristic_split(n: int, delta: float, tol: float = 1e-12, periodic: bool = False):
    """A+, A-, eigenvalues of A = -i D, D = tridiag(-1, 0, 1)/(2 delta).

    A is Hermitian Toeplitz; its eigenpairs are closed form:
    lambda_k = cos(k pi/(n+1))/delta with eigenvector i^j sin(j k pi/(n+1)).
    Eigenvalues with |lambda| <= tol go to neither part and entries below
    1e-12 are zeroed, as in splitCharacteristicMatrix
    (get_Diff_full2D.m:718-730).
    """
    Phi, lam = rho_modes_periodic(n, delta) if periodic else rho_modes(n, delta)
    lp = np.where(lam > tol, lam, 0.0)
    lm = np.where(lam < -tol, lam, 0.0)
    Ap = (Phi * lp) @ Phi.conj().T
    Am = (Phi * lm) @ Phi.conj().T
    Ap[np.abs(Ap) < 1e-12] = 0
    Am[np.abs(Am) < 1e-12] = 0
    return Ap, Am, lam


def cap_profile(q, length, delta):
    """Polynomial absorber W(q) of DG_code/core/get_CAP.m (W0 = 1, n = 1, width 0.4 L)."""
    q = np.asarray(q, float)
    L0 = length / 2 - delta
    width = 0.4 * length
    right = np.where(q >= L0 - width, ((q - L0 + width) / width) ** 2, 0.0)
    left = np.where(q <= -L0 + width, ((q + L0 - width) / width) ** 2, 0.0)
    return right + left


# ---------------------------------------------------------------------------
# The structured operator
# ---------------------------------------------------------------------------

@dataclass
class Part:
    """coef * kron(C, R) with C along `axis` (X or Y) and R along the same
    relative axis (get_Diff_full2D.m:463-474 addOperatorPart)."""
    name: str
    axis: int               # 0: X / rho_x, 1: Y / rho_y
    center: int             # _lib.WDG2_CENTER_BANDED or _FACE
    rel: int                # _lib.WDG2_REL_CENTRAL, _IDENTITY or _DENSE
    coef: float
    blocks: AxisBlocks | None = None
    face: int = 0           # FACE: 0 low end, 1 high end
    lift: np.ndarray | None = None
    mask: np.ndarray | None = None      # FACE: per transverse centre DOF
    dense: np.ndarray | None = None     # DENSE: complex (n_rho, n_rho)


def _flux(value, allowed, kind):
    f = str(value).strip().lower()
    if f in ("lf", "lax-friedrichs", "rusanov-local"):
        f = "rusanov"
    elif f in ("matrix-upwind", "characteristic-upwind") and "upwind" in allowed:
        f = "upwind"
    if f not in allowed:
        raise ValueError(f"DG:Full2D:UnknownFluxType: unknown {kind} flux '{value}'")
    return f


def y_boundary_type(params: dict) -> str:
    """get_YBoundaryType_full2D.m:14-45."""
    v = str(read_param(params, "full2D_Y_boundary", "specular")).strip().lower()
    aliases = {"specular-reflection": "specular", "reflection": "specular", "reflecting": "specular",
               "zero_inflow": "zero-inflow", "zeroinflow": "zero-inflow", "zero": "zero-inflow",
               "hardwall": "hard-wall", "hard_wall": "hard-wall", "dirichlet": "hard-wall"}
    v = aliases.get(v, v)
    if v not in ("specular", "zero-inflow", "hard-wall", "periodic"):
        raise ValueError(f"DG:Full2D:UnknownYBoundary: {v}")
    return v


class Full2DSetup:
    """Everything the device needs for one (mat, Vxy, EfL, EfR) instance."""

    def __init__(self, mat: dict, Vxy=None, EfL=None, EfR=None, host_reservoirs: bool = True):
        """host_reservoirs=False keeps only the host-solved contact modes of
        contact-mode reservoirs; the reservoirs and the inflow rhs are then
        built on the device (Full2DOperator.upload_contact_rhs)."""
        params = dg_params(mat)
        self.mat, self.params = mat, params
        self.host_reservoirs = host_reservoirs
        self.EfL, self.EfR = EfL, EfR
        self._vxy = Vxy
        ytype = y_boundary_type(params)
        if ytype == "hard-wall":
            raise NotImplementedError(f"DG:Full2D:Unsupported: Y boundary '{ytype}' is not built on the device")
        self.periodic = ytype == "periodic"
        mass_model = str(read_param(params, "full2D_massModel", "constant")).strip().lower()
        if mass_model not in ("constant", "scalar", "legacy", "reference"):
            raise NotImplementedError("DG:Full2D:Unsupported: position-dependent mass is not built on the device")
        self.ytype = ytype
        scale = float(read_param(params, "full2D_coordinate_scale", 1e-9))
        self.coord_scale = scale
        gx = material_axis(mat, "x", "dx", "Nx") * scale
        gy = material_axis(mat, "y", "dy", "Ny") * scale
        self.grid_x, self.grid_y = gx, gy
        nk_def = read_param(params, "N_K_chi", 3)
        self.X = dg_axis(gx, int(read_param(params, "N_K_X", nk_def)), read_param(params, "N_el_X"), "X")
        self.Y = dg_axis(gy, int(read_param(params, "N_K_Y", nk_def)), read_param(params, "N_el_Y"), "Y")
        nrx = int(read_param(params, "N_rho_x", read_param(params, "N_xi", 40)))
        nry = int(read_param(params, "N_rho_y", nrx))
        l_def = read_param(params, "Ly", max(abs(gy[-1] - gy[0]), np.finfo(float).eps))
        self.RX = RelAxis(nrx, float(read_param(params, "L_rho_x", l_def)))
        self.RY = RelAxis(nry, float(read_param(params, "L_rho_y", l_def)))
        if self.periodic:     # initParams_full2D.m:65-76
            width = gy[-1] - gy[0]
            self.RY = RelAxis(nry, float(read_param(params, "L_rho_y", 2 * width)))
            if abs(self.RY.length - 2 * width) > 1e-10 * width:
                raise ValueError("DG:Full2D:PeriodicRelativeLength: periodic Y requires L_rho_y = 2*(Ymax-Ymin)")
            if self.Y.ne % 2 or nry % 2 or nry < 4:
                raise ValueError("DG:Full2D:PeriodicGrid: periodic Y requires even N_el_Y and even N_rho_y >= 4")
        self.nrel = nrx * nry
        self.ncen = self.X.ndof * self.Y.ndof
        self.n_total = self.nrel * self.ncen
        self._transport()
        self._boundary(Vxy)
        self._potential(Vxy)

    # -- transport parts (get_Diff_full2D.m:36-204) --
    def _transport(self):
        params, mat = self.params, self.mat
        general = read_param(params, "full2D_fluxType",
                             read_param(params, "rho_flux", read_param(params, "fluxType", "rusanov")))
        xdef = _flux(general, ("central", "rusanov"), "general")
        self.flux_x = _flux(read_param(params, "full2D_X_flux", xdef), ("central", "rusanov", "upwind"), "X")
        self.flux_y = _flux(read_param(params, "full2D_Y_flux", self.flux_x), ("central", "rusanov", "upwind"), "Y")
        tol = float(read_param(params, "full2D_characteristicTol", 1e-12))
        me_x = mat.get("me_x_ch", 0.041)
        me_y = mat.get("me_y_ch", me_x)
        f = float(read_param(params, "full2D_diff_scale", read_param(params, "rho_diff_scale", 1)))
        sx = f * float(read_param(params, "full2D_Q_diff_x",
                                  read_param(params, "full2D_transportScaleX", HBAR / float(mass_kg(me_x)))))
        sy = f * float(read_param(params, "full2D_Q_diff_y",
                                  read_param(params, "full2D_transportScaleY", HBAR / float(mass_kg(me_y)))))
        self.scales = (sx, sy)
        ops = {0: axis_operators(self.X), 1: axis_operators(self.Y, self.periodic)}
        self.axis_ops = ops
        self.char = {}
        parts = []
        for ax, R, flux, s, key in ((0, self.RX, self.flux_x, sx, "X"), (1, self.RY, self.flux_y, sy, "Y")):
            Ap, Am, lam = characteristic_split(R.n, R.delta, tol, periodic=(ax == 1 and self.periodic))
            theta = 0.0
            if flux == "rusanov":
                theta = float(read_param(params, f"full2D_thetaLF_{key}",
                                         read_param(params, "full2D_thetaLF", read_param(params, "thetaLF", 0.01))))
            alpha = float(read_param(params, f"full2D_alphaLF_{key}", np.max(np.abs(lam))))
            beta = theta * alpha
            Aabs = Ap - Am
            flip = np.fliplr(np.eye(R.n))
            self.char[key] = {"Aplus": Ap, "Aminus": Am, "Aabs": Aabs, "lambda": lam, "beta": beta,
                              "reflect_plus": Ap @ flip, "reflect_minus": Am @ flip}
            o = ops[ax]
            C, F, D, I = _lib.WDG2_CENTER_BANDED, _lib.WDG2_CENTER_FACE, _lib.WDG2_REL_DENSE, _lib.WDG2_REL_IDENTITY
            parts.append(Part(f"{key} central rho transport", ax, C, _lib.WDG2_REL_CENTRAL, s, blocks=o["h1"]))
            if flux == "upwind":
                parts.append(Part(f"{key} interior characteristic upwind", ax, C, D, s, blocks=o["h2i"], dense=Aabs))
            elif beta != 0:
                parts.append(Part(f"{key} interior Rusanov", ax, C, I, s * beta, blocks=o["h2i"]))
            if ax == 1 and self.periodic:
                continue      # no physical Y faces
            # physical boundary penalty H2B (x) |A| = (lift_low e_first^T + lift_high e_last^T)/2 (x) |A|
            parts.append(Part(f"{key}-low physical boundary penalty", ax, F, D, 0.5 * s, face=0,
                              lift=o["lift_low"], dense=Aabs))
            parts.append(Part(f"{key}-high physical boundary penalty", ax, F, D, 0.5 * s, face=1,
                              lift=o["lift_high"], dense=Aabs))
        self.parts = parts

    # -- inflow data (get_Boundary_full2D.m, get_InflowBoundary_full2D.m) --
    def _contact_mask(self, side: str) -> np.ndarray:
        """get_ContactMask_full2D.m:14-90: explicit mask, doping edge, or all."""
        params, mat = self.params, self.mat
        yf = self.Y.nodes
        gy = self.grid_y

        def nearest(vals):
            idx = np.abs(yf[:, None] - gy[None, :]).argmin(axis=1)
            return np.asarray(vals, float)[idx] > 0.5

        explicit = read_param(params, f"full2D_{side}ContactMask")
        if explicit is not None:
            m = np.asarray(explicit).ravel()
            if m.size == yf.size:
                mask = m.astype(bool)
            elif m.size == int(mat["Ny"]):
                mask = nearest(m)
            else:
                raise ValueError("DG:Full2D:ContactMaskSize: contact mask must have N_Y_dof or Ny entries")
        elif not _empty(mat.get("Nd")):
            thr = float(read_param(params, "full2D_contactNdThreshold", 0))
            Nd = self._field_xy(mat["Nd"], 0.0)
            row = Nd[0 if side == "source" else int(mat["Nx"]) - 1, :]
            mask = nearest(row > thr)
        else:
            mask = np.ones(yf.size, bool)
        if not mask.any():
            mask = np.ones(yf.size, bool)
        return mask

    def _field_xy(self, field, default):
        """fieldToXY: scalars, (Nx,Ny), (Ny,Nx), X or Y profiles, first valley of (nv,Nx,Ny)."""
        nx, ny = int(self.mat["Nx"]), int(self.mat["Ny"])
        if _empty(field):
            return np.full((nx, ny), float(default))
        f = np.squeeze(np.asarray(field, float))
        if f.ndim == 0:
            return np.full((nx, ny), float(f))
        if f.shape == (nx, ny):
            return f
        if f.shape == (ny, nx) and nx != ny:
            return f.T
        if f.ndim == 1 and f.size == nx:
            return np.repeat(f[:, None], ny, axis=1)
        if f.ndim == 1 and f.size == ny:
            return np.repeat(f[None, :], nx, axis=0)
        if f.ndim == 3 and f.shape[1:] == (nx, ny):
            return f[0]
        raise ValueError("DG:Full2D:FieldShape: cannot interpret a material field")

    def _reservoir(self, side: str, Ef, Vxy) -> np.ndarray:
        """Material-default contact reservoir (get_DefaultReservoirRho_full2D.m:1-99):
        rho_b(rho_x, rho_y) = int f_FD(k) cos(k_x rho_x) cos(k_y rho_y) dk / (2 pi)^2
        on a midpoint k grid, per Y node of the contact face."""
        params, mat = self.params, self.mat
        if Ef is None or not np.isfinite(Ef) or not read_param(params, "full2D_useMaterialReservoir", True):
            return np.zeros((self.nrel, self.Y.ndof))
        temp = float(mat.get("Temp", 300))
        deg = float(np.atleast_1d(mat.get("deg_factor", 1))[0])
        nkx = int(read_param(params, "full2D_reservoir_Nkx", self.RX.n))
        nky = int(read_param(params, "full2D_reservoir_Nky", self.RY.n))
        kx = -2 * np.pi / self.RX.length * (np.arange(1, nkx + 1) - 0.5 * (nkx + 1))
        ky = -2 * np.pi / self.RY.length * (np.arange(1, nky + 1) - 0.5 * (nky + 1))
        dkx, dky = abs(kx[1] - kx[0]), abs(ky[1] - ky[0])
        cx = np.cos(np.outer(self.RX.cells, kx))
        cy = np.cos(np.outer(self.RY.cells, ky))
        yf = self.Y.nodes
        ix = 0 if side == "source" else int(mat["Nx"]) - 1
        gy = self.grid_y

        def edge(row):
            row = np.asarray(row, float)
            if np.all(row == row[0]):
                return np.full(yf.size, row[0])
            i = np.clip(np.searchsorted(gy, yf) - 1, 0, gy.size - 2)
            t = (yf - gy[i]) / (gy[i + 1] - gy[i])
            return row[i] + t * (row[i + 1] - row[i])

        Vf = self._field_xy(Vxy if Vxy is not None else mat.get("V"), 0.0)
        mx = mass_kg(edge(self._field_xy(mat.get("me_x"), mat.get("me_x_ch", 0.041))[ix]))
        my = mass_kg(edge(self._field_xy(mat.get("me_y"), mat.get("me_y_ch", 0.041))[ix]))
        Ve = edge(Vf[ix])
        out = np.zeros((self.nrel, yf.size))
        for iy in range(yf.size):
            E = Ve[iy] + (HBAR * kx[:, None]) ** 2 / (2 * mx[iy] * QE) + (HBAR * ky[None, :]) ** 2 / (2 * my[iy] * QE)
            arg = np.clip(QE * (E - Ef) / (KB * temp), -80, 80)
            occ = 2 * deg / (1 + np.exp(arg))
            grid = (cx @ occ @ cy.T) * (dkx * dky / (2 * np.pi) ** 2)   # (rho_x, rho_y)
            out[:, iy] = grid.T.ravel()                                  # rho_x fastest
        return out

    def _boundary(self, Vxy):
        params = self.params
        nyd = self.Y.ndof
        self.boundary = {}
        self.reservoir_info = {}
        self.mode_data = {}     # host-solved contact modes, per side (contact-mode reservoirs)
        for side, Ef in (("source", self.EfL), ("drain", self.EfR)):
            direct = read_param(params, f"full2D_{side}Rho")
            if direct is not None:
                d = np.asarray(direct, complex)
                rho = np.repeat(d.reshape(-1, 1), nyd, axis=1) if d.ndim == 1 else d
                if rho.shape != (self.nrel, nyd):
                    raise ValueError("DG:Full2D:BoundaryDataSize: boundary rho has the wrong shape")
            else:
                model = str(read_param(params, "full2D_reservoirModel", "material-default")).lower().replace("_", "-")
                if model in ("material", "material-default", "bulk", "bulk-local", "local-bulk"):
                    rho = self._reservoir(side, Ef, Vxy)
                elif model in ("contact-mode", "contact-modes", "modes", "transverse-modes"):
                    self.mode_data[side] = contact_mode_data(self, side, Ef, Vxy)
                    self.reservoir_info[side] = self.mode_data[side]["info"]
                    rho = None
                    if self.host_reservoirs:
                        rho, _ = contact_mode_reservoir(self, side, Ef, Vxy, data=self.mode_data[side])
                elif model in ("zero", "none", "off", "false"):
                    rho = np.zeros((self.nrel, nyd))
                else:
                    raise NotImplementedError(f"DG:Full2D:Unsupported: reservoir model '{model}'")
            mask = self._contact_mask(side)
            if rho is not None:
                rho = np.array(rho, complex)
                rho[:, ~mask] = 0
            self.boundary[side] = {"rho": rho, "mask": mask}
        sx, sy = self.scales
        ox = self.axis_ops[0]
        chx = self.char["X"]
        # closed (reflecting) pieces of the X faces (get_Diff_full2D.m:393-432)
        for side, face, lift, refl, coef in (("source", 0, ox["lift_low"], chx["reflect_plus"], -sx),
                                             ("drain", 1, ox["lift_high"], chx["reflect_minus"], sx)):
            closed = ~self.boundary[side]["mask"]
            if closed.any():
                self.parts.append(Part(f"X-{side} closed-face reflection correction", 0, _lib.WDG2_CENTER_FACE,
                                       _lib.WDG2_REL_DENSE, coef, face=face, lift=lift,
                                       mask=closed.astype(np.int32), dense=refl))
        if self.ytype == "specular":
            oy, chy = self.axis_ops[1], self.char["Y"]
            self.parts.append(Part("Y-bottom specular reflection correction", 1, _lib.WDG2_CENTER_FACE,
                                   _lib.WDG2_REL_DENSE, -sy, face=0, lift=oy["lift_low"], dense=chy["reflect_plus"]))
            self.parts.append(Part("Y-top specular reflection correction", 1, _lib.WDG2_CENTER_FACE,
                                   _lib.WDG2_REL_DENSE, sy, face=1, lift=oy["lift_high"], dense=chy["reflect_minus"]))

    def rhs(self) -> np.ndarray:
        """Inflow RHS (addSourceBoundary/addDrainBoundary rhs parts): on every
        contact Y node, lift (x) (+-A^+- rho_b) into the first/last X element."""
        sx = self.scales[0]
        nxd, nyd, nl = self.X.ndof, self.Y.ndof, self.X.nl
        out = np.zeros((nyd, nxd, self.nrel), complex)
        ch = self.char["X"]
        ox = self.axis_ops[0]
        for side, A, sign, lift, cols in (("source", ch["Aplus"], 1.0, ox["lift_low"], slice(0, nl)),
                                          ("drain", ch["Aminus"], -1.0, ox["lift_high"], slice(nxd - nl, nxd))):
            b = self.boundary[side]
            if b["rho"] is None:
                raise ValueError("the reservoirs of this setup are built on the device (host_reservoirs=False): "
                                 "use Full2DOperator.upload_contact_rhs")
            for iy in np.nonzero(b["mask"])[0]:
                rb = b["rho"][:, iy].reshape(self.RY.n, self.RX.n)      # (rho_y, rho_x)
                g = (rb @ A.T).ravel() * sign
                out[iy, cols, :] += sx * lift[:, None] * g[None, :]
        return out.ravel()

    # -- potential and CAP (get_Drift_full2D.m) --
    def _potential(self, Vxy):
        params, mat = self.params, self.mat
        mode = str(read_param(params, "full2D_potentialDiscretization", "fv-consistent")).strip().lower()
        if mode not in ("fv-consistent", "collocated"):
            raise ValueError(f"DG:Full2D:UnknownPotentialDiscretization: {mode}")
        self.pot_mode = mode
        if Vxy is None:
            Vxy = mat.get("V")
        Vf = self._field_xy(Vxy, 0.0)
        ix, iy = np.argsort(self.grid_x), np.argsort(self.grid_y)
        self.vgrid_x, self.vgrid_y = self.grid_x[ix], self.grid_y[iy]
        self.vfield = np.ascontiguousarray(Vf[np.ix_(ix, iy)])
        q = QE / HBAR
        qd = float(read_param(params, "full2D_Q_drift", q))
        ds = float(read_param(params, "full2D_drift_scale", 1))
        cs = float(read_param(params, "full2D_cap_scale", 1))
        self.pot_coef = ds * qd           # multiplies i * DeltaV
        self.cap_coef = ds * cs * qd
        self.cap_x = cap_profile(self.RX.cells, self.RX.length, self.RX.delta)
        self.cap_y = cap_profile(self.RY.cells, self.RY.length, self.RY.delta)
        if self.periodic:
            self.cap_y = np.zeros_like(self.cap_y)     # get_CAP_full2D.m:14-16
            if np.max(np.abs(self.vfield[:, 0] - self.vfield[:, -1])) > 1e-10 * max(1.0, np.abs(self.vfield).max()):
                raise ValueError("DG:Full2D:NonperiodicPotential: the first and last material Y columns must "
                                 "represent the same periodic potential")
        X, Y = self.X, self.Y
        if mode == "fv-consistent":
            order = read_param(params, "full2D_potentialQuadratureOrder")
            nqx = order if order is not None else read_param(params, "full2D_potentialQuadratureX", max(4, 2 * X.nl - 1))
            nqy = order if order is not None else read_param(params, "full2D_potentialQuadratureY", max(4, 2 * Y.nl - 1))
            nqx, nqy = max(X.nl, int(round(nqx))), max(Y.nl, int(round(nqy)))
            self.nq = (nqx, nqy)
            self.quad = []
            self.interp, self.proj = [], []
            for ax, nq in ((X, nqx), (Y, nqy)):
                rq, wq = lgl(nq)
                Vq, _ = onb_vandermonde(ax.nl, rq)
                ip = Vq @ np.linalg.inv(ax.V)                       # (nq, nl)
                Mi = np.linalg.inv(ax.M)
                self.interp.append(ip)
                self.proj.append((ip * wq[:, None]) @ Mi.T)         # (nq, nl)
                mid = 0.5 * (ax.vertices[:-1] + ax.vertices[1:])
                self.quad.append((mid[:, None] + ax.jac[:, None] * rq[None, :]).ravel())
            ey_ = self.RY.edges.copy()
            if self.periodic:
                ey_[-1] = ey_[0]    # the last vertex row is the first one (unique periodic vertices)
            self.rel_pts = (self.RX.edges, ey_)
        else:
            self.nq = (0, 0)
            self.quad = [X.nodes, Y.nodes]
            self.interp = self.proj = None
            self.rel_pts = (self.RX.cells, self.RY.cells)

    # -- observables (get_Observables_full2D.m:81-135) --
    @staticmethod
    def zero_weights(cells):
        cells = np.asarray(cells, float)
        near = int(np.argmin(np.abs(cells)))
        tol = max(1e-14, 1e-12 * max(1.0, np.abs(cells).max()))
        if abs(cells[near]) <= tol:
            return [near], [1.0]
        left, right = np.nonzero(cells < 0)[0], np.nonzero(cells > 0)[0]
        if left.size and right.size:
            l, r = int(left[-1]), int(right[0])
            h = cells[r] - cells[l]
            return [l, r], [cells[r] / h, -cells[l] / h]
        return [near], [1.0]

    @staticmethod
    def derivative_row(cells, delta):
        """Sparse derivative at zero separation: the central-difference row of
        the zero cell, else the bracketing two-point slope."""
        cells = np.asarray(cells, float)
        n = cells.size
        near = int(np.argmin(np.abs(cells)))
        tol = max(1e-14, 1e-12 * max(1.0, np.abs(cells).max()))
        if abs(cells[near]) <= tol:
            ids, w = [], []
            if near > 0:
                ids.append(near - 1); w.append(-1 / (2 * delta))
            if near < n - 1:
                ids.append(near + 1); w.append(1 / (2 * delta))
            return ids, w
        left, right = np.nonzero(cells < 0)[0], np.nonzero(cells > 0)[0]
        if left.size and right.size:
            l, r = int(left[-1]), int(right[0])
        elif near == 0:
            l, r = 0, 1
        elif near == n - 1:
            l, r = n - 2, n - 1
        else:
            l, r = near - 1, near + 1
        h = cells[r] - cells[l]
        return [l, r], [-1 / h, 1 / h]

    def observable_weights(self):
        """(index, weight) lists in the relative block for n, d/drho_x, d/drho_y."""
        nx = self.RX.n
        ixs, wxs = self.zero_weights(self.RX.cells)
        iys, wys = self.zero_weights(self.RY.cells)
        dxi, dxw = self.derivative_row(self.RX.cells, self.RX.delta)
        dyi, dyw = self.derivative_row(self.RY.cells, self.RY.delta)
        n = [(i + nx * j, wi * wj) for i, wi in zip(ixs, wxs) for j, wj in zip(iys, wys)]
        dx = [(i + nx * j, wi * wj) for i, wi in zip(dxi, dxw) for j, wj in zip(iys, wys)]
        dy = [(i + nx * j, wi * wj) for i, wi in zip(ixs, wxs) for j, wj in zip(dyi, dyw)]
        return n, dx, dy

    def current_scales(self):
        mat = self.mat
        mx = float(mass_kg(mat.get("me_x_ch", 0.041)))
        my = float(mass_kg(mat.get("me_y_ch", mat.get("me_x_ch", 0.041))))
        return QE * HBAR / mx, QE * HBAR / my


# ---------------------------------------------------------------------------
# Device operator
# ---------------------------------------------------------------------------

def _f64(a):
    return np.ascontiguousarray(np.asarray(a, np.float64))


def _c128(a):
    return np.ascontiguousarray(np.asarray(a, np.complex128))


class Full2DOperator:
    """The full-2D ``sysInfo`` handle surface on the GPU: ``apply``,
    ``assembleRhs``, ``getDiagonal``, ``getRowAbsSum`` (get_SysM_full2D.m:80-86)
    plus ``solve`` (runGmresWithProgress, solve_transport_DG_full2D.m:422-506)
    and ``observables``."""

    def __init__(self, setup: Full2DSetup, transport: bool = True, drift: bool = True):
        lib = _lib.load()
        self._lib = lib
        self.setup = s = setup
        self.drift_on = drift
        self.desc = _lib.wdg2_desc(s.RX.n, s.RY.n, s.X.nl, s.Y.nl, s.X.ne, s.Y.ne, s.RX.delta, s.RY.delta,
                                   1 if s.periodic else 0)
        keep = []
        parts = s.parts if transport else []
        cparts = (_lib.wdg2_part * max(len(parts), 1))()
        for i, p in enumerate(parts):
            cp = cparts[i]
            cp.axis, cp.center, cp.rel, cp.face, cp.coef = p.axis, p.center, p.rel, p.face, float(p.coef)
            if p.center == _lib.WDG2_CENTER_BANDED:
                arrs = [_f64(p.blocks.diag), _f64(p.blocks.lower), _f64(p.blocks.upper)]
                cp.c_diag, cp.c_lower, cp.c_upper = (dptr(a) for a in arrs)
            else:
                arrs = [_f64(p.lift)]
                cp.c_diag = dptr(arrs[0])
                if p.mask is not None:
                    m = np.ascontiguousarray(np.asarray(p.mask, np.int32))
                    arrs.append(m)
                    cp.mask = m.ctypes.data_as(C.POINTER(C.c_int32))
            if p.rel == _lib.WDG2_REL_DENSE:
                d = _c128(p.dense)
                arrs.append(d)
                cp.r_dense = dptr(d)
            keep.extend(arrs)
        ctx = C.c_void_p()
        check(lib.wdg2_create(C.byref(self.desc), cparts, len(parts), C.byref(ctx)))
        self.ctx = ctx
        self.n = s.n_total
        self.rhs_norm = 0.0
        self.set_potential()
        self._rhs = None

    def close(self):
        if getattr(self, "ctx", None):
            self._lib.wdg_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_potential(self, Vxy=None):
        """Upload V on the material grid, the quadrature factors and the CAP
        (get_Drift_full2D.m:133-209); Vxy replaces the field if given."""
        s = self.setup
        if Vxy is not None:
            s._vxy = Vxy
            s._potential(Vxy)
        fv = s.pot_mode == "fv-consistent"
        keep = [_f64(s.vgrid_x), _f64(s.vgrid_y), _f64(s.vfield.T), _f64(s.quad[0]), _f64(s.quad[1]),
                _f64(s.rel_pts[0]), _f64(s.rel_pts[1]), _f64(s.cap_x), _f64(s.cap_y)]
        if fv:
            keep += [_f64(s.interp[0]), _f64(s.interp[1]), _f64(s.proj[0]), _f64(s.proj[1])]
        pot = _lib.wdg2_potential()
        pot.mode = _lib.WDG2_POT_FV_CONSISTENT if fv else _lib.WDG2_POT_COLLOCATED
        pot.nq_x, pot.nq_y = s.nq
        pot.nx_grid, pot.ny_grid = s.vgrid_x.size, s.vgrid_y.size
        (pot.x_grid, pot.y_grid, pot.v, pot.center_x, pot.center_y, pot.rel_x, pot.rel_y,
         pot.cap_x, pot.cap_y) = (dptr(a) for a in keep[:9])
        if fv:
            pot.interp_x, pot.interp_y, pot.proj_x, pot.proj_y = (dptr(a) for a in keep[9:])
        pot.pot_coef, pot.cap_coef = (s.pot_coef, s.cap_coef) if self.drift_on else (0.0, 0.0)
        check(self._lib.wdg2_set_potential(self.ctx, C.byref(pot)), self.ctx)

    def assemble_rhs(self) -> np.ndarray:
        if self._rhs is None:
            self._rhs = _c128(self.setup.rhs())
        return self._rhs

    def upload_rhs(self, rhs=None):
        r = _c128(self.assemble_rhs() if rhs is None else rhs).reshape(-1)
        if r.size != self.n:
            raise ValueError("rhs has the wrong size")
        check(self._lib.wdg2_set_rhs(self.ctx, dptr(r)), self.ctx)
        self.rhs_norm = float(np.linalg.norm(r))

    def device_rhs(self) -> np.ndarray:
        """The inflow rhs the device holds, in physical units (wdg_rhs)."""
        r = np.empty(self.n, np.complex128)
        check(self._lib.wdg_rhs(self.ctx, dptr(r)), self.ctx)
        return r

    def upload_contact_rhs(self, want_reservoirs: bool = False):
        """Build the contact-mode reservoirs and the inflow rhs on the device
        (wdg2_set_contact_rhs) from the setup's host-solved modes; the
        alternative to upload_rhs for the contact-mode reservoir model.
        Returns the two reservoirs (rho_x*rho_y x Y dof) when asked."""
        s = self.setup
        if set(s.mode_data) != {"source", "drain"}:
            raise ValueError("upload_contact_rhs needs contact-mode reservoirs on both contacts "
                             "(full2D_reservoirModel = 'contact-modes', no explicit full2D_<side>Rho)")
        keep = []
        nrx, nyd = s.RX.n, s.Y.ndof
        ch, ox = s.char["X"], s.axis_ops[0]
        structs = []
        outs = {}
        for side, A, lift in (("source", ch["Aplus"], ox["lift_low"]), ("drain", ch["Aminus"], ox["lift_high"])):
            d = s.mode_data[side]
            cm = _lib.wdg2_contact_modes()
            A = _c128(np.asarray(A, complex).reshape(nrx, nrx))
            lift = _f64(np.asarray(lift, float))
            mask = np.ascontiguousarray(s.boundary[side]["mask"].astype(np.int32))
            keep += [A, lift, mask]
            cm.mask = mask.ctypes.data_as(C.POINTER(C.c_int32))
            cm.char_matrix = dptr(A)
            cm.lift = dptr(lift)
            if np.isfinite(d["Ef"]) and d["energy"].size:
                contact = d["contact"]
                nm = d["energy"].size
                cm.n_modes = nm
                if contact["periodic"]:
                    cm.kind = _lib.WDG2_MODES_FOURIER
                    grid = _f64(contact["k"])
                    vals = _c128(np.asarray(contact["coef"], complex).T)          # mode-major
                    cm.y0, cm.width = float(contact["gridY"][0]), float(contact["width"])
                else:
                    cm.kind = _lib.WDG2_MODES_BOX
                    gy = contact["gridY"]
                    dy = float(np.mean(np.diff(gy)))
                    grid = _f64(np.concatenate(([gy[0] - dy], gy, [gy[-1] + dy])))
                    modes = np.real(np.asarray(d["modes"]))
                    vals = _f64(np.concatenate((np.zeros((1, nm)), modes, np.zeros((1, nm))), axis=0).T)
                cm.n_grid = grid.size
                e, mx, mz = _f64(np.real(d["energy"])), _f64(d["mass_x"]), _f64(d["mass_z"])
                keep += [grid, vals, e, mx, mz]
                cm.grid, cm.values = dptr(grid), dptr(vals)
                cm.energy, cm.mass_x, cm.mass_z = dptr(e), dptr(mx), dptr(mz)
                cm.fermi_level = float(d["Ef"])
            if want_reservoirs:
                outs[side] = np.zeros((nyd, s.nrel), np.complex128)
                cm.reservoir_out = dptr(outs[side])
            structs.append(cm)
        d0 = s.mode_data["source"]
        g = _lib.wdg2_reservoir_grid()
        kx = _f64(d0["kx"])
        kz = d0.get("kz", {"integrate": False})
        kzv = _f64(kz["k"]) if kz.get("integrate") else None
        rx, ry, yn = _f64(s.RX.cells), _f64(s.RY.cells), _f64(s.Y.nodes)
        keep += [kx, rx, ry, yn] + ([kzv] if kzv is not None else [])
        g.temp, g.deg_factor = float(d0["temp"]), float(d0["deg"])
        g.n_kx, g.kx, g.dkx = kx.size, dptr(kx), float(d0["dkx"])
        if kzv is not None:
            g.n_kz, g.kz = kzv.size, dptr(kzv)
        g.rho_x, g.rho_y, g.y_nodes = dptr(rx), dptr(ry), dptr(yn)
        g.scale_x = float(s.scales[0])
        check(self._lib.wdg2_set_contact_rhs(self.ctx, C.byref(g), C.byref(structs[0]), C.byref(structs[1])),
              self.ctx)
        self._rhs = None
        self.rhs_norm = float(np.linalg.norm(self.device_rhs()))
        if want_reservoirs:
            return {side: outs[side].T.copy() for side in outs}
        return None

    def apply(self, u, precision: str = "fp64") -> np.ndarray:
        u = _as_c128(u, self.n)
        y = np.empty_like(u)
        prec = {"fp64": _lib.WDG_PREC_FP64, "mixed": _lib.WDG_PREC_MIXED}[precision]
        check(self._lib.wdg_apply(self.ctx, dptr(u), dptr(y), prec), self.ctx)
        return y

    def diagonal(self) -> np.ndarray:
        d = np.empty(self.n, np.complex128)
        check(self._lib.wdg_diagonal(self.ctx, dptr(d)), self.ctx)
        return d

    def set_dv_mode(self, mode: str):
        """'table' (Delta V streamed from a stored table, the default) or
        'interp' (Delta V interpolated in the apply): wdg2_set_dv_mode."""
        modes = {"interp": _lib.WDG2_DV_INTERP, "table": _lib.WDG2_DV_TABLE}
        if mode not in modes:
            raise ValueError(f"mode must be one of {sorted(modes)}")
        check(self._lib.wdg2_set_dv_mode(self.ctx, modes[mode]), self.ctx)

    def row_abs_sum(self) -> np.ndarray:
        d = np.empty(self.n, np.float64)
        check(self._lib.wdg2_rowabs(self.ctx, dptr(d)), self.ctx)
        return d

    def set_preconditioner(self, kind: str) -> str:
        """'jacobi' falls back to 'rowabs' on small pivots, as
        buildPreconditioner does (solve_transport_DG_full2D.m:788-925).
        Returns the method in use."""
        if kind not in PRECONDITIONERS:
            raise ValueError(f"DG:Full2D:UnknownPreconditioner: {kind}")
        used = C.c_int32(0)
        check(self._lib.wdg2_set_precond(self.ctx, PRECONDITIONERS[kind], C.byref(used)), self.ctx)
        return {v: k for k, v in PRECONDITIONERS.items()}[used.value]

    def set_line_transverse(self) -> dict:
        """Build the line-transverse factors on the host (the reference's CPU
        setup) and select them for wdg_solve; returns their summary."""
        lf = line_transverse_factors(self)
        self._line = lf   # keep the arrays alive while the library copies them
        f = _lib.wdg2_line_factors()
        f.sigma = lf["sigma"]
        f.modes = dptr(lf["modes"])
        f.forward = lf["forward"].ctypes.data_as(C.POINTER(C.c_int32))
        f.x_inverse = dptr(lf["x_inverse"])
        f.x_coupling = dptr(lf["x_coupling"])
        f.t_inverse = dptr(lf["t_inverse"])
        check(self._lib.wdg2_set_line_precond(self.ctx, C.byref(f)), self.ctx)
        return {k: lf[k] for k in ("sigma", "factor_bytes", "transverse_dof", "modes_count")}

    def transient_cn(self, dt: float, nsteps: int, a_scale: float = 1.0, b_scale: float = 1.0, tol: float = 1e-10,
                     restart: int = 30, precision: str = "fp64", rho0=None, return_rho: bool = True,
                     use_graph: bool = True, krylov_basis: str = "fp16"):
        """Crank-Nicolson steps of d rho/dt = b_scale*rhs - a_scale*A rho
        (wdg_transient_cn); the Jacobi / row-sum preconditioner in use is
        rebuilt for the shifted operator.  use_graph=False launches the restart
        cycles kernel by kernel instead of replaying captured graphs;
        krylov_basis is the mixed solver's basis storage ('fp16', 'bf16',
        'fp32').  Returns (rho, info)."""
        from .solver import transient_cn
        return transient_cn(self._lib, self.ctx, self.n, dt, nsteps, a_scale, b_scale, tol, restart, precision,
                            rho0, return_rho, use_graph=use_graph, krylov_basis=krylov_basis)

    def set_profiling(self, mask: int):
        check(self._lib.wdg_set_profiling(self.ctx, int(mask)), self.ctx)

    def profile(self) -> dict:
        from .solver import read_profile
        return read_profile(self._lib, self.ctx)

    def precond(self, r) -> np.ndarray:
        r = _as_c128(r, self.n)
        z = np.empty_like(r)
        check(self._lib.wdg_precond(self.ctx, dptr(r), dptr(z)), self.ctx)
        return z

    def solve(self, opts: SolveOptions | None = None, x0=None, return_x: bool = True):
        opts = opts or SolveOptions()
        info = _lib.wdg_solve_info()
        x = np.empty(self.n, np.complex128) if return_x else None
        x0p = None
        if x0 is not None:
            x0 = validate_initial_rho(x0, self.n)
            x0p = dptr(x0)
        check(self._lib.wdg_solve(self.ctx, C.byref(opts.to_c()), x0p, dptr(x) if return_x else None,
                                  C.byref(info)), self.ctx)
        return x, info

    def resvec(self) -> np.ndarray:
        """Residual history of the last solve relative to ||rhs|| ([0] initial,
        [k] after Arnoldi step k; wdg_solve_resvec)."""
        from .solver import read_resvec
        return read_resvec(self._lib, self.ctx)

    def observables(self, rho=None):
        """nDG, jxDG, jyDG on the (X dof, Y dof) grid (get_Observables_full2D.m:24-60);
        rho=None uses the device solution of the last solve."""
        rho = None if rho is None else _as_c128(rho, self.n)
        s = self.setup
        wn, wx, wy = s.observable_weights()
        idx = np.zeros(3 * 8, np.int32)
        w = np.zeros(3 * 8, np.float64)
        cnt = np.zeros(3, np.int32)
        for k, lst in enumerate((wn, wx, wy)):
            if len(lst) > 8:
                raise ValueError("too many zero-separation weights")
            cnt[k] = len(lst)
            for i, (j, v) in enumerate(lst):
                idx[8 * k + i] = j
                w[8 * k + i] = v
        out = np.empty(3 * s.ncen, np.complex128)
        rp = None
        if rho is not None:
            rp = dptr(rho)
        check(self._lib.wdg2_observables(self.ctx, rp, idx.ctypes.data_as(C.POINTER(C.c_int32)), dptr(w),
                                         cnt.ctypes.data_as(C.POINTER(C.c_int32)), dptr(out)), self.ctx)
        out = out.reshape(3, s.Y.ndof, s.X.ndof)
        cx, cy = s.current_scales()
        nDG = out[0].real.T
        jxDG = cx * out[1].imag.T
        jyDG = cy * out[2].imag.T
        return nDG, jxDG, jyDG


def _probe_block(op: "Full2DOperator", ids: np.ndarray) -> np.ndarray:
    """A[ids][:, ids] of a device operator by applying it to unit vectors."""
    out = np.zeros((ids.size, ids.size), complex)
    e = np.zeros(op.n, complex)
    for j, col in enumerate(ids):
        e[col] = 1.0
        out[:, j] = op.apply(e)[ids]
        e[col] = 0.0
    return out


def transverse_reference(setup: Full2DSetup, ix_ref: int):
    """transverseReference (get_LineTransversePreconditioner_full2D.m:128-172):
    the (rho_y, Y) blocks K (transport) and V (potential) of a two-cell rho_x,
    two-element X strip carrying the potential of material column ix_ref, with
    no X transport, no CAP and zero reservoirs, at rho_x cell 1 and X DOF 1.
    The blocks are read off the device operators of that strip."""
    import copy
    mat = setup.mat
    ref = copy.deepcopy(mat)
    Vf = setup._field_xy(mat.get("V") if setup._vxy is None else setup._vxy, 0.0)
    ref["V"] = np.repeat(Vf[ix_ref][None, :], int(mat["Nx"]), axis=0)
    q = dict(dg_params(mat))
    nyd, nry = setup.Y.ndof, setup.RY.n
    q.update({"N_rho_x": 2, "N_K_X": 2, "N_el_X": 2, "full2D_Q_diff_x": 0, "full2D_cap_scale": 0,
              "full2D_reservoirModel": "zero", "full2D_sourceRho": np.zeros((2 * nry, nyd)),
              "full2D_drainRho": np.zeros((2 * nry, nyd))})
    for key in ("full2D_sourceContactMask", "full2D_drainContactMask"):
        q.pop(key, None)
    ref.setdefault("dg", {})
    ref["dg"] = dict(ref["dg"])
    ref["dg"]["params"] = q
    strip = Full2DSetup(ref, ref["V"], None, None)
    ids = np.array([2 * (ry + nry * (strip.X.ndof * iyd)) for iyd in range(nyd) for ry in range(nry)])
    with Full2DOperator(strip, transport=True, drift=False) as opk:
        K = _probe_block(opk, ids)
    with Full2DOperator(strip, transport=False, drift=True) as opv:
        V = _probe_block(opv, ids)
    if strip.pot_mode == "fv-consistent":
        V = 2 * V       # the rho_x diagonal of P_x^T P_x / 4 on two cells is 1/2
    return K, V


def line_transverse_factors(op: "Full2DOperator") -> dict:
    """Host setup of get_LineTransversePreconditioner_full2D.m:17-111: memory
    guard, modes of -i D_rhoX, frozen-X transverse blocks, CAP and potential
    mode weights, shift sigma, X factors and the per-mode transverse inverses."""
    s = op.setup
    params = s.params
    nrx, nl, ne = s.RX.n, s.X.nl, s.X.ne
    nT = s.RY.n * s.Y.ndof
    factor_bytes = 32 * nT ** 2 * nrx + 32 * nl ** 2 * nrx * ne + 16 * nrx ** 2 + 8 * nT * nrx
    max_bytes = read_param(params, "full2D_lineMaxFactorBytes", 4 * 1024 ** 3)
    if not (np.isscalar(max_bytes) and np.isfinite(max_bytes) and max_bytes > 0):
        raise _lib.WdgError(_lib.WDG_ERR_LINE_MEMORY, "lineMaxFactorBytes must be positive and finite.")
    if factor_bytes > max_bytes:
        raise _lib.WdgError(_lib.WDG_ERR_LINE_MEMORY, f"Estimated factors {factor_bytes / 2**30:.3g} GiB exceed "
                          f"full2D_lineMaxFactorBytes {max_bytes / 2**30:.3g} GiB")
    ix_ref = read_param(params, "full2D_lineReferenceXIndex", 1)
    if not (float(ix_ref).is_integer() and 1 <= ix_ref <= int(s.mat["Nx"])):
        raise ValueError("DG:Full2D:LineReferenceIndex: reference X index must be an integer in 1:mat.Nx")
    U, lam = rho_modes(nrx, s.RX.delta)
    lam = lam.copy()
    lam[np.abs(lam) < 100 * np.spacing(np.max(np.abs(lam)))] = 0
    K, V = transverse_reference(s, int(ix_ref) - 1)
    capY = np.tile(s.cap_y, s.Y.ndof) * s.cap_coef
    capX = np.real(np.sum(np.conj(U) * (s.cap_x[:, None] * U), axis=0)) * s.cap_coef
    if s.pot_mode == "fv-consistent":
        Px = np.zeros((nrx + 1, nrx))
        Px[np.arange(nrx), np.arange(nrx)] = 1
        Px[np.arange(1, nrx + 1), np.arange(nrx)] = 1
        weight = np.real(np.sum(np.conj(U) * ((Px.T @ Px / 4) @ U), axis=0))
    else:
        weight = np.ones(nrx)
    Qx = s.scales[0]
    xrate = abs(Qx) * np.max(np.abs(lam)) / np.min(s.X.jac)
    trate = (np.abs(K).sum(axis=1).max() + np.abs(V).sum(axis=1).max() + np.max(np.abs(capY))
             + np.max(np.abs(capX)))
    rel_shift = read_param(params, "full2D_lineRelativeShift", 0.003)
    sigma = float(read_param(params, "full2D_lineShift", rel_shift * max(xrate, trate, np.finfo(float).tiny)))
    if not (np.isfinite(sigma) and sigma > 0):
        raise ValueError("DG:Full2D:LineShift: the preconditioner shift must be positive, real and finite")
    # xFactors: per mode k and element e, (I + a K1 + |a| K2)^-1 and the upwind coupling
    Mi = np.linalg.inv(s.X.M)
    e0 = np.zeros((nl, nl)); e0[0, 0] = 1
    en = np.zeros((nl, nl)); en[-1, -1] = 1
    K1 = Mi @ (s.X.S + 0.5 * e0 - 0.5 * en)
    K2 = Mi @ (0.5 * e0 + 0.5 * en)
    speeds = Qx * lam
    forward = (speeds >= 0).astype(np.int32)
    xinv = np.zeros((nrx, ne, nl, nl))
    xcoup = np.zeros((nrx, ne, nl, nl))
    for k in range(nrx):
        for e in range(ne):
            a = speeds[k] / (sigma * s.X.jac[e])
            xinv[k, e] = np.linalg.inv(np.eye(nl) + a * K1 + abs(a) * K2)
            if forward[k]:
                xcoup[k, e][:, nl - 1] = -a * Mi[:, 0]
            else:
                xcoup[k, e][:, 0] = a * Mi[:, nl - 1]
    tinv = np.empty((nrx, nT, nT), complex)
    for k in range(nrx):
        B = (K + weight[k] * V + np.diag(capY + capX[k])) / sigma + np.eye(nT)
        tinv[k] = np.linalg.inv(B)
    return {"sigma": sigma, "modes": np.ascontiguousarray(U), "forward": forward,
            "x_inverse": np.ascontiguousarray(xinv), "x_coupling": np.ascontiguousarray(xcoup),
            "t_inverse": np.ascontiguousarray(tinv), "factor_bytes": factor_bytes, "transverse_dof": nT,
            "modes_count": nrx}


def project_to_material_y(values_dg, y_dg, target_y):
    """projectValuesToY (get_Observables_full2D.m:161-174): duplicate DG Y
    nodes averaged, then linear inter/extrapolation."""
    uy, group = np.unique(np.asarray(y_dg), return_inverse=True)
    avg = np.stack([values_dg[:, group == k].mean(axis=1) for k in range(uy.size)], axis=1)
    if uy.size == 1:
        return np.repeat(avg, np.size(target_y), axis=1)
    t = np.asarray(target_y, float)
    i = np.clip(np.searchsorted(uy, t) - 1, 0, uy.size - 2)
    a = (t - uy[i]) / (uy[i + 1] - uy[i])
    return avg[:, i] * (1 - a) + avg[:, i + 1] * a


@dataclass
class Full2DResult:
    status: str
    rho: np.ndarray
    n: np.ndarray
    jx: np.ndarray
    jy: np.ndarray
    j: np.ndarray
    nDG: np.ndarray
    jxDG: np.ndarray
    jyDG: np.ndarray
    info: dict = field(default_factory=dict)
    setup: "Full2DSetup | None" = None    # the host setup of the solve (grids, DG axes, reservoirs)


def validate_initial_rho(rho0, n_total: int) -> np.ndarray:
    """rho0 as a flat complex vector of n_total finite entries, or
    DG:Full2D:InvalidInitialRho (validateInitialRho,
    DG_code/solve_transport_DG_full2D.m:508-525: a full vector with nTotal
    entries in rho_x, rho_y, X, Y order, never a reshaped 4-D array; finite)."""
    a = np.asarray(rho0)
    if a.dtype.kind not in "fc" or (a.ndim > 1 and sum(d > 1 for d in a.shape) > 1) or a.ndim > 2 \
            or a.size != n_total:
        raise _lib.WdgError(_lib.WDG_ERR_INITIAL_RHO, f"rho0 must be a full single/double vector with {n_total} "
                                                      "entries in rho_x,rho_y,X,Y order.")
    a = np.ascontiguousarray(a.reshape(-1), dtype=np.complex128)
    if not np.all(np.isfinite(a)):
        raise _lib.WdgError(_lib.WDG_ERR_INITIAL_RHO, "rho0 must contain only finite values.")
    return a


def gmres_basis_guard(params: dict, n_total: int, restart: int, pmode: str):
    """DG:Full2D:GmresMemoryLimit of solve_transport_DG_full2D.m:183-203: with
    the line-transverse preconditioner the estimated basis 16 nTotal
    (min(restart, nTotal) + 1) bytes may not exceed full2D_gmresMaxBasisBytes
    (9 GiB on the GPU).  Returns (estimate, limit)."""
    est = 16 * n_total * (min(restart, n_total) + 1)
    limit = float("inf")
    if pmode == "line-transverse":
        limit = float(read_param(params, "full2D_gmresMaxBasisBytes", 9 * 1024 ** 3))
        if est > limit:
            raise _lib.WdgError(_lib.WDG_ERR_GMRES_MEMORY,
                                f"Estimated GMRES basis {est / 1024 ** 3:.3g} GiB exceeds full2D_gmresMaxBasisBytes. "
                                "Reduce restart or explicitly increase the memory budget.")
    return est, limit


def _transport_options(setup, solver: SolveOptions | None):
    """Preconditioner mode and GMRES options of a full-2D solve from the
    setup's parameters (solve_transport_DG_full2D.m:132-308)."""
    params = setup.params
    if str(read_param(params, "full2D_solver", "gmres")).lower() not in ("gmres", "auto"):
        raise NotImplementedError("DG:Full2D:Unsupported: only the GMRES solver runs on the device")
    pmode = str(read_param(params, "full2D_preconditioner", "auto")).strip().lower()
    pmode = {"auto": "jacobi", "diagonal": "jacobi", "off": "none", "false": "none", "no": "none",
             "rowsum": "rowabs", "row-sum": "rowabs", "row-scaling": "rowabs",
             "blockjacobi": "rowabs", "block-jacobi": "rowabs", "rho-block-jacobi": "rowabs",
             "relative-block-jacobi": "rowabs", "rho-blockjacobi": "rowabs"}.get(pmode, pmode)
    if solver is None:
        # MATLAB's gmres(restart) counts full2D_solverMaxIt in restart cycles
        restart = int(read_param(params, "full2D_gmresRestart", read_param(params, "full2D_gpuGmresRestart", 20)))
        solver = SolveOptions(tol=float(read_param(params, "full2D_solverTol", 1e-4)), restart=restart,
                              max_iter=int(read_param(params, "full2D_solverMaxIt", 1400)) * restart)
    return pmode, solver


def _set_preconditioner(op, pmode: str) -> str:
    if pmode == "line-transverse":
        op.set_line_transverse()
        return pmode
    return op.set_preconditioner(pmode)


def _configure(op, setup, pmode: str, basis_limit: float) -> str:
    """Memory guards, preconditioner and inflow rhs of a fresh operator."""
    params = setup.params
    if np.isfinite(basis_limit):
        check(op._lib.wdg_set_memory_limits(op.ctx, basis_limit,
                                            float(read_param(params, "full2D_lineMaxFactorBytes", 4 * 1024 ** 3))),
              op.ctx)
    used = _set_preconditioner(op, pmode)
    if set(setup.mode_data) == {"source", "drain"}:
        op.upload_contact_rhs()      # contact-mode reservoirs and the inflow rhs built on the device
    else:
        op.upload_rhs()
    return used


def _result(op, setup, rho, info, warm: bool, used: str) -> Full2DResult:
    """The reference's outputs of one solve (solve_transport_DG_full2D.m:67-129,310-334)."""
    # resvec * ||rhs|| as the reference reports it (solve_transport_DG_full2D.m:324,333)
    resvec = op.resvec() * (op.rhs_norm if op.rhs_norm > 0 else 1.0)
    nDG, jxDG, jyDG = op.observables()
    target_y = setup.grid_y if "Ny" in setup.mat else setup.Y.nodes
    n = project_to_material_y(nDG, setup.Y.nodes, target_y)
    jx = project_to_material_y(jxDG, setup.Y.nodes, target_y)
    jy = project_to_material_y(jyDG, setup.Y.nodes, target_y)
    j = np.trapezoid(jx, target_y, axis=1) if np.size(target_y) > 1 else jx[:, 0]
    status = "solved-gmres" if info.flag == 0 else "gmres-not-converged"
    return Full2DResult(status, rho, n, jx, jy, j, nDG, jxDG, jyDG,
                        {"solve": {"flag": info.flag, "relres": info.true_relres, "trueRelres": info.true_relres,
                                   "initialTrueRelres": info.initial_true_relres, "warmStartUsed": warm,
                                   "iterations": info.iterations, "resvec": resvec, "preconditioner": used,
                                   "seconds": info.seconds}}, setup)


def solve_transport_DG_full2D(mat: dict, Vxy=None, EfL=None, EfR=None, rho0=None,
                              solver: SolveOptions | None = None) -> Full2DResult:
    """Stationary full-2D solve with the reference's signature and outputs
    (solve_transport_DG_full2D.m:1,67-129): GMRES on the device with the
    preconditioner named by full2D_preconditioner (auto/jacobi/rowabs/none),
    tolerance full2D_solverTol, restart full2D_gmresRestart."""
    setup = Full2DSetup(mat, Vxy, EfL, EfR, host_reservoirs=False)
    pmode, solver = _transport_options(setup, solver)
    # both checks come before the device operator is built, as in the reference
    if rho0 is not None:
        rho0 = validate_initial_rho(rho0, setup.n_total)
    _, basis_limit = gmres_basis_guard(setup.params, setup.n_total, max(solver.restart, solver.restart_max), pmode)
    with Full2DOperator(setup) as op:
        used = _configure(op, setup, pmode, basis_limit)
        solver.precondition = used != "none"
        rho, info = op.solve(solver, x0=rho0)
        return _result(op, setup, rho, info, rho0 is not None, used)


class TransportSession:
    """One device operator for the Gummel iterations of one bias point of
    the self-consistent loop: between iterations only the potential changes
    (set_potential and a rebuilt preconditioner), while the contact data
    (reservoirs, inflow rhs) stay valid because the contacts hold phi = 0 and
    the bias; each solve warm-starts from the previous device solution, as
    the loop's rho0 = rho of the previous iteration
    (RTD/solve2D_rtd_self_consistent.m:116-121).  Outputs per solve as
    solve_transport_DG_full2D, without copying rho to the host."""

    def __init__(self, mat: dict, Vxy, EfL, EfR, rho0=None, solver: SolveOptions | None = None):
        self.setup = Full2DSetup(mat, Vxy, EfL, EfR, host_reservoirs=False)
        self.pmode, self.solver = _transport_options(self.setup, solver)
        self.rho0 = None if rho0 is None else validate_initial_rho(rho0, self.setup.n_total)
        _, limit = gmres_basis_guard(self.setup.params, self.setup.n_total,
                                     max(self.solver.restart, self.solver.restart_max), self.pmode)
        self.op = Full2DOperator(self.setup)
        self.used = _configure(self.op, self.setup, self.pmode, limit)
        self.solves = 0

    def solve(self, Vxy=None) -> Full2DResult:
        if Vxy is not None:
            self.op.set_potential(Vxy)
            self.used = _set_preconditioner(self.op, self.pmode)
        opts = dataclasses.replace(self.solver, precondition=self.used != "none", warm_device=self.solves > 0)
        x0 = self.rho0 if self.solves == 0 else None
        _, info = self.op.solve(opts, x0=x0, return_x=False)
        warm = self.solves > 0 or x0 is not None
        self.solves += 1
        return _result(self.op, self.setup, None, info, warm, self.used)

    def solution(self) -> np.ndarray:
        """The device solution in physical units (host copy)."""
        x = np.empty(self.op.n, np.complex128)
        check(self.op._lib.wdg_solution(self.op.ctx, dptr(x)), self.op.ctx)
        return x

    def close(self):
        self.op.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
