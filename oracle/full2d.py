"""FP64 CPU restatement of the reference's full-2D Wigner-DG operator (rho basis).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Unknown F(rho_x, rho_y, X, Y) vectorised with rho_x fastest, then rho_y, the
X-DG DOF and the Y-DG DOF (DG_code/initParams_full2D.m:4-8,105-107); an
operator part coefficient * kron(C_center, R_relative) acts as
Y = c R U C^T on U = reshape(u, nRelative, nCenter).

Restates the 'specular', 'zero-inflow' and 'periodic' Y boundaries:
  DG_code/initParams_full2D.m:1-243                 -> init_params_full2d
  DG_code/core/get_Diff_full2D.m:1-813              -> get_diff_full2d (parts, rhs parts)
  DG_code/core/get_Drift_full2D.m:1-880             -> DriftFull2D ('fv-consistent', 'collocated')
  DG_code/core/get_CAP_full2D.m:1-39                -> cap_full2d
  DG_code/core/get_SpecularReflection_full2D.m      -> reflection
  DG_code/core/get_Boundary_full2D.m, get_InflowBoundary_full2D.m,
  get_ContactMask_full2D.m, get_DefaultReservoirRho_full2D.m -> boundary_full2d
  DG_code/core/get_Observables_full2D.m:1-135       -> observables_full2d

Pinning (tests/test_oracle_full2d.py): the reference's own test assertions
(DG_code/tests/full2DDiffOperatorTest.m, full2DDriftSystemSolveTest.m,
test_full2D_boundaries.m) re-expressed on their small setups; A+ + A- = A
and A+ A- = 0 for the characteristic split; constant potential without CAP
gives a zero drift; zero separation has no potential difference; the
FV-consistent relative block is a 3x3 stencil; a flat-band constant state
is transported without a residual in the interior; the direct solve has a
vanishing residual; with periodic Y the operator commutes with the
double-cover half-period shift rho(Y+W/2, rho_y+W) (get_PeriodicYStateError_full2D.m)
and a Y-constant state has no Y-transport residual anywhere.  Parity
unpinned: the 'hard-wall' Y boundary, contact-mode reservoirs,
position-dependent masses (not restated).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
from scipy.interpolate import RegularGridInterpolator

from . import wigner1d as W

HBAR = 1.054571817e-34   # get_Diff_full2D.m:810-813
M0 = 9.1093837015e-31
QE = 1.602176634e-19     # get_Drift_full2D.m / get_Observables_full2D.m constants
KB_RES = 1.38064852e-23  # get_DefaultReservoirRho_full2D.m:220


def _param(params, name, default=None):
    """readParam of the reference: params.name, else params.full2D.name."""
    if isinstance(params, dict):
        v = params.get(name)
        if v is not None and not (hasattr(v, "__len__") and not isinstance(v, str) and len(v) == 0):
            return v
        sub = params.get("full2D")
        if isinstance(sub, dict) and sub.get(name) is not None:
            return sub[name]
    return default


def _params(mat):
    return mat.get("dg", {}).get("params", {}) if isinstance(mat, dict) else {}


def coordinate_vector(mat, name, dname, nname):
    if mat.get(name) is not None:
        return np.asarray(mat[name], float).ravel()
    return np.arange(mat[nname]) * mat[dname]


def mass_to_kg(m):
    m = np.asarray(m, float)
    out = np.where(np.abs(m) > 1e-25, m * M0, m)
    return np.where(out <= 0, 0.041 * M0, out)


# ---------------------------------------------------------------------------
# initParams_full2D
# ---------------------------------------------------------------------------

def make_rect_dg_axis(coord, n_local, params, label):
    """initParams_full2D.m:143-197."""
    if n_local < 2:
        raise ValueError("N_K must be at least 2")
    n_int = coord.size - 1
    n_el = _param(params, f"N_el_{label}")
    if n_el is None:
        if n_int % (n_local - 1):
            raise ValueError(f"{label} grid incompatible with {n_local} local DG nodes")
        n_el = n_int // (n_local - 1)
    n_el = int(n_el)
    r, w = W.jacobi_gl(0, 0, n_local - 1)
    V = W.vandermonde1d(n_local - 1, r)
    invV = np.linalg.inv(V)
    M = invV.T @ invV
    Dr = W.dmatrix1d(n_local - 1, r, V)
    S = M @ Dr
    vertices = np.linspace(coord[0], coord[-1], n_el + 1)
    jac = (vertices[1:] - vertices[:-1]) / 2
    nodes = np.concatenate([(vertices[e] + vertices[e + 1]) / 2 + jac[e] * r for e in range(n_el)])
    return dict(label=label, nLocal=n_local, nElements=n_el, nDof=n_el * n_local, r=r, w=w, V=V, M=M,
                Dr=Dr, S=S, vertices=vertices, nodes=nodes, jacobian=jac)


def make_fv_relative_axis(n, length):
    """initParams_full2D.m:199-214: cell centres, central difference D."""
    edges = np.linspace(-length / 2, length / 2, n + 1)
    cells = 0.5 * (edges[:-1] + edges[1:])
    delta = edges[1] - edges[0]
    D = sp.diags([-np.ones(n - 1) / (2 * delta), np.ones(n - 1) / (2 * delta)], [-1, 1], shape=(n, n))
    return dict(nCells=n, length=length, edges=edges, cells=cells, delta=delta, D=D.tocsr())


def init_params_full2d(mat, Vxy=None, EfL=None, EfR=None):
    """initParams_full2D.m:1-116 (non-periodic Y)."""
    params = _params(mat)
    p = {"EfL": EfL, "EfR": EfR}
    scale = _param(params, "full2D_coordinate_scale", 1e-9)
    p["coordinateScale"] = scale
    x = coordinate_vector(mat, "x", "dx", "Nx") * scale
    y = coordinate_vector(mat, "y", "dy", "Ny") * scale
    nk_def = _param(params, "N_K_chi", 3)
    X = make_rect_dg_axis(x, int(_param(params, "N_K_X", nk_def)), params, "X")
    Y = make_rect_dg_axis(y, int(_param(params, "N_K_Y", nk_def)), params, "Y")
    p["dg"] = {"X": X, "Y": Y, "nLocalDof": X["nLocal"] * Y["nLocal"],
               "centerSize": (X["nDof"], Y["nDof"]), "nCenterDof": X["nDof"] * Y["nDof"],
               "MXY": np.kron(Y["M"], X["M"])}
    nrx = int(_param(params, "N_rho_x", _param(params, "N_xi", 40)))
    nry = int(_param(params, "N_rho_y", nrx))
    l_def = _param(params, "Ly", max(abs(y[-1] - y[0]), np.finfo(float).eps))
    lrx = float(_param(params, "L_rho_x", l_def))
    lry = float(_param(params, "L_rho_y", l_def))
    periodic = y_boundary_type(mat) == "periodic"
    if periodic:                                    # initParams_full2D.m:65-76
        W_y = y[-1] - y[0]
        lry = float(_param(params, "L_rho_y", 2 * W_y))
        if abs(lry - 2 * W_y) > 1e-10 * W_y:
            raise ValueError("DG:Full2D:PeriodicRelativeLength")
        if Y["nElements"] % 2 or nry % 2 or nry < 4:
            raise ValueError("DG:Full2D:PeriodicGrid")
    rx = make_fv_relative_axis(nrx, lrx)
    ry = make_fv_relative_axis(nry, lry)
    ry["periodic"] = periodic
    if periodic:                                    # :82-87
        ry["D"] = (ry["D"] + sp.csr_matrix(([-1.0, 1.0], ([0, nry - 1], [nry - 1, 0])), shape=(nry, nry))
                   / (2 * ry["delta"])).tocsr()
    Ix, Iy = sp.identity(nrx, format="csr"), sp.identity(nry, format="csr")
    rel = {"rhoX": rx, "rhoY": ry, "NrhoX": nrx, "NrhoY": nry, "nDof": nrx * nry,
           "Ax1D": (-1j * rx["D"]).tocsr(), "Ay1D": (-1j * ry["D"]).tocsr()}
    rel["Ax"] = sp.kron(Iy, rel["Ax1D"]).tocsr()
    rel["Ay"] = sp.kron(rel["Ay1D"], Ix).tocsr()
    p["relative"] = rel
    p["arraySize"] = (nrx, nry, X["nDof"], Y["nDof"])
    p["nTotal"] = nrx * nry * X["nDof"] * Y["nDof"]
    p["domain"] = {"X": (x[0], x[-1]), "Y": (y[0], y[-1])}
    p["periodicY"] = {"enabled": periodic}
    if periodic:                                    # :109-115
        p["periodicY"].update({"width": y[-1] - y[0],
                               "centerShift": np.roll(np.arange(Y["nDof"]), Y["nDof"] // 2),
                               "relativeShift": np.roll(np.arange(nry), nry // 2)})
    return p


def close_periodic_axis(ops):
    """close_PeriodicDGAxis_full2D.m: the last/first DG faces become one
    interior interface with the interior central and dissipative fluxes."""
    n = ops["nDof"]
    left = sp.csr_matrix(([1.0], ([0], [0])), shape=(1, n))
    right = sp.csr_matrix(([1.0], ([0], [n - 1])), shape=(1, n))
    ll = sp.csr_matrix(ops["leftLift"][:, None])
    rl = sp.csr_matrix(ops["rightLift"][:, None])
    out = dict(ops)
    out["H1"] = (ops["H1"] - 0.5 * ll @ right + 0.5 * rl @ left).tocsr()
    seam = 0.5 * (ll - rl) @ (left - right)
    out["H2Interior"] = (ops["H2Interior"] + seam).tocsr()
    out["H2Boundary"] = sp.csr_matrix((n, n))
    out["H2"] = out["H2Interior"]
    return out


def periodic_y_shift(u, p):
    """The double-cover half-period shift rho(Y+W/2, rho_y+W) of
    get_PeriodicYStateError_full2D.m (relative and centre Y index shifts)."""
    nrx, nry, nxd, nyd = p["arraySize"]
    R4 = np.asarray(u).reshape(nyd, nxd, nry, nrx)
    R4 = R4[p["periodicY"]["centerShift"]][:, :, p["periodicY"]["relativeShift"], :]
    return R4.reshape(-1)


def y_boundary_type(mat):
    """get_YBoundaryType_full2D.m:14-45."""
    v = _param(_params(mat), "full2D_Y_boundary", "specular")
    v = str(v).strip().lower()
    v = {"specular-reflection": "specular", "reflection": "specular", "reflecting": "specular",
         "zero_inflow": "zero-inflow", "zeroinflow": "zero-inflow", "zero": "zero-inflow",
         "hardwall": "hard-wall", "hard_wall": "hard-wall", "dirichlet": "hard-wall"}.get(v, v)
    if v not in ("specular", "zero-inflow", "hard-wall", "periodic"):
        raise ValueError(f"unknown full2D_Y_boundary {v}")
    return v


# ---------------------------------------------------------------------------
# get_Diff_full2D
# ---------------------------------------------------------------------------

def one_dimensional_dg_operators(axis):
    """get_Diff_full2D.m:206-266 (H1 central, H2 |.|-flux, boundary split, lifts)."""
    n, ne, nd = axis["nLocal"], axis["nElements"], axis["nDof"]
    invM = np.linalg.inv(axis["M"])
    S = axis["S"]
    jac = axis["jacobian"]
    P1 = np.zeros((n, n)); P1[0, 0] = 1
    P2 = np.zeros((n, n)); P2[0, n - 1] = 1
    P3 = np.zeros((n, n)); P3[n - 1, n - 1] = 1
    P4 = np.zeros((n, n)); P4[n - 1, 0] = 1
    K1 = invM @ (S + 0.5 * P1 - 0.5 * P3)
    K2 = invM @ (0.5 * P1 + 0.5 * P3)
    K2L = invM @ (0.5 * P1)
    K2R = invM @ (0.5 * P3)
    K3 = -0.5 * invM @ P2
    K4 = K3
    K5 = 0.5 * invM @ P4
    K6 = -K5
    # vertices are equispaced (linspace), so every element has the same jacobian
    upper = sp.diags([1.0 / jac[1:]], [1], shape=(ne, ne))
    lower = sp.diags([1.0 / jac[1:]], [-1], shape=(ne, ne))
    diag = sp.diags([1.0 / jac], [0], shape=(ne, ne))
    H1 = sp.kron(upper, K5) + sp.kron(lower, K3) + sp.kron(diag, K1)
    H2 = sp.kron(upper, K6) + sp.kron(lower, K4) + sp.kron(diag, K2)
    left_el = sp.csr_matrix(([1.0], ([0], [0])), shape=(ne, ne))
    right_el = sp.csr_matrix(([1.0], ([ne - 1], [ne - 1])), shape=(ne, ne))
    H2B = sp.kron(left_el @ diag, K2L) + sp.kron(right_el @ diag, K2R)
    left_lift = np.zeros(nd); left_lift[:n] = invM[:, 0] / jac[0]
    right_lift = np.zeros(nd); right_lift[nd - n:] = invM[:, n - 1] / jac[-1]
    left_trace = np.zeros(nd); left_trace[0] = 1
    right_trace = np.zeros(nd); right_trace[nd - 1] = 1
    return dict(H1=H1.tocsr(), H2=H2.tocsr(), H2Boundary=H2B.tocsr(), H2Interior=(H2 - H2B).tocsr(),
                leftBoundary=sp.csr_matrix(np.outer(left_lift, left_trace)),
                rightBoundary=sp.csr_matrix(np.outer(right_lift, right_trace)),
                leftLift=left_lift, rightLift=right_lift, nDof=nd)


def split_characteristic(A, tol=1e-12):
    """get_Diff_full2D.m:718-730: A+ and A- from eig(A), entries below 1e-12 zeroed."""
    lam, Phi = np.linalg.eig(np.asarray(A.todense() if sp.issparse(A) else A))
    lp = np.where(lam.real <= tol, 0, lam)
    lm = np.where(lam.real >= -tol, 0, lam)
    Pinv = np.linalg.inv(Phi)
    Ap = Phi @ np.diag(lp) @ Pinv
    Am = Phi @ np.diag(lm) @ Pinv
    Ap[np.abs(Ap) < 1e-12] = 0
    Am[np.abs(Am) < 1e-12] = 0
    return Ap, Am, lam


def _norm_flux(value, allowed, kind):
    f = str(value).strip().lower()
    if f in ("lf", "lax-friedrichs", "rusanov-local"):
        f = "rusanov"
    elif f in ("matrix-upwind", "characteristic-upwind") and "upwind" in allowed:
        f = "upwind"
    if f not in allowed:
        raise ValueError(f"unknown {kind} flux {f}")
    return f


def flux_types(params):
    """get_Diff_full2D.m:57-62, :349-391."""
    general = _param(params, "full2D_fluxType", _param(params, "rho_flux", _param(params, "fluxType", "rusanov")))
    xdef = _norm_flux(general, ("central", "rusanov"), "general")
    xf = _norm_flux(_param(params, "full2D_X_flux", xdef), ("central", "rusanov", "upwind"), "X")
    yf = _norm_flux(_param(params, "full2D_Y_flux", xf), ("central", "rusanov", "upwind"), "Y")
    return xf, yf


def relative_transport_operators(p, params, xf, yf):
    """get_Diff_full2D.m:295-347."""
    tol = _param(params, "full2D_characteristicTol", 1e-12)
    rel = p["relative"]
    Ix, Iy = sp.identity(rel["NrhoX"]), sp.identity(rel["NrhoY"])
    IR = sp.identity(rel["nDof"])
    Apx, Amx, lx = split_characteristic(rel["Ax1D"], tol)
    Apy, Amy, ly = split_characteristic(rel["Ay1D"], tol)
    th_def = 0.01
    thx = _param(params, "full2D_thetaLF_X", _param(params, "full2D_thetaLF", _param(params, "thetaLF", th_def))) \
        if xf == "rusanov" else 0.0
    thy = _param(params, "full2D_thetaLF_Y", _param(params, "full2D_thetaLF", _param(params, "thetaLF", th_def))) \
        if yf == "rusanov" else 0.0
    ax = _param(params, "full2D_alphaLF_X", np.max(np.abs(lx)))
    ay = _param(params, "full2D_alphaLF_Y", np.max(np.abs(ly)))
    out = {}
    for key, Ap, Am, th, al, kr in (("X", Apx, Amx, thx, ax, lambda m: sp.kron(Iy, m)),
                                     ("Y", Apy, Amy, thy, ay, lambda m: sp.kron(m, Ix))):
        d = {"AplusChar": kr(sp.csr_matrix(Ap)).tocsr(), "AminusChar": kr(sp.csr_matrix(Am)).tocsr(),
             "alphaLF": float(al), "thetaLF": float(th), "betaLF": float(th) * float(al)}
        d["AabsChar"] = (d["AplusChar"] - d["AminusChar"]).tocsr()
        d["AabsLF"] = (d["betaLF"] * IR).tocsr()
        out[key] = d
    return out


def transport_scales(mat, params):
    """get_Diff_full2D.m:732-755."""
    me_x = mat.get("me_x_ch", 0.041)
    me_y = mat.get("me_y_ch", me_x)
    sx = _param(params, "full2D_Q_diff_x", _param(params, "full2D_transportScaleX", HBAR / mass_to_kg(me_x)))
    sy = _param(params, "full2D_Q_diff_y", _param(params, "full2D_transportScaleY", HBAR / mass_to_kg(me_y)))
    f = _param(params, "full2D_diff_scale", _param(params, "rho_diff_scale", 1))
    return float(f) * float(sx), float(f) * float(sy)


def reflection(p):
    """get_SpecularReflection_full2D.m: R_x = I_y (x) flip_x, R_y = flip_y (x) I_x."""
    nx, ny = p["relative"]["NrhoX"], p["relative"]["NrhoY"]
    fx = sp.csr_matrix(np.fliplr(np.eye(nx)))
    fy = sp.csr_matrix(np.fliplr(np.eye(ny)))
    return sp.kron(sp.identity(ny), fx).tocsr(), sp.kron(fy, sp.identity(nx)).tocsr()


def get_diff_full2d(mat, p, boundary):
    """get_Diff_full2D.m:36-204.  Returns (parts, rhs_parts, info); a part is
    (name, coefficient, center sparse, relative sparse)."""
    params = _params(mat)
    xf, yf = flux_types(params)
    ytype = y_boundary_type(mat)
    if ytype == "hard-wall":
        raise NotImplementedError(ytype)
    ox = one_dimensional_dg_operators(p["dg"]["X"])
    oy = one_dimensional_dg_operators(p["dg"]["Y"])
    if ytype == "periodic":
        oy = close_periodic_axis(oy)
    nxd, nyd = p["dg"]["X"]["nDof"], p["dg"]["Y"]["nDof"]
    IX, IY = sp.identity(nxd), sp.identity(nyd)
    C = {"X": {"H1": sp.kron(IY, ox["H1"]), "H2I": sp.kron(IY, ox["H2Interior"]),
               "H2B": sp.kron(IY, ox["H2Boundary"])},
         "Y": {"H1": sp.kron(oy["H1"], IX), "H2I": sp.kron(oy["H2Interior"], IX),
               "H2B": sp.kron(oy["H2Boundary"], IX), "bottom": sp.kron(oy["leftBoundary"], IX),
               "top": sp.kron(oy["rightBoundary"], IX)}}
    R = relative_transport_operators(p, params, xf, yf)
    sx, sy = transport_scales(mat, params)
    rel = p["relative"]
    parts, rhs_parts = [], []

    def add(name, c, Cm, Rm):
        parts.append((name, c, sp.csr_matrix(Cm), sp.csr_matrix(Rm)))

    add("X central rho transport", sx, C["X"]["H1"], rel["Ax"])
    if xf == "upwind":
        add("X interior characteristic upwind", sx, C["X"]["H2I"], R["X"]["AabsChar"])
    elif R["X"]["betaLF"] != 0:
        add("X interior Rusanov", sx, C["X"]["H2I"], R["X"]["AabsLF"])
    add("X physical boundary characteristic penalty", sx, C["X"]["H2B"], R["X"]["AabsChar"])
    Rx, Ry = reflection(p)
    # Source (X-left): closed pieces reflect rho_x, contact pieces inject A+ rho_source
    for side, lift, vec, charm, sign_c, sign_r in (
            ("source", ox["leftBoundary"], ox["leftLift"], R["X"]["AplusChar"], -1.0, 1.0),
            ("drain", ox["rightBoundary"], ox["rightLift"], R["X"]["AminusChar"], 1.0, -1.0)):
        b = boundary[side]
        mask = b["contactMask"]
        closed = ~mask
        if closed.any():
            add(f"X-{side} closed-face reflection correction", sign_c * sx,
                sp.kron(sp.diags(closed.astype(float)), lift), charm @ Rx)
        for iy in np.nonzero(mask)[0]:
            e = np.zeros(nyd); e[iy] = 1
            rhs_parts.append((sx, np.kron(e, vec), sign_r * (charm @ b["rhoBoundary"][:, iy])))
    add("Y central rho transport", sy, C["Y"]["H1"], rel["Ay"])
    if yf == "upwind":
        add("Y interior characteristic upwind", sy, C["Y"]["H2I"], R["Y"]["AabsChar"])
    elif R["Y"]["betaLF"] != 0:
        add("Y interior Rusanov", sy, C["Y"]["H2I"], R["Y"]["AabsLF"])
    if ytype != "periodic":
        add("Y physical boundary characteristic penalty", sy, C["Y"]["H2B"], R["Y"]["AabsChar"])
    if ytype == "specular":
        add("Y-bottom specular reflection correction", -sy, C["Y"]["bottom"], R["Y"]["AplusChar"] @ Ry)
        add("Y-top specular reflection correction", sy, C["Y"]["top"], R["Y"]["AminusChar"] @ Ry)
    info = {"fluxTypeX": xf, "fluxTypeY": yf, "betaLF": (R["X"]["betaLF"], R["Y"]["betaLF"]),
            "yBoundaryType": ytype, "scales": (sx, sy), "relativeOps": R}
    return parts, rhs_parts, info


def assemble_parts(parts, p):
    n = p["nTotal"]
    A = sp.csr_matrix((n, n), dtype=complex)
    for _, c, Cm, Rm in parts:
        A = A + c * sp.kron(Cm, Rm, format="csr")
    return A.tocsr()


def assemble_rhs(rhs_parts, p):
    r = np.zeros(p["nTotal"], complex)
    for c, cv, rv in rhs_parts:
        r += c * np.kron(cv, rv)
    return r


def kron_diagonal(parts, p):
    """assembleKronDiagonal (get_Diff_full2D.m:495-514): sum c kron(diag C, diag R)."""
    d = np.zeros(p["nTotal"], complex)
    for _, c, Cm, Rm in parts:
        d += c * np.kron(Cm.diagonal(), Rm.diagonal())
    return d


def kron_row_abs_sum(parts, p):
    """assembleKronRowAbsSum (get_Diff_full2D.m:516-535): sum |c| kron(rowabs C, rowabs R)."""
    s = np.zeros(p["nTotal"])
    for _, c, Cm, Rm in parts:
        s += abs(c) * np.kron(np.asarray(abs(Cm).sum(axis=1)).ravel(), np.asarray(abs(Rm).sum(axis=1)).ravel())
    return s


def apply_parts(parts, u, p):
    """applyKronOperator (get_Diff_full2D.m:618-639): Y = sum c R U C^T."""
    nr, nc = p["relative"]["nDof"], p["dg"]["nCenterDof"]
    U = np.asarray(u).reshape(nc, nr).T
    Y = np.zeros((nr, nc), complex)
    for _, c, Cm, Rm in parts:
        Y += c * (Cm @ (Rm @ U).T).T
    return Y.T.reshape(-1)


# ---------------------------------------------------------------------------
# Boundary data
# ---------------------------------------------------------------------------

def field_to_xy(field, mat, default=0.0):
    """fieldToXY of the reference (material X/Y field conventions)."""
    nx, ny = mat["Nx"], mat["Ny"]
    if field is None:
        return default * np.ones((nx, ny))
    f = np.squeeze(np.asarray(field, float))
    if f.ndim == 0:
        return float(f) * np.ones((nx, ny))
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
    raise ValueError("cannot interpret material field")


def contact_mask(mat, p, side, y_face):
    """get_ContactMask_full2D.m:14-90 (explicit mask, Nd edge, all)."""
    params = _params(mat)
    explicit = _param(params, f"full2D_{side}ContactMask")
    scale = p["coordinateScale"]
    grid_y = coordinate_vector(mat, "y", "dy", "Ny") * scale

    def nearest(vals):
        idx = np.abs(y_face[:, None] - grid_y[None, :]).argmin(axis=1)
        return np.asarray(vals, float)[idx] > 0.5

    if explicit is not None:
        m = np.asarray(explicit).ravel()
        if m.size == y_face.size:
            mask = m.astype(bool)
        elif m.size == mat["Ny"]:
            mask = nearest(m)
        else:
            raise ValueError("contact mask size")
    elif mat.get("Nd") is not None:
        thr = _param(params, "full2D_contactNdThreshold", 0)
        Nd = field_to_xy(mat["Nd"], mat, 0)
        row = Nd[0 if side == "source" else mat["Nx"] - 1, :]
        mask = nearest(row > thr)
    else:
        mask = np.ones(y_face.size, bool)
    if not mask.any():
        mask = np.ones(y_face.size, bool)
    return mask


def default_reservoir(mat, p, side, Ef, Vxy):
    """get_DefaultReservoirRho_full2D.m:1-99: separable cosine quadrature of a
    contact Fermi-Dirac distribution on a (k_x, k_y) grid."""
    params = _params(mat)
    rel = p["relative"]
    if Ef is None or not np.isfinite(Ef) or not _param(params, "full2D_useMaterialReservoir", True):
        return np.zeros((rel["nDof"], p["dg"]["Y"]["nDof"]))
    temp = mat.get("Temp", 300)
    deg = float(np.atleast_1d(mat.get("deg_factor", 1))[0])
    nkx = int(_param(params, "full2D_reservoir_Nkx", rel["NrhoX"]))
    nky = int(_param(params, "full2D_reservoir_Nky", rel["NrhoY"]))
    kx = -2 * np.pi / rel["rhoX"]["length"] * (np.arange(1, nkx + 1) - 0.5 * (nkx + 1))
    ky = -2 * np.pi / rel["rhoY"]["length"] * (np.arange(1, nky + 1) - 0.5 * (nky + 1))
    dkx, dky = abs(kx[1] - kx[0]), abs(ky[1] - ky[0])
    cosx = np.cos(np.outer(rel["rhoX"]["cells"], kx))
    cosy = np.cos(np.outer(rel["rhoY"]["cells"], ky))
    y_face = p["dg"]["Y"]["nodes"]
    ix = 0 if side == "source" else mat["Nx"] - 1
    grid_y = coordinate_vector(mat, "y", "dy", "Ny") * p.get("coordinateScale", 1)
    Vf = field_to_xy(Vxy if Vxy is not None else mat.get("V"), mat, 0)
    mxf = field_to_xy(mat.get("me_x"), mat, mat.get("me_x_ch", 0.041))
    myf = field_to_xy(mat.get("me_y"), mat, mat.get("me_y_ch", 0.041))

    def edge(row):
        row = np.asarray(row, float)
        if np.all(row == row[0]):
            return np.full(y_face.size, row[0])
        # interp1 'linear','extrap'
        return np.interp(y_face, grid_y, row, left=None, right=None) if (
            y_face.min() >= grid_y.min() and y_face.max() <= grid_y.max()) else _interp_extrap(grid_y, row, y_face)

    Ve, mxe, mye = edge(Vf[ix]), mass_to_kg(edge(mxf[ix])), mass_to_kg(edge(myf[ix]))
    out = np.zeros((rel["nDof"], y_face.size))
    KX, KY = np.meshgrid(kx, ky, indexing="ij")
    for iy in range(y_face.size):
        E = Ve[iy] + HBAR ** 2 * KX ** 2 / (2 * mxe[iy] * QE) + HBAR ** 2 * KY ** 2 / (2 * mye[iy] * QE)
        arg = np.clip(QE * (E - Ef) / (KB_RES * temp), -80, 80)
        f = 2 * deg / (1 + np.exp(arg))
        grid = cosx @ f @ cosy.T * dkx * dky / (2 * np.pi) ** 2
        out[:, iy] = grid.reshape(-1, order="F")
    return out


def _interp_extrap(x, y, q):
    x = np.asarray(x, float); y = np.asarray(y, float)
    i = np.clip(np.searchsorted(x, q) - 1, 0, x.size - 2)
    t = (q - x[i]) / (x[i + 1] - x[i])
    return y[i] + t * (y[i + 1] - y[i])


def boundary_full2d(mat, p, EfL, EfR, Vxy=None):
    """get_Boundary_full2D.m + get_InflowBoundary_full2D.m (direct rho data,
    material-default or zero reservoirs; contact masks)."""
    params = _params(mat)
    y_face = p["dg"]["Y"]["nodes"]
    nrel, nyd = p["relative"]["nDof"], p["dg"]["Y"]["nDof"]
    out = {}
    for side, Ef in (("source", EfL), ("drain", EfR)):
        direct = _param(params, f"full2D_{side}Rho")
        if direct is not None:
            d = np.asarray(direct, complex)
            rho = np.repeat(d.reshape(-1, 1), nyd, axis=1) if d.ndim == 1 else d
            if rho.shape != (nrel, nyd):
                raise ValueError("boundary data shape")
        else:
            model = str(_param(params, "full2D_reservoirModel", "material-default")).lower().replace("_", "-")
            if model in ("material", "material-default", "bulk", "bulk-local", "local-bulk"):
                rho = default_reservoir(mat, p, side, Ef, Vxy)
            elif model in ("zero", "none", "off", "false"):
                rho = np.zeros((nrel, nyd))
            else:
                raise NotImplementedError(f"reservoir model {model}")
        mask = contact_mask(mat, p, side, y_face)
        rho = np.array(rho, complex)
        rho[:, ~mask] = 0
        out[side] = {"rhoBoundary": rho, "contactMask": mask}
    return out


# ---------------------------------------------------------------------------
# get_Drift_full2D and CAP
# ---------------------------------------------------------------------------

def cap_full2d(p):
    """get_CAP_full2D.m: C_rho = I_y (x) C_x + C_y (x) I_x with the 1D CAP profile."""
    rx, ry = p["relative"]["rhoX"], p["relative"]["rhoY"]
    wx = W.get_cap(rx["cells"], rx["length"], rx["delta"])
    wy = W.get_cap(ry["cells"], ry["length"], ry["delta"])
    if ry.get("periodic"):
        wy = np.zeros_like(wy)       # no absorbing edge on a periodic coherence coordinate
    return (wx[None, :] + wy[:, None]).reshape(-1)   # rho_x fastest


class DriftFull2D:
    """get_Drift_full2D.m: potential + CAP, matrix-free apply."""

    def __init__(self, mat, p, Vxy=None):
        params = _params(mat)
        if Vxy is None:
            Vxy = mat.get("V", np.zeros((mat["Nx"], mat["Ny"])))
        self.mode = str(_param(params, "full2D_potentialDiscretization", "fv-consistent")).lower()
        if self.mode not in ("fv-consistent", "collocated"):
            raise ValueError(self.mode)
        self.p = p
        scale = p["coordinateScale"]
        gx = coordinate_vector(mat, "x", "dx", "Nx") * scale
        gy = coordinate_vector(mat, "y", "dy", "Ny") * scale
        Vf = field_to_xy(Vxy, mat, 0)
        ix, iy = np.argsort(gx), np.argsort(gy)
        self.gx, self.gy = gx[ix], gy[iy]
        self.Vf = Vf[np.ix_(ix, iy)]
        self.periodic = bool(p.get("periodicY", {}).get("enabled"))
        if self.periodic and np.max(np.abs(self.Vf[:, 0] - self.Vf[:, -1])) > 1e-10 * max(1, np.abs(self.Vf).max()):
            raise ValueError("DG:Full2D:NonperiodicPotential")
        self.interp = RegularGridInterpolator((self.gx, self.gy), self.Vf, method="linear")
        q_def = QE / HBAR
        self.Qdrift = float(_param(params, "full2D_Q_drift", q_def))
        ds = float(_param(params, "full2D_drift_scale", 1))
        cs = float(_param(params, "full2D_cap_scale", 1))
        self.pot_coef = 1j * ds * self.Qdrift
        self.cap_coef = ds * cs * self.Qdrift
        self.cap = cap_full2d(p)
        rel = p["relative"]
        self.nrel, self.ncen = rel["nDof"], p["dg"]["nCenterDof"]
        X, Y = p["dg"]["X"], p["dg"]["Y"]
        self.centerX = np.tile(X["nodes"], Y["nDof"])
        self.centerY = np.repeat(Y["nodes"], X["nDof"])
        self.relX = np.tile(rel["rhoX"]["cells"], rel["NrhoY"])
        self.relY = np.repeat(rel["rhoY"]["cells"], rel["NrhoX"])
        if self.mode == "fv-consistent":
            self._fv_setup(params)

    def V(self, qx, qy):
        """evaluateExtendedPotential: clamp X and Y independently, then bilinear."""
        qx = np.clip(qx, self.gx[0], self.gx[-1])
        if self.periodic:
            qy = self.gy[0] + np.mod(qy - self.gy[0], self.gy[-1] - self.gy[0])
        else:
            qy = np.clip(qy, self.gy[0], self.gy[-1])
        return self.interp(np.stack([np.ravel(qx), np.ravel(qy)], axis=-1)).reshape(np.shape(qx))

    def _fv_setup(self, params):
        p = self.p
        rel = p["relative"]
        nx, ny = rel["NrhoX"], rel["NrhoY"]
        Ex = sp.diags([np.ones(nx), np.ones(nx)], [0, -1], shape=(nx + 1, nx))
        Ey = sp.diags([np.ones(ny), np.ones(ny)], [0, -1], shape=(ny + 1, ny))
        edges_y = rel["rhoY"]["edges"]
        if self.periodic:                           # unique periodic vertices, get_Drift_full2D.m:180-184
            Ey = sp.identity(ny) + sp.csr_matrix((np.ones(ny), (np.arange(ny), np.r_[ny - 1, np.arange(ny - 1)])),
                                                 shape=(ny, ny))
            edges_y = edges_y[:-1]
        self.P = sp.kron(Ey, Ex).tocsr()
        vx, vy = np.meshgrid(rel["rhoX"]["edges"], edges_y, indexing="ij")
        self.vx, self.vy = vx.reshape(-1, order="F"), vy.reshape(-1, order="F")
        X, Y = p["dg"]["X"], p["dg"]["Y"]
        nlx, nly = X["nLocal"], Y["nLocal"]
        order = _param(params, "full2D_potentialQuadratureOrder")
        nqx = order if order is not None else _param(params, "full2D_potentialQuadratureX", max(4, 2 * nlx - 1))
        nqy = order if order is not None else _param(params, "full2D_potentialQuadratureY", max(4, 2 * nly - 1))
        nqx, nqy = max(nlx, int(round(nqx))), max(nly, int(round(nqy)))
        qx, wx = W.jacobi_gl(0, 0, nqx - 1)
        qy, wy = W.jacobi_gl(0, 0, nqy - 1)
        ipx = W.vandermonde1d(nlx - 1, qx) @ np.linalg.inv(X["V"])
        ipy = W.vandermonde1d(nly - 1, qy) @ np.linalg.inv(Y["V"])
        self.interpolation = np.kron(ipy, ipx)
        weights = np.kron(wy, wx)
        invM = np.linalg.inv(p["dg"]["MXY"])
        self.centerProjection = (self.interpolation * weights[:, None]) @ invM.T
        self.elements = []
        for ey in range(Y["nElements"]):
            yq = 0.5 * (Y["vertices"][ey] + Y["vertices"][ey + 1]) + Y["jacobian"][ey] * qy
            yids = ey * nly + np.arange(nly)
            for ex in range(X["nElements"]):
                xq = 0.5 * (X["vertices"][ex] + X["vertices"][ex + 1]) + X["jacobian"][ex] * qx
                xids = ex * nlx + np.arange(nlx)
                lx, ly = np.meshgrid(xids, yids, indexing="ij")
                ids = (lx + ly * X["nDof"]).reshape(-1, order="F")
                Xq, Yq = np.meshgrid(xq, yq, indexing="ij")
                self.elements.append((ids, Xq.reshape(-1, order="F"), Yq.reshape(-1, order="F")))

    def delta_v(self, Xq, Yq):
        """DeltaV at the relative vertices (rows) x quadrature points (cols)."""
        rx, ry = 0.5 * self.vx[:, None], 0.5 * self.vy[:, None]
        return self.V(Xq[None, :] + rx, Yq[None, :] + ry) - self.V(Xq[None, :] - rx, Yq[None, :] - ry)

    def apply(self, u):
        U = np.asarray(u).reshape(self.ncen, self.nrel).T
        if self.mode == "collocated":
            return (self.diagonal() * np.asarray(u).reshape(-1))
        Y = self.cap_coef * (self.cap[:, None] * U)
        for ids, Xq, Yq in self.elements:
            Uq = U[:, ids] @ self.interpolation.T
            dV = self.delta_v(Xq, Yq)
            res = self.P.T @ (dV * (self.P @ Uq))
            res = self.pot_coef / 16.0 * res
            Y[:, ids] += res @ self.centerProjection
        return Y.T.reshape(-1)

    def diagonal(self):
        if self.mode != "collocated":
            return self._fv_diagonal()
        Xc = np.repeat(self.centerX, self.nrel)
        Yc = np.repeat(self.centerY, self.nrel)
        rx = np.tile(self.relX, self.ncen)
        ry = np.tile(self.relY, self.ncen)
        dv = self.V(Xc + 0.5 * rx, Yc + 0.5 * ry) - self.V(Xc - 0.5 * rx, Yc - 0.5 * ry)
        return self.pot_coef * dv + self.cap_coef * np.tile(self.cap, self.ncen)

    def _fv_weights(self):
        """centerDiagonalWeights / centerRowAbsWeights (get_Drift_full2D.m:240-252)."""
        Ip = self.interpolation
        Pi = self.centerProjection
        return Pi * Ip, np.abs(Pi) * np.abs(Ip).sum(axis=1)[:, None]

    def _fv_diagonal(self):
        """fvConsistentDiagonal (get_Drift_full2D.m:466-488)."""
        dw, _ = self._fv_weights()
        D = self.cap_coef * np.repeat(self.cap[:, None], self.ncen, axis=1).astype(complex)
        for ids, Xq, Yq in self.elements:
            rel = (self.P.T @ self.delta_v(Xq, Yq)) / 16.0         # (nrel, nq)
            D[:, ids] += self.pot_coef * (rel @ dw)
        return D.T.reshape(-1)

    def row_abs(self):
        """getDriftRowAbsSum (get_Drift_full2D.m:490-523): |diag| (collocated) or
        the triangle-inequality bound with vertex degrees (FV-consistent)."""
        if self.mode == "collocated":
            return np.abs(self.diagonal())
        _, rw = self._fv_weights()
        deg = np.asarray(self.P.sum(axis=1)).ravel()
        S = abs(self.cap_coef) * np.repeat(np.abs(self.cap)[:, None], self.ncen, axis=1)
        for ids, Xq, Yq in self.elements:
            rel = (self.P.T @ (np.abs(self.delta_v(Xq, Yq)) * deg[:, None])) / 16.0
            S[:, ids] += abs(self.pot_coef) * (rel @ rw)
        return S.T.reshape(-1)

    def assemble(self):
        n = self.nrel * self.ncen
        cols = [self.apply(np.eye(1, n, k).ravel().astype(complex)) for k in range(n)]
        return sp.csr_matrix(np.array(cols).T)


# ---------------------------------------------------------------------------
# Observables (get_Observables_full2D.m:1-135)
# ---------------------------------------------------------------------------

def _zero_weights(cells):
    cells = np.asarray(cells, float)
    near = np.argmin(np.abs(cells))
    tol = max(1e-14, 1e-12 * max(1, np.abs(cells).max()))
    if abs(cells[near]) <= tol:
        return [near], [1.0]
    left = np.nonzero(cells < 0)[0]
    right = np.nonzero(cells > 0)[0]
    if left.size and right.size:
        l, r = left[-1], right[0]
        h = cells[r] - cells[l]
        return [l, r], [cells[r] / h, -cells[l] / h]
    return [near], [1.0]


def _derivative_row(cells, D):
    cells = np.asarray(cells, float)
    near = np.argmin(np.abs(cells))
    tol = max(1e-14, 1e-12 * max(1, np.abs(cells).max()))
    if abs(cells[near]) <= tol:
        return np.asarray(D.todense())[near]
    row = np.zeros(cells.size)
    left = np.nonzero(cells < 0)[0]
    right = np.nonzero(cells > 0)[0]
    l, r = left[-1], right[0]
    h = cells[r] - cells[l]
    row[l], row[r] = -1 / h, 1 / h
    return row


def observables_full2d(rho, p, mat):
    """Returns (nDG, jxDG, jyDG) on the DG centre grid (X dof x Y dof)."""
    rel = p["relative"]
    nx, ny = rel["NrhoX"], rel["NrhoY"]
    ncx, ncy = p["dg"]["X"]["nDof"], p["dg"]["Y"]["nDof"]
    R4 = np.asarray(rho).reshape((ncy, ncx, ny, nx))       # C order of (Y, X, rho_y, rho_x)
    ixs, wxs = _zero_weights(rel["rhoX"]["cells"])
    iys, wys = _zero_weights(rel["rhoY"]["cells"])
    n = np.zeros((ncx, ncy))
    for i, wi in zip(ixs, wxs):
        for j, wj in zip(iys, wys):
            n += wi * wj * R4[:, :, j, i].real.T
    drx = _derivative_row(rel["rhoX"]["cells"], rel["rhoX"]["D"])
    dry = _derivative_row(rel["rhoY"]["cells"], rel["rhoY"]["D"])
    dx = np.zeros((ncy, ncx), complex)
    for j, wj in zip(iys, wys):
        dx += wj * (R4[:, :, j, :] @ drx)
    dy = np.zeros((ncy, ncx), complex)
    for i, wi in zip(ixs, wxs):
        dy += wi * (R4[:, :, :, i] @ dry)
    mx = mass_to_kg(mat.get("me_x_ch", 0.041))
    my = mass_to_kg(mat.get("me_y_ch", mat.get("me_x_ch", 0.041)))
    jx = QE * HBAR / mx * dx.imag.T
    jy = QE * HBAR / my * dy.imag.T
    return n, jx, jy


# ---------------------------------------------------------------------------
# Normal flux split and boundary flux (get_NormalFlux_full2D.m,
# apply_InflowFlux_full2D.m)
# ---------------------------------------------------------------------------

def normal_flux(rel, normal, tol=1e-12):
    """F_n = A_n^+ f_inside + A_n^- f_boundary for an axis-aligned outward
    normal; rel holds NrhoX, NrhoY, Ax1D, Ay1D."""
    nX, nY = normal
    if abs(nY) <= tol:
        Ap, Am, lam = split_characteristic(nX * rel["Ax1D"], tol)
        lift = lambda m: sp.kron(sp.identity(rel["NrhoY"]), sp.csr_matrix(m))
        lam = np.tile(lam, rel["NrhoY"])
    elif abs(nX) <= tol:
        Ap, Am, lam = split_characteristic(nY * rel["Ay1D"], tol)
        lift = lambda m: sp.kron(sp.csr_matrix(m), sp.identity(rel["NrhoX"]))
        lam = np.repeat(lam, rel["NrhoX"])
    else:
        raise NotImplementedError("only axis-aligned normals")
    return {"Aplus": lift(Ap).tocsr(), "Aminus": lift(Am).tocsr(), "lambda": lam,
            "inflowMask": lam.real < -tol, "outflowMask": lam.real > tol}


def apply_inflow_flux(side, rho_inside, ghost=None):
    """apply_InflowFlux_full2D.m: closed (masked) columns use the ghost
    operator (rho_x -> -rho_x reflection) of the inside state."""
    rb = np.array(side["rhoBoundary"], complex)
    ri = np.asarray(rho_inside, complex)
    if ri.ndim == 1:
        ri = np.repeat(ri[:, None], rb.shape[1], axis=1)
    mask = np.asarray(side.get("contactMask", np.ones(rb.shape[1], bool)), bool)
    if (~mask).any():
        rb[:, ~mask] = (ghost @ ri[:, ~mask]) if ghost is not None else ri[:, ~mask]
    nf = side["normalFlux"]
    return nf["Aplus"] @ ri + nf["Aminus"] @ rb


def inflow_sides(mat, p, EfL, EfR, Vxy=None):
    """get_InflowBoundary_full2D.m:1-57 (sides with normal flux and masks)."""
    b = boundary_full2d(mat, p, EfL, EfR, Vxy)
    rel = p["relative"]
    b["source"]["normalFlux"] = normal_flux(rel, (-1.0, 0.0))
    b["drain"]["normalFlux"] = normal_flux(rel, (1.0, 0.0))
    for s in ("source", "drain"):
        b[s]["inflowComponents"] = int(np.count_nonzero(b[s]["normalFlux"]["inflowMask"]))
    return b


def solve_direct_full2d(mat, Vxy, EfL, EfR):
    """solve_transport_DG_full2D.m direct branch: A = diff + drift, rho = A\\rhs."""
    import scipy.sparse.linalg as spla
    p = init_params_full2d(mat, Vxy, EfL, EfR)
    bnd = boundary_full2d(mat, p, EfL, EfR, Vxy)
    parts, rparts, info = get_diff_full2d(mat, p, bnd)
    A = assemble_parts(parts, p) + DriftFull2D(mat, p, Vxy).assemble()
    rhs = assemble_rhs(rparts, p)
    rho = spla.spsolve(A.tocsc(), rhs)
    return p, A, rhs, rho


# ---------------------------------------------------------------------------
# Line-transverse preconditioner (get_LineTransversePreconditioner_full2D.m)
# ---------------------------------------------------------------------------

def _transverse_reference(mat, p, Vxy, ix_ref):
    """transverseReference (:128-172): a 2-cell rho_x, 2-element X strip with
    the potential of material column ix_ref, no X transport, no CAP, zero
    reservoirs; returns the (rho_y, Y) block K of the transport and V of the
    potential at rho_x cell 1 and X DOF 1 (V doubled for 'fv-consistent')."""
    import copy
    ref = copy.deepcopy(mat)
    Vf = field_to_xy(Vxy, mat, 0)
    ref["V"] = np.repeat(Vf[ix_ref][None, :], mat["Nx"], axis=0)
    q = dict(_params(mat))
    if isinstance(q.get("full2D"), dict):
        q["full2D"] = dict(q["full2D"])
    nyd = p["dg"]["Y"]["nDof"]
    q.update({"N_rho_x": 2, "N_K_X": 2, "N_el_X": 2, "full2D_Q_diff_x": 0, "full2D_cap_scale": 0,
              "full2D_reservoirModel": "zero", "full2D_sourceRho": np.zeros((2 * p["relative"]["NrhoY"], nyd)),
              "full2D_drainRho": np.zeros((2 * p["relative"]["NrhoY"], nyd))})
    ref.setdefault("dg", {})["params"] = q
    pr = init_params_full2d(ref, ref["V"], None, None)
    b = boundary_full2d(ref, pr, None, None, ref["V"])
    parts, _, _ = get_diff_full2d(ref, pr, b)
    Kfull = assemble_parts(parts, pr)
    drift = DriftFull2D(ref, pr, ref["V"])
    Vfull = drift.assemble()
    nrx, nry = pr["relative"]["NrhoX"], pr["relative"]["NrhoY"]
    nxd = pr["dg"]["X"]["nDof"]
    ids = np.array([0 + nrx * (ry + nry * (0 + nxd * iyd)) for iyd in range(nyd) for ry in range(nry)])
    K = Kfull[ids][:, ids].toarray()
    V = Vfull[ids][:, ids].toarray()
    if drift.mode == "fv-consistent":
        V = 2 * V
    return K, V


class LineTransverse:
    """get_LineTransversePreconditioner_full2D.m:1-239 with dense LU pages:
    M = (sigma I + Xref)(sigma I + Tref)/sigma in the eigenbasis U of -i D_rhoX;
    apply(r) = M^-1 r (applyInverse, :199-239)."""

    def __init__(self, mat, p, Vxy, info, drift):
        params = _params(mat)
        nrx, nry = p["relative"]["NrhoX"], p["relative"]["NrhoY"]
        X, Y = p["dg"]["X"], p["dg"]["Y"]
        self.sz = (nrx, nry, X["nDof"], Y["nDof"])
        self.nT = nry * Y["nDof"]
        self.nl, self.ne = X["nLocal"], X["nElements"]
        max_bytes = _param(params, "full2D_lineMaxFactorBytes", 4 * 1024 ** 3)
        self.factor_bytes = (32 * self.nT ** 2 * nrx + 32 * self.nl ** 2 * nrx * self.ne + 16 * nrx ** 2
                             + 8 * self.nT * nrx)
        if self.factor_bytes > max_bytes:
            raise MemoryError("DG:Full2D:LineMemoryLimit")
        ix_ref = int(_param(params, "full2D_lineReferenceXIndex", 1)) - 1
        lam, U = np.linalg.eig(np.asarray(p["relative"]["Ax1D"].todense()))
        lam = lam.real
        lam[np.abs(lam) < 100 * np.spacing(np.max(np.abs(lam)))] = 0
        self.U, self.lam = U, lam
        K, V = _transverse_reference(mat, p, Vxy, ix_ref)
        self.K, self.V = K, V
        wy = W.get_cap(p["relative"]["rhoY"]["cells"], p["relative"]["rhoY"]["length"], p["relative"]["rhoY"]["delta"])
        if p["relative"]["rhoY"].get("periodic"):
            wy = np.zeros_like(wy)          # sys.drift.CAP.profileY is zero for a periodic rho_y
        capY = np.tile(wy, Y["nDof"]) * drift.cap_coef
        wx = W.get_cap(p["relative"]["rhoX"]["cells"], p["relative"]["rhoX"]["length"], p["relative"]["rhoX"]["delta"])
        capX = np.real(np.sum(np.conj(U) * (wx[:, None] * U), axis=0)) * drift.cap_coef
        if drift.mode == "fv-consistent":
            Pm = sp.diags([np.ones(nrx), np.ones(nrx)], [0, -1], shape=(nrx + 1, nrx)).toarray()
            weight = np.real(np.sum(np.conj(U) * ((Pm.T @ Pm / 4) @ U), axis=0))
        else:
            weight = np.ones(nrx)
        self.capY, self.capX, self.weight = capY, capX, weight
        Qx = info["scales"][0]
        self.Qx = Qx
        xrate = abs(Qx) * np.max(np.abs(lam)) / np.min(X["jacobian"])
        trate = (np.abs(K).sum(axis=1).max() + np.abs(V).sum(axis=1).max() + np.max(np.abs(capY))
                 + np.max(np.abs(capX)))
        rel_shift = _param(params, "full2D_lineRelativeShift", 0.003)
        self.sigma = float(_param(params, "full2D_lineShift", rel_shift * max(xrate, trate, np.finfo(float).tiny)))
        self._x_factors(X, Qx * lam)
        self.B = []
        for k in range(nrx):
            Bk = K + weight[k] * V + np.diag(capY + capX[k])
            self.B.append(Bk / self.sigma + np.eye(self.nT))

    def _x_factors(self, X, speeds):
        """xFactors (:174-197)."""
        n, ne = self.nl, self.ne
        M = X["M"]
        e0 = np.zeros((n, n)); e0[0, 0] = 1
        en = np.zeros((n, n)); en[-1, -1] = 1
        from_left = np.zeros((n, n)); from_left[0, -1] = 1
        from_right = np.zeros((n, n)); from_right[-1, 0] = 1
        K1 = np.linalg.solve(M, X["S"] + 0.5 * e0 - 0.5 * en)
        K2 = np.linalg.solve(M, 0.5 * e0 + 0.5 * en)
        self.forward = speeds >= 0
        self.xinv = np.zeros((ne, speeds.size, n, n))
        self.xcoup = np.zeros((ne, speeds.size, n, n))
        for e in range(ne):
            for k, s in enumerate(speeds):
                a = s / (self.sigma * X["jacobian"][e])
                self.xinv[e, k] = np.linalg.inv(np.eye(n) + a * K1 + abs(a) * K2)
                self.xcoup[e, k] = (-a * np.linalg.solve(M, from_left) if self.forward[k]
                                    else a * np.linalg.solve(M, from_right))
        self.speeds = speeds

    def apply(self, r):
        """applyInverse (:199-239): modes of rho_x, the per-mode upwind X
        sweep, the per-mode transverse (rho_y, Y) solve, back to rho_x, / sigma."""
        nrx, nry, nxd, nyd = self.sz
        n, ne, nT = self.nl, self.ne, self.nT
        R4 = np.asarray(r).reshape(nyd, nxd, nry, nrx)                      # [Y, X, rho_y, rho_x]
        w = np.einsum("jk,yxrj->kxyr", self.U.conj(), R4).reshape(nrx, nxd, nT)   # t = rho_y + nry*Y
        v = np.zeros_like(w)
        for k in range(nrx):
            prev = np.zeros((n, nT), complex)
            order = range(ne) if self.forward[k] else range(ne - 1, -1, -1)
            for e in order:
                blk = w[k, e * n:(e + 1) * n] - self.xcoup[e, k] @ prev
                prev = self.xinv[e, k] @ blk
                v[k, e * n:(e + 1) * n] = prev
        out = np.stack([np.linalg.solve(self.B[k], v[k].T).T for k in range(nrx)])   # [k, X, t]
        out = out.reshape(nrx, nxd, nyd, nry)
        z = np.einsum("jk,kxyr->yxrj", self.U, out) / self.sigma
        return z.reshape(-1)
