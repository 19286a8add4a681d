"""2D electrostatics and the self-consistent full-2D RTD loop.

* `eval_poisson2d` assembles M ~ div(eps grad phi) on the material grid
  with harmonic-mean face permittivities, half control volumes at free X
  ends, Dirichlet contact rows and an optional periodic Y duplicate column
  (RTD/eval_Poisson2D_rtd.m:1-205).
* `project_dg_density` integrates the raw DG density exactly over the
  Poisson dual volumes, element by element (RTD/project_DG_density_RTD.m).
* `poisson_step` is one direct or approximate-Newton update,
  J = M - diag(q^2 n / kT) (RTD/solve_Poisson2D_rtd.m:85-130).
* `solve2d_rtd_self_consistent` is the warm-started transport/Poisson loop
  with linear mixing and the unmixed-defect, density-change and Poisson
  residual tests (RTD/solve2D_rtd_self_consistent.m:88-205).
* `transient_cn_self_consistent` interleaves Crank-Nicolson steps of the
  device operator with a Poisson update every few steps.

`DevicePoisson2D` runs the projection and the update on the device (banded
FP64 LU of the 5-point operator in one workgroup, wdg2_poisson_step) from the
density of the operators' device solutions, and every loop uses it: the
stationary loop keeps one device operator per bias point
(full2d.TransportSession), the time-dependent loops their operators, so
between transport solves only the potential crosses the boundary.
`project_dg_density` and `poisson_step` are the same steps in host FP64.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

EPS0 = 8.8541878128e-12     # eval_Poisson2D_rtd.m:9
QE = 1.602176634e-19        # solve_Poisson2D_rtd.m:23-24
KB = 1.380649e-23


def _periodic_y(mat) -> bool:
    p = mat.get("dg", {}).get("params", {})
    yb = str(p.get("full2D_Y_boundary", "")).lower() == "periodic"
    pois = mat.get("poisson", {})
    if isinstance(pois, dict) and "yBoundary" in pois:
        yb = str(pois["yBoundary"]).lower() == "periodic"
    return yb


def x_boundary(mat) -> str:
    """get_RTD_poisson_x_boundary.m: 'dirichlet' unless mat.poisson.xBoundary."""
    pois = mat.get("poisson", {})
    v = str(pois.get("xBoundary", "dirichlet")).lower() if isinstance(pois, dict) else "dirichlet"
    if v not in ("dirichlet", "neumann"):
        raise ValueError("RTD:PoissonXBoundary: use dirichlet or neumann")
    return v


def eval_poisson2d(mat):
    """Returns (M [N x N] csr, node types [Nx, Ny] (1 free, 2 Dirichlet, 3
    periodic duplicate), periodic flag).  M acts on phi in volts with x, y
    in nm (the factor 1e18 of eps0/dx^2)."""
    nx, ny = int(mat["Nx"]), int(mat["Ny"])
    dx, dy = float(mat["dx"]), float(mat["dy"])
    er = np.asarray(mat.get("eps", mat.get("Eps")), float)
    if er.shape != (nx, ny) or not np.all(np.isfinite(er)) or np.any(er <= 0):
        raise ValueError("eval_Poisson:MissingEps: mat.eps (relative permittivity) [Nx, Ny] is required")
    for name, n, h in (("x", nx, dx), ("y", ny, dy)):
        g = mat.get(name)
        if g is not None and (np.size(g) != n or np.any(np.abs(np.diff(np.ravel(g)) - h) > 1e-10 * h)):
            raise ValueError(f"RTD:PoissonUniformGrid: mat.{name} must agree with its uniform spacing in nm")
    periodic = _periodic_y(mat)
    if periodic and (ny < 3 or np.max(np.abs(er[:, -1] - er[:, 0])) > 1e-12 * er.max()):
        raise ValueError("RTD:PoissonPeriodicMaterial: periodic Y requires >=3 nodes and matching endpoint eps")
    xb = x_boundary(mat)
    if mat.get("boundary") is not None and np.size(mat["boundary"]):
        bnd = np.asarray(mat["boundary"], bool)
    else:
        bnd = np.zeros((nx, ny), bool)
        if xb == "dirichlet":
            bnd[[0, nx - 1], :] = True
    ids = np.arange(nx * ny).reshape(ny, nx).T            # column-major node numbers
    types = 1 + bnd.astype(int)
    if periodic:
        types[:, -1] = 3
    cx = EPS0 * 1e18 / dx ** 2
    cy = EPS0 * 1e18 / dy ** 2
    hm = lambda a, b: 2 * a * b / (a + b)
    rows, cols, vals = [], [], []
    for iy in range(ny):
        for ix in range(nx):
            k = ids[ix, iy]
            if periodic and iy == ny - 1:
                rows += [k, k]
                cols += [k, ids[ix, 0]]
                vals += [1.0, -1.0]
                continue
            if bnd[ix, iy]:
                rows.append(k); cols.append(k); vals.append(1.0)
                continue
            diag = 0.0
            e0 = er[ix, iy]
            if ix > 0:
                c = cx * hm(e0, er[ix - 1, iy]) * (2 if ix == nx - 1 else 1)
                rows.append(k); cols.append(ids[ix - 1, iy]); vals.append(c); diag -= c
            if ix < nx - 1:
                c = cx * hm(e0, er[ix + 1, iy]) * (2 if ix == 0 else 1)
                rows.append(k); cols.append(ids[ix + 1, iy]); vals.append(c); diag -= c
            if iy > 0 or periodic:
                jm = iy - 1 if iy > 0 else ny - 2
                c = cy * hm(e0, er[ix, jm]) * (2 if (not periodic and iy == ny - 1) else 1)
                rows.append(k); cols.append(ids[ix, jm]); vals.append(c); diag -= c
            if iy < ny - 1:
                jp = iy + 1
                if periodic and jp == ny - 1:
                    jp = 0
                c = cy * hm(e0, er[ix, jp]) * (2 if (not periodic and iy == 0) else 1)
                rows.append(k); cols.append(ids[ix, jp]); vals.append(c); diag -= c
            rows.append(k); cols.append(k); vals.append(diag)
    M = sp.csr_matrix((vals, (rows, cols)), shape=(nx * ny, nx * ny))
    return M, types, periodic


def doping_si(mat):
    """get_RTD_doping_SI.m: Nd/Na in m^-3 (legacy N_d/N_a in cm^-3, checked when both exist)."""
    nx, ny = int(mat["Nx"]), int(mat["Ny"])

    def field(si, legacy, required):
        has_si, has_leg = mat.get(si) is not None, mat.get(legacy) is not None
        if not has_si and not has_leg:
            if required:
                raise ValueError("RTD:MissingDoping: provide mat.Nd [m^-3] or mat.N_d [cm^-3]")
            return np.zeros((nx, ny))
        v = np.asarray(mat[si], float) if has_si else np.asarray(mat[legacy], float) * 1e6
        if has_si and has_leg:
            leg = np.asarray(mat[legacy], float) * 1e6
            if np.max(np.abs(v - leg)) > 1e-12 * max(1.0, float(np.max(v))):
                raise ValueError(f"RTD:InconsistentDopingUnits: mat.{si} and mat.{legacy} disagree")
        return v

    return field("Nd", "N_d", True), field("Na", "N_a", False)


def _integration_map(axis, nodes):
    """integrationMap (project_DG_density_RTD.m:44-69): T[row, dof] = integral
    of the Lagrange basis over the dual volume of material node `row`."""
    nodes = np.asarray(nodes, float)
    v = axis.vertices
    if nodes.size < 2 or np.any(np.diff(nodes) <= 0) or max(abs(nodes[0] - v[0]), abs(nodes[-1] - v[-1])) > 1e-18:
        raise ValueError("RTD:ChargeGridMismatch: DG and Poisson grids must cover the same interval")
    edges = np.concatenate(([nodes[0]], 0.5 * (nodes[:-1] + nodes[1:]), [nodes[-1]]))
    widths = np.diff(edges)
    nl = axis.nl
    nq = int(np.ceil(nl / 2))
    pts, wts = np.polynomial.legendre.leggauss(nq)
    T = np.zeros((nodes.size, axis.ndof))
    jac = axis.jac
    for e in range(axis.ne):
        left, right = v[e], v[e + 1]
        rows = np.nonzero((edges[:-1] < right) & (edges[1:] > left))[0]
        for row in rows:
            a, b = max(edges[row], left), min(edges[row + 1], right)
            s = ((a + b) / 2 + (b - a) / 2 * pts - (left + right) / 2) / jac[e]
            basis = np.ones((nq, nl))
            for j in range(nl):
                for k in range(nl):
                    if k != j:
                        basis[:, j] *= (s - axis.r[k]) / (axis.r[j] - axis.r[k])
            T[row, e * nl:(e + 1) * nl] = (b - a) / 2 * (wts @ basis)
    return T, widths


def project_dg_density(nDG, setup, periodic: bool):
    """Conservative DG -> Poisson dual-volume density (m^-3) on the material
    grid; with periodic Y the endpoint half volumes merge and the output
    repeats the first column as the last.  Returns (density, info)."""
    mat = setup.mat
    nDG = np.asarray(nDG, float)
    if nDG.shape != (setup.X.ndof, setup.Y.ndof):
        raise ValueError("nDG must have shape [X DOF, Y DOF]")
    if abs(setup.coord_scale - 1e-9) > 1e-20:
        raise ValueError("RTD:PoissonCoordinateUnits: the RTD Poisson material coordinates must be in nm")
    Tx, wx = _integration_map(setup.X, np.ravel(mat["x"]) * 1e-9)
    Ty, wy = _integration_map(setup.Y, np.ravel(mat["y"]) * 1e-9)
    if periodic:
        Ty[0] += Ty[-1]
        wy = wy.copy()
        wy[0] += wy[-1]
        Ty, wy = Ty[:-1], wy[:-1]
    volume = np.outer(wx, wy)
    integral = Tx @ nDG @ Ty.T
    dens = integral / volume
    src_x = np.concatenate([setup.X.M.sum(axis=0) * setup.X.jac[e] for e in range(setup.X.ne)])
    src_y = np.concatenate([setup.Y.M.sum(axis=0) * setup.Y.jac[e] for e in range(setup.Y.ne)])
    source = src_x @ nDG @ src_y
    target = integral.sum()
    info = {"sourceIntegral": source, "targetIntegral": target,
            "relativeConservationError": abs(target - source) / max(abs(source), np.finfo(float).tiny),
            "uniqueVolume": volume}
    if periodic:
        dens = np.concatenate([dens, dens[:, :1]], axis=1)
    return dens, info


def poisson_step(M, types, n, mat, phi0, bias: float, method: str = "direct", alpha: float = 1.0):
    """One direct or approximate-Newton Poisson update (volts); returns
    (phi, info) with the physical residual before the step."""
    nx, ny = int(mat["Nx"]), int(mat["Ny"])
    N = nx * ny
    Nd, Na = doping_si(mat)
    neumann = x_boundary(mat) == "neumann"
    periodic = bool(np.any(types == 3))
    nu = ny - int(periodic)
    bc = np.zeros((nx, ny))
    bc[-1, :] = bias
    if mat.get("Vbc") is not None and np.size(mat["Vbc"]):
        bc = np.asarray(mat["Vbc"], float)
    col = lambda a: np.asarray(a, float)[:, :nu].ravel(order="F")
    ut = types[:, :nu].ravel(order="F")
    free, fixed = np.nonzero(ut == 1)[0], np.nonzero(ut == 2)[0]
    if free.size + fixed.size != nx * nu or (not neumann and fixed.size == 0) or (neumann and fixed.size):
        raise ValueError("RTD:PoissonNodeMap: invalid or ungrounded Poisson node types")
    mp = np.arange(nx * nu).reshape(nu, nx).T
    if periodic:
        mp = np.concatenate([mp, mp[:, :1]], axis=1)
    expand = sp.csr_matrix((np.ones(N), (np.arange(N), mp.ravel(order="F"))), shape=(N, nx * nu))
    A = (M[:nx * nu] @ expand).tocsr()
    old = col(phi0)
    nf = col(n)
    net = col(Nd - Na)
    charge = QE * (nf - net)
    v = old.copy()
    v[fixed] = col(bc)[fixed]
    Av = A @ v
    F = Av - charge
    scale = max(np.max(np.abs(QE * nf[free])), np.max(np.abs(QE * net[free])), np.max(np.abs(Av[free])), QE)
    Aff = A[free][:, free]
    method = method.lower()
    if method == "direct":
        if neumann:
            raise ValueError("RTD:NeumannNeedsNewton: Neumann X requires newton-raphson")
        J = Aff
    elif method in ("newton-raphson", "newton-rhapson"):
        J = Aff - sp.diags(QE ** 2 * nf[free] / (KB * float(mat["Temp"])))
    else:
        raise ValueError(f"RTD:UnknownPoissonMethod: {method}")
    delta = -spla.spsolve(J.tocsc(), F[free])
    lin = np.max(np.abs(J @ delta + F[free])) / max(np.max(np.abs(F[free])), QE)
    upd = v.copy()
    upd[free] = v[free] + alpha * delta
    cand = v.copy()
    cand[free] = v[free] + delta
    phi = (expand @ upd).reshape(ny, nx).T
    after = A @ upd - charge
    info = {"converged": bool(np.all(np.isfinite(upd)) and np.isfinite(lin) and lin < 1e-8),
            "linearRelres": lin, "physicalRelresBefore": np.max(np.abs(F[free])) / scale,
            "physicalRelresAfter": np.max(np.abs(after[free])) / scale, "unmixedStep": np.max(np.abs(delta)),
            "candidateUnmixed": (expand @ cand).reshape(ny, nx).T, "method": method}
    return phi, info


class DevicePoisson2D:
    """The Poisson update of `poisson_step` and the projection of
    `project_dg_density` on the device (wdg2_poisson_setup /
    wdg2_poisson_step): the density comes from the device solutions of
    full-2D operators, so a self-consistent loop keeps its transport states
    in HBM and only the potential (nx*ny values) passes the boundary.  The
    Poisson data live on `op`, whose grid every operator of a step shares."""

    def __init__(self, op, mat=None):
        from . import _lib as L
        from .solver import check
        self._lib, self._check, self.op = L.load(), check, op
        setup = op.setup
        mat = setup.mat if mat is None else mat
        self.mat = mat
        M, types, periodic = eval_poisson2d(mat)
        nx, ny = int(mat["Nx"]), int(mat["Ny"])
        nu = ny - int(periodic)
        self.nx, self.ny, self.nu, self.periodic = nx, ny, nu, periodic
        neumann = x_boundary(mat) == "neumann"
        ut = types[:, :nu].ravel(order="F")
        free, fixed = np.nonzero(ut == 1)[0], np.nonzero(ut == 2)[0]
        if free.size + fixed.size != nx * nu or (not neumann and fixed.size == 0) or (neumann and fixed.size):
            raise ValueError("RTD:PoissonNodeMap: invalid or ungrounded Poisson node types")
        self.neumann = neumann
        mp = np.arange(nx * nu).reshape(nu, nx).T
        if periodic:
            mp = np.concatenate([mp, mp[:, :1]], axis=1)
        N = nx * ny
        expand = sp.csr_matrix((np.ones(N), (np.arange(N), mp.ravel(order="F"))), shape=(N, nx * nu))
        A = (M[:nx * nu] @ expand).tocoo()
        Nd, Na = doping_si(mat)
        if abs(setup.coord_scale - 1e-9) > 1e-20:
            raise ValueError("RTD:PoissonCoordinateUnits: the RTD Poisson material coordinates must be in nm")
        Tx, wx = _integration_map(setup.X, np.ravel(mat["x"]) * 1e-9)
        Ty, wy = _integration_map(setup.Y, np.ravel(mat["y"]) * 1e-9)
        if periodic:
            Ty[0] += Ty[-1]
            wy = wy.copy()
            wy[0] += wy[-1]
            Ty, wy = Ty[:-1], wy[:-1]
        wn = setup.observable_weights()[0]
        if not 1 <= len(wn) <= 8:
            raise ValueError("1..8 zero-separation density weights expected")
        self._keep = dict(row=np.ascontiguousarray(A.row, np.int32), col=np.ascontiguousarray(A.col, np.int32),
                          val=np.ascontiguousarray(A.data, np.float64),
                          types=np.ascontiguousarray(ut, np.int32),
                          net=np.ascontiguousarray((Nd - Na)[:, :nu].ravel(order="F"), np.float64),
                          px=np.ascontiguousarray(Tx, np.float64), py=np.ascontiguousarray(Ty, np.float64),
                          vol=np.ascontiguousarray(np.outer(wx, wy).ravel(order="F"), np.float64),
                          widx=np.ascontiguousarray([j for j, _ in wn], np.int32),
                          wval=np.ascontiguousarray([w for _, w in wn], np.float64))
        k = self._keep
        i32 = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
        d = L.wdg2_poisson_desc(nx, nu, int(k["row"].size), i32(k["row"]), i32(k["col"]), L.dptr(k["val"]),
                                i32(k["types"]), L.dptr(k["net"]), float(mat["Temp"]), setup.X.ndof, setup.Y.ndof,
                                L.dptr(k["px"]), L.dptr(k["py"]), L.dptr(k["vol"]), int(k["widx"].size),
                                i32(k["widx"]), L.dptr(k["wval"]))
        check(self._lib.wdg2_poisson_setup(op.ctx, C.byref(d)), op.ctx)

    def _full(self, u):
        """Unique-node values (x fastest) -> [Nx, Ny] with the periodic duplicate row."""
        full = np.empty((self.nx, self.ny))
        full[:, :self.nu] = np.asarray(u).reshape(self.nu, self.nx).T
        if self.periodic:
            full[:, -1] = full[:, 0]
        return full

    def step(self, ops, phi0, bias: float, method: str = "direct", alpha: float = 1.0):
        """One update from the summed density of `ops`' device solutions:
        (phi, info) with the keys of `poisson_step` plus 'density' (the
        projected density [m^-3] on the material grid)."""
        method = method.lower()
        if method == "direct":
            if self.neumann:
                raise ValueError("RTD:NeumannNeedsNewton: Neumann X requires newton-raphson")
            newton = 0
        elif method in ("newton-raphson", "newton-rhapson"):
            newton = 1
        else:
            raise ValueError(f"RTD:UnknownPoissonMethod: {method}")
        mat, nu = self.mat, self.nu
        bc = np.zeros((self.nx, self.ny))
        bc[-1, :] = bias
        if mat.get("Vbc") is not None and np.size(mat["Vbc"]):
            bc = np.asarray(mat["Vbc"], float)
        col = lambda a: np.ascontiguousarray(np.asarray(a, float)[:, :nu].ravel(order="F"))
        phi_u, bc_u = col(phi0), col(bc)
        m = self.nx * nu
        out, cand, dens = np.empty(m), np.empty(m), np.empty(m)
        info = self._lib_info()
        ctxs = (C.c_void_p * len(ops))(*[o.ctx for o in ops])
        from . import _lib as L
        self._check(self._lib.wdg2_poisson_step(ctxs, len(ops), L.dptr(phi_u), L.dptr(bc_u), newton, float(alpha),
                                                L.dptr(out), L.dptr(cand), L.dptr(dens), C.byref(info)), self.op.ctx)
        lin = info.linear_relres
        upd = out
        return self._full(upd), {
            "converged": bool(np.all(np.isfinite(upd)) and np.isfinite(lin) and lin < 1e-8),
            "linearRelres": lin, "physicalRelresBefore": info.physical_relres_before,
            "physicalRelresAfter": info.physical_relres_after, "unmixedStep": info.unmixed_step,
            "candidateUnmixed": self._full(cand), "method": method, "density": self._full(dens)}

    @staticmethod
    def _lib_info():
        from . import _lib as L
        return L.wdg2_poisson_info()


def solve2d_rtd_self_consistent(mat: dict, max_iterations: int = 100, mixing: float = 0.1,
                                potential_tolerance: float = 5e-4, density_tolerance: float = 5e-4,
                                poisson_residual_tolerance: float = 8e-2, poisson_method: str = "direct",
                                initial_phi=None, biases=None, solver=None):
    """The warm-started transport/Poisson loop of
    RTD/solve2D_rtd_self_consistent.m:88-205 (Dirichlet X contacts) over
    mat['Vdd'] (or `biases`).  Returns the sweep (per bias: current,
    density, potential, phi, flag, relres, converged, iterations, status)
    and the per-iteration histories."""
    from .full2d import TransportSession
    if x_boundary(mat) != "dirichlet":
        raise NotImplementedError("RTD:Unsupported: the Neumann-X contact calibration is not built")
    if not (0 < mixing <= 1):
        raise ValueError("mixing must lie in (0, 1]")
    bias = np.ravel(mat["Vdd"] if biases is None else biases).astype(float)
    Vc = np.asarray(mat["Vc"], float)
    x = np.ravel(mat["x"]).astype(float)
    ramp_len = x[-1] - float(mat["Lc1"]) - float(mat["Lc2"])
    if ramp_len <= 0:
        raise ValueError("RTD:ContactGeometry: contact regions must leave a positive central length")
    ramp = np.repeat(np.clip((x - float(mat["Lc1"])) / ramp_len, 0, 1)[:, None], int(mat["Ny"]), axis=1)
    eval_poisson2d(mat)   # the Poisson data's checks before any device work
    sweep = {k: [] for k in ("bias", "current", "density", "potential", "phi", "flag", "relres", "converged",
                             "iterations", "status")}
    histories = []
    rho_guess, prev_phi = None, None
    for ib, b in enumerate(bias):
        if ib == 0:
            phi = b * ramp if initial_phi is None else np.asarray(initial_phi, float)
        else:
            phi = prev_phi + (b - bias[ib - 1]) * ramp
        efl, efr = Vc[0, 0] + 0.05, Vc[-1, 0] + 0.05 - b
        if "EfL" in mat:
            efl = float(mat["EfL"])
        if "EfR" in mat:
            efr = float(mat["EfR"]) - b
        prev_n, hist, status, converged = None, [], "iteration-limit", False
        # one device operator per bias: the iterations change its potential,
        # warm-start from its device solution and update phi on the device
        with TransportSession(mat, Vc - phi, efl, efr, rho0=rho_guess, solver=solver) as ses:
            dev = DevicePoisson2D(ses.op)
            for it in range(1, max_iterations + 1):
                V = Vc - phi
                tr = ses.solve(V if it > 1 else None)
                s = tr.info["solve"]
                h = {"iteration": it, "transportFlag": s["flag"], "transportRelres": s["relres"],
                     "warmStartUsed": bool(s["warmStartUsed"]), "initialTrueRelres": s["initialTrueRelres"],
                     "gmresIterations": s["iterations"], "densityChange": np.inf}
                if s["flag"] != 0 or not np.isfinite(s["relres"]):
                    status = "transport-not-converged"
                    hist.append(h)
                    break
                if prev_n is not None:
                    h["densityChange"] = np.linalg.norm(tr.n - prev_n) / max(np.linalg.norm(tr.n),
                                                                             np.linalg.norm(prev_n),
                                                                             np.finfo(float).tiny)
                cand, pinfo = dev.step([ses.op], phi, b, poisson_method)
                h.update({"potentialDefect": float(np.max(np.abs(cand - phi))),
                          "poissonConverged": pinfo["converged"],
                          "poissonPhysicalRelres": pinfo["physicalRelresBefore"]})
                hist.append(h)
                if not pinfo["converged"]:
                    status = "poisson-not-converged"
                    break
                if (h["potentialDefect"] <= potential_tolerance and h["densityChange"] <= density_tolerance
                        and h["poissonPhysicalRelres"] <= poisson_residual_tolerance):
                    status, converged = "fixed-point-converged", True
                    break
                if it == max_iterations:
                    break
                prev_n = tr.n
                phi = phi + mixing * (cand - phi)
            rho_guess = ses.solution()
        for k, v in (("bias", b), ("current", tr.j), ("density", tr.n), ("potential", V), ("phi", phi),
                     ("flag", s["flag"]), ("relres", s["relres"]), ("converged", converged),
                     ("iterations", len(hist)), ("status", status)):
            sweep[k].append(v)
        histories.append(hist)
        if not converged:
            break
        prev_phi = phi
    return {"sweep": sweep, "histories": histories, "converged": all(sweep["converged"]) and len(hist) > 0}


def transient_cn_self_consistent(op, phi0, bias: float, dt: float, steps: int, poisson_every: int = 10,
                                 precision: str = "mixed", tol: float = 1e-8, restart: int = 30,
                                 poisson_method: str = "newton-raphson", rho0=None, use_graph: bool = True,
                                 on_block=None):
    """Crank-Nicolson steps of a full-2D operator (`Full2DOperator`) with a
    2D Poisson update of the potential every `poisson_every` steps: after
    each block of steps the device density is projected onto the Poisson
    dual volumes, one Poisson step from the current phi gives the new phi
    (DevicePoisson2D: density, projection and update on the device), and
    the operator's potential becomes Vc - phi (the contacts keep
    phi = 0 and bias, so the inflow data stay valid).  rho0=None continues
    from the operator's device state (e.g. a stationary solve).  on_block:
    called with each block's history entry as soon as the block is done.
    Returns (phi, history) with one entry per block."""
    mat = op.setup.mat
    Vc = np.asarray(mat["Vc"], float)
    dev = DevicePoisson2D(op)
    phi = np.asarray(phi0, float).copy()
    done, hist = 0, []
    while done < steps:
        k = min(int(poisson_every), steps - done)
        _, info = op.transient_cn(dt, k, tol=tol, restart=restart, precision=precision,
                                  rho0=rho0 if done == 0 else None, return_rho=False, use_graph=use_graph)
        done += k
        new, pinfo = dev.step([op], phi, bias, poisson_method)
        hist.append({"steps": done, "gmresIterations": info["iterations"], "maxStepRelres": info["max_relres"],
                     "flag": info["flag"], "potentialUpdate": float(np.max(np.abs(new - phi))),
                     "poissonPhysicalRelres": pinfo["physicalRelresBefore"], "poissonConverged": pinfo["converged"]})
        if on_block is not None:
            on_block(hist[-1])
        phi = new
        op.set_potential(Vc - phi)
    return phi, hist


def transient_cn_batch_self_consistent(ops, offsets, phi0, bias: float, dt: float, steps: int,
                                       poisson_every: int = 10, precision: str = "mixed", tol: float = 1e-8,
                                       restart: int = 30, poisson_method: str = "newton-raphson", rho0=None,
                                       use_graph: bool = True, on_block=None):
    """Crank-Nicolson steps of several full-2D instances of one device (the
    transverse modes: instance i sees the potential Vc - phi + offsets[i]),
    all stepped in one batched call (wdg_transient_cn_batch), with a 2D
    Poisson update every `poisson_every` steps driven by the summed density
    of the instances: project the summed device density onto the Poisson dual
    volumes, one Poisson step from the current phi, and every instance's
    potential becomes Vc - phi_new + offset (the contacts keep phi = 0 and
    bias, so the inflow data stay valid).  rho0: None (continue from the
    device states) or one initial state per instance.  on_block: called with
    each block's history entry as soon as the block is done.  Returns (phi,
    history) with one entry per block, holding the phi the block ran with."""
    from .solver import transient_cn_batch
    ops = list(ops)
    if len(offsets) != len(ops):
        raise ValueError("one potential offset per instance")
    mat = ops[0].setup.mat
    Vc = np.asarray(mat["Vc"], float)
    dev = DevicePoisson2D(ops[0])
    phi = np.asarray(phi0, float).copy()
    for op, off in zip(ops, offsets):
        op.set_potential(Vc - phi + float(off))
    done, hist = 0, []
    while done < steps:
        k = min(int(poisson_every), steps - done)
        _, infos = transient_cn_batch(ops, dt, k, tol=tol, restart=restart, precision=precision,
                                      rho0=rho0 if done == 0 else None, return_rho=False, use_graph=use_graph)
        done += k
        # the summed density, its projection and the update stay on the device
        new, pinfo = dev.step(ops, phi, bias, poisson_method)
        hist.append({"steps": done, "phi": phi, "gmresIterations": [i["iterations"] for i in infos],
                     "maxStepRelres": max(i["max_relres"] for i in infos), "flag": max(i["flag"] for i in infos),
                     "potentialUpdate": float(np.max(np.abs(new - phi))),
                     "poissonPhysicalRelres": pinfo["physicalRelresBefore"], "poissonConverged": pinfo["converged"]})
        if on_block is not None:
            on_block(hist[-1])
        phi = new
        for op, off in zip(ops, offsets):
            op.set_potential(Vc - phi + float(off))
    return phi, hist
