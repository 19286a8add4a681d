"""Contact-mode reservoirs of the full-2D inflow boundary (host, FP64).

The Source/Drain reservoir density matrix is assembled from the transverse
contact modes phi_n(y) of the contact edge and a parabolic longitudinal
occupation per mode (get_ContactModeReservoirRho_full2D.m:1-137):

    rho_c(Y, rho_y, rho_x) = sum_n phi_n(Y + rho_y/2) conj(phi_n(Y - rho_y/2)) F_n(rho_x),
    F_n(rho_x) = sum_kx f_n(kx) cos(rho_x kx) dkx / (2 pi),

with f_n the Fermi occupation integrated over the unresolved k_z continuum
when requested, and the Fermi level solved from contact charge neutrality
with the same modes and mode-averaged masses.  Two mode models:

* periodic Y (get_PeriodicContactModes_full2D.m:1-75): Fourier-Galerkin
  modes of the contact potential on the periodic cell, exact plane waves
  for a Y-constant potential;
* otherwise the box problem: the finite-difference transverse Hamiltonian
  of the material Y grid with the boundary rows of
  eval_transversal_Hamiltonian.m (transverseHamiltonian, :285-313), its
  lowest modes normalised in the continuum, evaluated with zero ghost nodes
  one grid step outside each Y edge (transverseModeDensity, :458-485).

The Y boundary type 'hard-wall' is not built (the full-2D setup rejects it).
"""
from __future__ import annotations

import numpy as np

HBAR = 1.054571817e-34      # physicalConstants, get_ContactModeReservoirRho_full2D.m:808-813
QE = 1.602176634e-19
KB = 1.38064852e-23
M0 = 9.1093837015e-31


def _param(params, name, default=None):
    v = params.get(name) if isinstance(params, dict) else None
    if v is None or (hasattr(v, "size") and np.size(v) == 0):
        return default
    return v


def _logical(params, name, default):
    v = _param(params, name, default)
    if isinstance(v, str):
        return v.strip().lower() in ("true", "yes", "on", "1")
    return bool(np.asarray(v).ravel()[0]) if np.size(v) else bool(default)


def _scalar(v, default=np.nan):
    try:
        a = np.asarray(v, float).ravel()
    except (TypeError, ValueError):
        return default
    return float(a[0]) if a.size == 1 and np.isfinite(a[0]) else default


def trapezoid_weights(y):
    """quadratureWeights (:538-551): trapezoid weights of a nonuniform grid."""
    y = np.asarray(y, float).ravel()
    if y.size < 2:
        return np.ones(1)
    d = np.diff(y)
    w = 0.5 * (np.concatenate(([0.0], d)) + np.concatenate((d, [0.0])))
    w[0], w[-1] = 0.5 * (y[1] - y[0]), 0.5 * (y[-1] - y[-2])
    return w


def sanitize_relative_mass(m, default):
    """sanitizeRelativeMass (:754-757): non-finite or non-positive -> default."""
    m = np.real(np.asarray(m, float).ravel())
    return np.where(np.isfinite(m) & (m > 0), m, default)


def mode_averaged_mass(modes, mass_rel, default_rel, y):
    """modeAveragedMass (:521-536): the |phi|^2-weighted harmonic mean of the
    relative mass per mode, in kg."""
    mass_rel = sanitize_relative_mass(mass_rel, default_rel)
    w = trapezoid_weights(y)
    weight = np.abs(modes) ** 2
    wsum = (weight * w[:, None]).sum(axis=0)
    inv = (weight.T @ (w / mass_rel)) / np.maximum(wsum, np.finfo(float).eps)
    rel = 1.0 / np.maximum(inv, np.finfo(float).eps)
    return sanitize_relative_mass(rel, default_rel) * M0


def k_grid(nk, length):
    """makeKGrid (:719-723): midpoint grid k = -2 pi/L ((1:nk) - (nk+1)/2)."""
    nk = max(2, int(round(nk)))
    k = -2 * np.pi / length * (np.arange(1, nk + 1) - 0.5 * (nk + 1))
    return k, abs(k[1] - k[0])


def stable_softplus(x):
    """stableSoftplus (:637-645): log(1 + e^x) with the large/small branches."""
    x = np.asarray(x, float)
    out = np.empty_like(x)
    big, small = x > 40, x < -40
    mid = ~(big | small)
    out[big] = x[big]
    out[small] = np.exp(x[small])
    out[mid] = np.log1p(np.exp(x[mid]))
    return out


def sheet_density(Ef, energy, mx, mz, temp, deg):
    """contactSheetDensity (:627-636): sum_n deg sqrt(mx mz) kT/(pi hbar^2) softplus(q(Ef-E_n)/kT)."""
    kbt = KB * temp
    eta = QE * (Ef - np.real(energy)) / kbt
    dos = np.sqrt(np.real(mx) * np.real(mz))
    by_mode = deg * dos * kbt / (np.pi * HBAR ** 2) * stable_softplus(eta)
    return float(by_mode.sum()), by_mode


def solve_neutrality(energy, mx, mz, target, temp, deg, params, guess):
    """solveContactNeutrality (:561-626): bracket, then bisection to the
    relative density tolerance."""
    rtol = float(_param(params, "full2D_contactModeFermiTolerance", 1e-10))
    maxit = int(_param(params, "full2D_contactModeFermiMaxIterations", 200))
    kt_ev = KB * temp / QE
    emin = float(np.min(np.real(energy)))
    step = max(0.5, 20 * kt_ev)
    lo, hi = emin - max(2.0, 80 * kt_ev), emin + max(2.0, 80 * kt_ev)
    if np.isfinite(guess):
        lo, hi = min(lo, guess - step), max(hi, guess + step)
    dlo = sheet_density(lo, energy, mx, mz, temp, deg)[0]
    dhi = sheet_density(hi, energy, mx, mz, temp, deg)[0]
    for _ in range(50):
        if dlo <= target <= dhi:
            break
        step *= 2
        if dlo > target:
            lo -= step
            dlo = sheet_density(lo, energy, mx, mz, temp, deg)[0]
        if dhi < target:
            hi += step
            dhi = sheet_density(hi, energy, mx, mz, temp, deg)[0]
    if dlo > target or dhi < target:
        raise RuntimeError(f"DG:Full2D:ContactFermiBracketFailed: cannot bracket the Fermi level for {target:g} m^-2")
    converged, it = False, 0
    for it in range(1, maxit + 1):
        mid = 0.5 * (lo + hi)
        dmid = sheet_density(mid, energy, mx, mz, temp, deg)[0]
        if abs(dmid - target) / target <= rtol:
            lo = hi = mid
            converged = True
            break
        if dmid < target:
            lo = mid
        else:
            hi = mid
    Ef = 0.5 * (lo + hi)
    dens, by_mode = sheet_density(Ef, energy, mx, mz, temp, deg)
    res = abs(dens - target) / target
    return Ef, {"achieved": dens, "by_mode": by_mode, "relative_residual": res, "iterations": it,
                "converged": converged or res <= rtol, "bracket": (lo, hi)}


def fermi_model(raw):
    """normalizeFermiModel (:503-514)."""
    m = str(raw).strip().lower().replace("_", "-").replace(" ", "-")
    if m in ("contact", "contact-modes", "neutrality", "contact-neutrality", "solve"):
        return "contact-neutrality"
    if m in ("external", "legacy", "supplied", "input"):
        return "external"
    raise ValueError(f"DG:Full2D:UnknownContactFermiModel: unknown full2D_contactModeFermiModel '{raw}'")


def target_sheet_density(mat, params, side, coord_scale):
    """contactTargetSheetDensity (:515-560): explicit sheet density, else
    doping [cm^-3] times the contact width."""
    v = _scalar(_param(params, f"full2D_{side}ContactSheetDensity"))
    if np.isfinite(v):
        if v <= 0:
            raise ValueError(f"full2D_{side}ContactSheetDensity must be positive")
        return v, f"full2D_{side}ContactSheetDensity"
    dop_name = {"source": "N_s", "drain": "N_d"}.get(side)
    dop = _scalar(mat.get(dop_name)) if dop_name else np.nan
    wc = _scalar(mat.get("W_c"))
    if np.isfinite(dop) and np.isfinite(wc):
        width = _scalar(_param(params, f"full2D_{side}ContactWidth", wc))
        return dop * 1e6 * width * coord_scale, f"{dop_name}*full2D_{side}ContactWidth"
    return np.nan, "unavailable"


def kz_info(params, mz, energy, Ef, temp, integrate):
    """makeKzInfo (:663-718): k_z grid up to max(user window, Ef - min E_n + thermal tail)."""
    if not integrate:
        return {"integrate": False, "k": np.zeros(1)}
    nkz = int(_param(params, "full2D_contactMode_Nkz", _param(params, "full2D_reservoir_Nkz", 1001)))
    wmin = float(_param(params, "full2D_contactModeKzEnergyWindow", 0.3))
    tail_kt = float(_param(params, "full2D_contactModeKzThermalTail", 10))
    if wmin < 0 or tail_kt < 0:
        raise ValueError("k_z energy window and thermal tail must be non-negative")
    tail = tail_kt * KB * temp / QE
    occupied = max(float(np.max(np.real(Ef))) - float(np.min(np.real(energy))), 0.0)
    adaptive = occupied + tail
    window = max(wmin, adaptive)
    nkz = max(2, int(round(nkz)))
    kmax = np.sqrt(max(window, np.finfo(float).eps) * 2 * float(np.max(np.real(mz))) * QE) / HBAR
    return {"integrate": True, "k": np.linspace(0.0, kmax, nkz), "nKz": nkz, "energyWindowEV": window,
            "minimumEnergyWindowEV": wmin, "adaptiveEnergyWindowEV": adaptive, "thermalTailKBT": tail_kt,
            "occupiedEnergyWindowEV": occupied}


def longitudinal_occupation(kx, e_mode, mx, mz, Ef, temp, deg, kz):
    """longitudinalOccupation (:406-424): 2 deg f(E_n + hbar^2 kx^2/2mx [+ hbar^2 kz^2/2mz]),
    k_z integrated by the trapezoid rule over [0, kz_max] times 2/pi."""
    ekx = HBAR ** 2 * np.asarray(kx, float) ** 2 / (2 * mx * QE)
    if kz["integrate"]:
        k = kz["k"]
        ekz = HBAR ** 2 * k ** 2 / (2 * mz * QE)
        arg = np.clip(QE * (e_mode + ekx[:, None] + ekz[None, :] - Ef) / (KB * temp), -80, 80)
        return 2 * deg * np.trapezoid(1.0 / (1.0 + np.exp(arg)), k, axis=1) / np.pi
    arg = np.clip(QE * (e_mode + ekx - Ef) / (KB * temp), -80, 80)
    return 2 * deg / (1.0 + np.exp(arg))


def _constant_mass(mat, axis):
    """constantMass (get_PeriodicContactModes_full2D.m:77-95): the reference
    mass me_<axis>_ch (me_x_ch fallback), relative; a material field of that
    mass must be constant."""
    ref = mat.get(f"me_{axis}_ch", mat.get("me_x_ch"))
    m = float(np.asarray(ref, float).ravel()[0])
    if m < 1e-25:
        m = m / M0
    field = mat.get(f"me_{axis}")
    if field is not None and np.size(field):
        v = np.asarray(field, float).ravel()
        v = np.where(np.abs(v) < 1e-25, v / M0, v)
        if not np.all(np.isfinite(v)) or np.max(np.abs(v - m)) > 1e-10 * m:
            raise ValueError(f"DG:Full2D:PeriodicContactMass: periodic contact modes require constant me_{axis}")
    if not (np.isfinite(m) and m > 0):
        raise ValueError("DG:Full2D:PeriodicContactMass: periodic contact masses must be positive scalars")
    return m


def periodic_modes(setup, side, Vxy):
    """get_PeriodicContactModes_full2D.m:1-75: 2K+1 Fourier basis functions
    on the periodic cell; exact +/- plane waves when the contact potential is
    Y-constant, else the eigenvectors of H in that basis."""
    mat, params = setup.mat, setup.params
    gy = np.asarray(mat["y"], float).ravel() * setup.coord_scale
    width = gy[-1] - gy[0]
    nu = gy.size - 1
    dy = width / nu
    if nu < 2 or np.max(np.abs(np.diff(gy) - dy)) > 1e-10 * dy:
        raise ValueError("DG:Full2D:PeriodicContactGrid: periodic contact modes require a uniform material Y grid "
                         "including both cell endpoints")
    kmax = min((nu - 1) // 2, (setup.RY.n - 2) // 4)
    order = _param(params, "full2D_periodicModeOrder", kmax)
    order_f = float(np.asarray(order, float).ravel()[0])
    if not np.isfinite(order_f) or order_f < 0 or order_f != int(order_f) or order_f > kmax:
        raise ValueError(f"DG:Full2D:PeriodicModeOrder: periodicModeOrder must be an integer in 0:{kmax}")
    order = int(order_f)
    ix = 0 if side == "source" else int(mat["Nx"]) - 1
    V = setup._field_xy(Vxy if Vxy is not None else mat.get("V"), 0.0)
    edge = np.asarray(V[ix], float)
    if abs(edge[0] - edge[-1]) > 1e-10 * max(1.0, float(np.max(np.abs(edge)))):
        raise ValueError("DG:Full2D:NonperiodicPotential: contact potential endpoints must agree")
    mx, my, mz = (_constant_mass(mat, a) for a in ("x", "y", "z"))
    k = (2 * np.pi / width) * np.arange(-order, order + 1)
    basis = np.exp(1j * np.outer(gy[:-1] - gy[0], k)) / np.sqrt(width)
    H = np.diag(HBAR ** 2 * k ** 2 / (2 * my * M0 * QE)) + dy * (basis.conj().T @ (edge[:-1, None] * basis))
    H = 0.5 * (H + H.conj().T)
    if np.max(np.abs(edge - edge[0])) < 1e-13 * max(1.0, abs(edge[0])):
        energy = edge[0] + HBAR ** 2 * k ** 2 / (2 * my * M0 * QE)
        ids = np.argsort(energy, kind="stable")
        energy = energy[ids]
        coef = np.eye(k.size)[:, ids]
    else:
        energy, coef = np.linalg.eigh(H)
        ids = np.argsort(energy, kind="stable")
        energy, coef = energy[ids], coef[:, ids]
    modes = np.exp(1j * np.outer(gy - gy[0], k)) @ coef / np.sqrt(width)
    contact = {"gridY": gy, "width": width, "k": k, "coef": coef, "H": H, "Vedge": edge,
               "mXRelative": np.full(gy.size, mx), "mZRelative": np.full(gy.size, mz), "periodic": True}
    norm = {"normalizedIntegral": np.trapezoid(np.abs(modes) ** 2, gy, axis=0)}
    return contact, energy, modes, norm


def box_hamiltonian(gy, Vedge, my_rel):
    """transverseHamiltonian (:285-313): FD kinetic energy with harmonic-mean
    hopping; the end rows keep the first/last interior diagonal and omit the
    exterior neighbour (zero one grid step outside)."""
    ny = gy.size
    if ny < 3:
        raise ValueError("DG:Full2D:ContactModeGridTooSmall: the contact-mode reservoir needs at least three Y points")
    dy = float(np.mean(np.diff(gy)))
    t0 = HBAR ** 2 / (2 * M0 * QE * dy ** 2)
    i = np.arange(1, ny - 1)
    left = -2 * t0 / (my_rel[i - 1] + my_rel[i])
    right = -2 * t0 / (my_rel[i + 1] + my_rel[i])
    center = -left - right + Vedge[i]
    H = np.zeros((ny, ny))
    H[i, i - 1] = left
    H[i, i] = center
    H[i, i + 1] = right
    H[0, 0], H[0, 1] = center[0], right[0]
    H[-1, -1], H[-1, -2] = center[-1], left[-1]
    return H


def box_modes(setup, side, Vxy):
    """contactProblem + solveContactModes (:151-281, :328-374): the nModes
    eigenpairs of smallest |E| (eigs 'sm'; all when nModes equals the grid
    size) sorted ascending, normalised to int |phi|^2 dy = 1, sign fixed by
    the last grid value."""
    mat, params = setup.mat, setup.params
    ix = 0 if side == "source" else int(mat["Nx"]) - 1
    gy = np.asarray(setup.grid_y, float)
    order = np.argsort(gy, kind="stable")
    gy = gy[order]
    V = setup._field_xy(Vxy if Vxy is not None else mat.get("V"), 0.0)
    me_x = float(np.asarray(mat.get("me_x_ch", 0.041), float).ravel()[0])
    me_y = float(np.asarray(mat.get("me_y_ch", 0.041), float).ravel()[0])
    me_z = float(np.asarray(mat.get("me_z_ch", me_x), float).ravel()[0])
    field = lambda name, d: setup._field_xy(mat.get(name), d)[ix][order]
    Vedge = np.asarray(V[ix], float)[order]
    my = sanitize_relative_mass(field("me_y", me_y), me_y)
    mx = sanitize_relative_mass(field("me_x", me_x), me_x)
    mz = sanitize_relative_mass(field("me_z", me_z), me_z)
    H = box_hamiltonian(gy, Vedge, my)
    ng = H.shape[0]
    default = min(4, ng - 1)
    if mat.get("n_of_modes") is not None and np.size(mat.get("n_of_modes")):
        default = float(np.asarray(mat["n_of_modes"]).ravel()[0])
    nmodes = int(round(float(np.asarray(_param(params, "full2D_contactModeCount", default)).ravel()[0])))
    nmodes = max(1, min(nmodes, ng))
    E, vec = np.linalg.eig(H)
    E = np.real(E)
    if nmodes < ng:   # eigs(H, nModes, 'sm')
        keep = np.argsort(np.abs(E), kind="stable")[:nmodes]
        E, vec = E[keep], vec[:, keep]
    ids = np.argsort(E, kind="stable")[:nmodes]
    energy = E[ids]
    modes = np.real(vec[:, ids])
    w = trapezoid_weights(gy)
    raw = (w[:, None] * np.abs(modes) ** 2).sum(axis=0)
    modes = modes / np.where(raw > 0, np.sqrt(raw), 1.0)
    modes = np.where(modes[-1] < 0, -modes, modes)
    contact = {"gridY": gy, "H": H, "Vedge": Vedge, "mXRelative": mx, "mYRelative": my, "mZRelative": mz,
               "periodic": False, "defaults": (me_x, me_z), "HNnz": int(np.count_nonzero(H)),
               "modeZeroBoundaryY": (gy[0] - np.mean(np.diff(gy)), gy[-1] + np.mean(np.diff(gy)))}
    norm = {"rawIntegral": raw, "normalizedIntegral": (w[:, None] * np.abs(modes) ** 2).sum(axis=0)}
    return contact, energy, modes, norm


def _interp_zero(xg, fg, x):
    """interp1(xg, fg, x, 'linear', 0)."""
    out = np.interp(x, xg, fg)
    return np.where((x < xg[0]) | (x > xg[-1]), 0.0, out)


def contact_mode_reservoir(setup, side: str, Ef, Vxy, data=None):
    """get_ContactModeReservoirRho_full2D.m:35-137 for one contact: returns
    (rho [N_rho_x*N_rho_y, N_Y_dof], rho_x fastest, and an info dict).  The
    host-solved part (modes, masses, Fermi level, k grids) comes from
    contact_mode_data; the device builds the same reservoir from it
    (wdg2_set_contact_rhs)."""
    d = contact_mode_data(setup, side, Ef, Vxy) if data is None else data
    info = d["info"]
    if not np.isfinite(d["Ef"]):
        return np.zeros((setup.nrel, setup.Y.ndof), complex), info
    energy, mass_x, mass_z, modes, contact = d["energy"], d["mass_x"], d["mass_z"], d["modes"], d["contact"]
    kx, dkx, kz, Ef_used = d["kx"], d["dkx"], d["kz"], d["Ef"]
    temp, deg, gy = d["temp"], d["deg"], contact["gridY"]
    cosx = np.cos(np.outer(setup.RX.cells, kx))
    yface = np.asarray(setup.Y.nodes, float)
    rhoy_cells = np.asarray(setup.RY.cells, float)
    out = np.zeros((setup.nrel, yface.size), complex)
    profiles = np.zeros((setup.RX.n, energy.size))
    for m in range(energy.size):
        f = longitudinal_occupation(kx, energy[m], mass_x[m], mass_z[m], Ef_used, temp, deg, kz)
        rx = cosx @ f * dkx / (2 * np.pi)
        profiles[:, m] = rx
        yp = yface[None, :] + 0.5 * rhoy_cells[:, None]
        ym = yface[None, :] - 0.5 * rhoy_cells[:, None]
        if contact["periodic"]:
            c = contact["coef"][:, m] / np.sqrt(contact["width"])
            php = np.exp(1j * (yp[..., None] - gy[0]) * contact["k"]) @ c
            phm = np.exp(1j * (ym[..., None] - gy[0]) * contact["k"]) @ c
        else:
            dy = float(np.mean(np.diff(gy)))
            grid = np.concatenate(([gy[0] - dy], gy, [gy[-1] + dy]))
            vals = np.concatenate(([0.0], modes[:, m], [0.0]))
            php = _interp_zero(grid, vals, yp)
            phm = _interp_zero(grid, vals, ym)
        rhoy = php * np.conj(phm)                            # (rho_y, Y node)
        for iy in range(yface.size):
            out[:, iy] += np.outer(rhoy[:, iy], rx).ravel()  # rho_x fastest
    info["rhoXModeProfiles"] = profiles
    return out, info


def contact_mode_data(setup, side: str, Ef, Vxy):
    """The host-solved part of get_ContactModeReservoirRho_full2D.m:35-137
    for one contact: transverse modes (box or periodic Fourier), mode-averaged
    masses, the Fermi level (contact neutrality when requested) and the k_x /
    k_z grids.  Ef = NaN in the result means a zero reservoir."""
    mat, params = setup.mat, setup.params
    temp = float(np.asarray(mat.get("Temp", 300), float).ravel()[0])
    deg = float(np.asarray(mat.get("deg_factor", 1), float).ravel()[0]) if mat.get("deg_factor") is not None else 1.0
    if setup.periodic:
        contact, energy, modes, norm = periodic_modes(setup, side, Vxy)
        mx_def = float(contact["mXRelative"][0])
        mz_def = float(contact["mZRelative"][0])
    else:
        contact, energy, modes, norm = box_modes(setup, side, Vxy)
        mx_def, mz_def = contact["defaults"]
    gy = contact["gridY"]
    mass_x = mode_averaged_mass(modes, contact["mXRelative"], mx_def, gy)
    mass_z = mode_averaged_mass(modes, contact["mZRelative"], mz_def, gy)
    nkx = _param(params, "full2D_contactMode_Nkx", _param(params, "full2D_reservoir_Nkx", setup.RX.n))
    kx, dkx = k_grid(float(np.asarray(nkx).ravel()[0]), setup.RX.length)
    integrate = _logical(params, "full2D_contactModeIntegrateKz", ("me_z" in mat) or ("me_z_ch" in mat))
    model = fermi_model(_param(params, "full2D_contactModeFermiModel", "contact-neutrality"))
    target, origin = target_sheet_density(mat, params, side, setup.coord_scale)
    Ef_in = _scalar(Ef)
    fermi = {"requestedModel": model, "model": model, "inputEV": Ef_in, "targetSheetDensityM2": target,
             "targetOrigin": origin, "relativeResidual": np.nan}
    Ef_used = Ef_in
    if model == "contact-neutrality" and integrate and np.isfinite(target) and target > 0:
        Ef_used, sol = solve_neutrality(energy, mass_x, mass_z, target, temp, deg, params, Ef_in)
        fermi.update({"achievedSheetDensityM2": sol["achieved"], "relativeResidual": sol["relative_residual"],
                      "iterations": sol["iterations"], "converged": sol["converged"]})
    elif model == "contact-neutrality":
        fermi["model"] = "external-fallback"
    info = {"origin": "contact-modes", "modeCount": int(energy.size), "modeEnergy": energy, "modeMassXKg": mass_x,
            "modeMassZKg": mass_z, "modeNormalization": norm, "contact": contact, "fermiLevelInput": Ef_in,
            "fermiLevel": fermi}
    out = {"energy": energy, "mass_x": mass_x, "mass_z": mass_z, "modes": modes, "contact": contact,
           "kx": kx, "dkx": dkx, "temp": temp, "deg": deg, "info": info, "Ef": Ef_used}
    if not np.isfinite(Ef_used):
        info.update({"origin": "zero-placeholder", "fermiLevelUsed": np.nan})
        out["Ef"] = np.nan
        return out
    fermi["usedEV"] = Ef_used
    kz = kz_info(params, mass_z, energy, Ef_used, temp, integrate)
    out["kz"] = kz
    info.update({"fermiLevelUsed": Ef_used, "kz": kz, "kxRange": (kx.min(), kx.max()), "dkx": dkx})
    return out
