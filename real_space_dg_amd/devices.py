"""Device descriptions (the reference's ``mat`` struct) for resonant-tunnelling
diodes on the 1D-x x 1D-k path.

The reference builds ``mat`` in RTD/buildDevice2D_rtd.m (uniform material
grid ``mat.x = (0:Nx-1)*dx`` in nm, :640-648) and hands the subband energy
``Esub(valley, x, mode)`` and transverse wavefunctions ``Vsub`` to
``solve_transport_DG_2D(mat, Esub, Vsub, EfL, EfR)``
(DG_code/solve_transport_DG_2D.m:1).  This module produces exactly those
inputs for the double-barrier RTD workloads: a 60 nm device with two 3 nm
barriers around a 5 nm well, effective mass 0.067 m0, T = 300 K.
"""
from __future__ import annotations

import numpy as np


def rtd_potential(x_nm, length_nm=60.0, barrier_ev=0.3, barrier_nm=3.0, well_nm=5.0, bias_v=0.0):
    """Conduction-band edge [eV] of a centred double barrier with a linear bias drop."""
    x = np.asarray(x_nm, float)
    c = length_nm / 2
    lo1, hi1 = c - well_nm / 2 - barrier_nm, c - well_nm / 2
    lo2, hi2 = c + well_nm / 2, c + well_nm / 2 + barrier_nm
    v = np.where(((x >= lo1) & (x <= hi1)) | ((x >= lo2) & (x <= hi2)), barrier_ev, 0.0)
    return v - bias_v * x / length_nm


def rtd_1d(n_elem: int, degree: int, nk: int, bias_v: float = 0.0, barrier_ev: float = 0.3,
           length_nm: float = 60.0, fermi_ev: float = 0.1, mass: float = 0.067, temp: float = 300.0,
           fermi_model: str = "DG_MSA_3D", l_xi: float | None = None):
    """Return (mat, Esub, Vsub, EfL, EfR) for a 1D RTD.

    Discretisation: ``n_elem`` DG elements of degree ``degree`` in x (the
    material grid has n_elem*degree+1 points, so N_chi = n_elem), ``nk``
    FV cells in xi / plane waves in k.  Unless given, L_xi = (nk-1)*dx so
    that k spans [-pi/dx, pi/dx] with dx the material spacing.
    """
    nloc = degree + 1
    nx = n_elem * degree + 1
    dx = length_nm / (n_elem * degree)
    x = np.arange(nx) * dx
    if l_xi is None:
        l_xi = (nk - 1) * dx * 1e-9
    mat = {
        "type": "RTD",
        "fermi_model": fermi_model,
        "Nx": nx,
        "L_x": length_nm,
        "x": x,
        "dx": dx,
        "Ny": 1,
        "dy": 1.0,
        "Temp": temp,
        "n_of_valleys": 1,
        "n_of_modes": 1,
        "deg_factor": np.array([1.0]),
        "me_x": np.full((1, nx, 1), mass),
        "me_z_ch": mass,
        "dg": {"params": {
            "N_K_chi": nloc, "N_xi": nk, "N": nk, "Ly": l_xi,
            "Xi_solver": "FV", "Q": "EXP", "permute_shell": True,
            "MPC": 1.0, "st_init": False, "sl": 0,
        }},
    }
    esub = rtd_potential(x, length_nm, barrier_ev, bias_v=bias_v)[None, :, None]
    vsub = np.ones((1, nx, 1, 1))
    return mat, esub, vsub, fermi_ev, fermi_ev - bias_v


def rtd_sweep(n_instances: int = 64, seed: int = 0, degree: int = 3,
              n_elems=(128, 256, 384, 512), nks=(256, 512, 768, 1024),
              bias_range=(0.0, 0.4), barrier_range=(0.25, 0.35)):
    """Batched bias sweep of RTD instances with uneven grids, all from ``seed``.

    Per instance: bias ~ U[bias_range] V, elements ~ n_elems, plane waves ~ nks,
    barrier height ~ U[barrier_range] eV (numpy default_rng(seed), drawn in
    that order per instance).  Returns a list of (mat, Esub, Vsub, EfL, EfR).
    """
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_instances):
        bias = float(rng.uniform(*bias_range))
        ne = int(rng.choice(n_elems))
        nk = int(rng.choice(nks))
        barrier = float(rng.uniform(*barrier_range))
        out.append(rtd_1d(ne, degree, nk, bias_v=bias, barrier_ev=barrier))
    return out


def rtd_self_consistent(n_elem: int = 256, degree: int = 3, nk: int = 512, n_bias: int = 16,
                        bias_max: float = 0.3, contact_nm: float = 20.0, nd_contact: float = 1e24,
                        nd_intrinsic: float = 1e21, eps_r: float = 12.9, barrier_ev: float = 0.3,
                        fermi_offset_ev: float = 0.05, length_nm: float = 60.0):
    """Self-consistent RTD (Gummel loop) description.

    Returns ``mat`` with the band-edge offset ``Vc`` [eV] (double barrier),
    donor density ``Nd`` [m^-3] (n+ in the two contact regions of length
    ``contact_nm``, ``nd_intrinsic`` elsewhere), ``eps`` (relative
    permittivity), contact lengths ``Lc1``/``Lc2`` [nm], the bias list
    ``Vdd`` and the Fermi-level offset of the reference default
    EfL = Vc(1) + 0.05 eV, EfR = Vc(end) + 0.05 eV - bias
    (RTD/solve2D_rtd_self_consistent.m:109-110).  The transport contacts
    use the transverse-integrated (Tsu-Esaki) inflow so that the density is
    a volume density comparable with Nd.
    """
    mat, _, _, _, _ = rtd_1d(n_elem, degree, nk, bias_v=0.0, barrier_ev=barrier_ev, length_nm=length_nm,
                             fermi_model="TSU_ESAKI")
    x = mat["x"]
    mat["Vc"] = rtd_potential(x, length_nm, barrier_ev, bias_v=0.0)
    mat["Nd"] = np.where((x <= contact_nm) | (x >= length_nm - contact_nm), nd_contact, nd_intrinsic)
    mat["eps"] = np.full(x.size, eps_r)
    mat["Lc1"] = contact_nm
    mat["Lc2"] = contact_nm
    mat["Vdd"] = np.linspace(0.0, bias_max, n_bias)
    mat["fermi_offset"] = fermi_offset_ev
    return mat


def fet_slab(length_nm: float = 40.0, width_nm: float = 20.0, n_el_x: int = 64, n_el_y: int = 32, degree: int = 1,
             n_rho: int = 64, barrier_ev: float = 0.2, sigma_nm: float = 5.0, mass: float = 0.067,
             temp: float = 300.0, flux: str = "rusanov"):
    """Full-2D FET slab (the ``mat`` of solve_transport_DG_full2D,
    DG_code/initParams_full2D.m:42-115): ``length_nm`` x ``width_nm`` with
    n_el_x x n_el_y rectangular DG elements of ``degree`` (material grid
    n_el*degree+1 points per axis, nm), n_rho x n_rho relative FV cells over
    the default relative lengths (the Y extent), and a Gaussian gate barrier
    V = barrier * exp(-((x-L/2)^2 + (y-W/2)^2)/(2 sigma^2)) [eV].
    Returns (mat, V)."""
    nx, ny = n_el_x * degree + 1, n_el_y * degree + 1
    x = np.linspace(0.0, length_nm, nx)
    y = np.linspace(0.0, width_nm, ny)
    X, Y = np.meshgrid(x, y, indexing="ij")
    V = barrier_ev * np.exp(-((X - length_nm / 2) ** 2 + (Y - width_nm / 2) ** 2) / (2 * sigma_nm ** 2))
    mat = {"type": "FET-slab", "Nx": nx, "Ny": ny, "x": x, "y": y, "dx": x[1] - x[0], "dy": y[1] - y[0],
           "Temp": temp, "deg_factor": 1, "me_x_ch": mass, "me_y_ch": mass, "V": V,
           "dg": {"params": {"N_K_X": degree + 1, "N_K_Y": degree + 1, "N_rho_x": n_rho, "N_rho_y": n_rho,
                             "full2D_fluxType": flux}}}
    return mat, V


def fet_slab_self_consistent(length_nm: float = 40.0, width_nm: float = 20.0, n_el_x: int = 64, n_el_y: int = 32,
                             degree: int = 1, n_rho: int = 64, barrier_ev: float = 0.2, sigma_nm: float = 5.0,
                             mass: float = 0.067, temp: float = 300.0, eps_r: float = 12.9,
                             contact_nm: float = 10.0, nd_contact: float = 1e24, nd_channel: float = 1e21,
                             n_modes: int = 4, nkz: int = 101):
    """The FET slab of ``fet_slab`` with the data of a self-consistent run
    (the RTD conventions of RTD/solve2D_rtd_self_consistent.m): band offset
    ``Vc`` = the Gaussian gate barrier [eV], relative permittivity ``eps``,
    donors ``Nd`` [m^-3] (``nd_contact`` within ``contact_nm`` of either
    contact, ``nd_channel`` between), and contact-mode reservoirs (the
    ``n_modes`` lowest transverse modes of each contact edge, k_z-integrated,
    Fermi levels from contact neutrality with the contact donors' sheet
    density ``nd_contact * width``).  The electron energy is Vc - phi.
    Returns ``mat``."""
    mat, V = fet_slab(length_nm, width_nm, n_el_x, n_el_y, degree, n_rho, barrier_ev, sigma_nm, mass, temp)
    x, y = mat["x"], mat["y"]
    X = np.repeat(x[:, None], y.size, axis=1)
    shape = (x.size, y.size)
    mat.update({"Vc": V, "eps": np.full(shape, eps_r),
                "Nd": np.where((X <= contact_nm) | (X >= length_nm - contact_nm), nd_contact, nd_channel),
                "me_z_ch": mass, "me_x": np.full(shape, mass), "me_y": np.full(shape, mass),
                "me_z": np.full(shape, mass), "n_of_modes": n_modes})
    sheet = nd_contact * width_nm * 1e-9
    mat["dg"]["params"].update({"full2D_reservoirModel": "contact-modes",
                                "full2D_contactModeFermiModel": "contact-neutrality",
                                "full2D_contactModeIntegrateKz": 1, "full2D_contactModeCount": n_modes,
                                "full2D_contactMode_Nkz": nkz,
                                "full2D_sourceContactSheetDensity": sheet, "full2D_drainContactSheetDensity": sheet})
    return mat
