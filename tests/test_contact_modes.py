"""Contact-mode reservoirs of the full-2D inflow boundary (host setup), pinned
by the assertions of the reference's contactModeReservoirFull2DTest.m on its
own small setups, and the periodic Fourier branch by its closed forms
(plane-wave energies, Y-independent density, double-cover periodicity)."""
import copy

import numpy as np
import pytest

from real_space_dg_amd import contact_modes as CM
from real_space_dg_amd import full2d as P

HBAR, M0, QE, KB = 1.054571817e-34, 9.1093837015e-31, 1.602176634e-19, 1.38064852e-23


def contact_mode_mat():
    """makeContactModeMat (contactModeReservoirFull2DTest.m:280-317)."""
    nx, ny = 5, 9
    x, y = np.linspace(0, 4, nx), np.linspace(-2, 2, ny)
    X, Y = np.meshgrid(x, y, indexing="ij")
    V = 0.05 + 0.08 * Y ** 2
    V[:, [0, -1]] = 1.0
    params = {"full2D_coordinate_scale": 1e-9, "N_K_chi": 3, "N_K_X": 3, "N_K_Y": 3, "N_rho_x": 5, "N_rho_y": 5,
              "L_rho_x": 8e-9, "L_rho_y": 4e-9, "full2D_reservoirModel": "contact-modes",
              "full2D_contactModeCount": 2, "full2D_contactModeIntegrateKz": False, "full2D_contactMode_Nkx": 7}
    return {"type": "ContactModeReservoirFull2DTest", "Nx": nx, "Ny": ny, "x": x, "y": y, "dx": x[1] - x[0],
            "dy": y[1] - y[0], "Temp": 300, "deg_factor": 1, "me_x_ch": 1.0, "me_y_ch": 1.0, "me_z_ch": 1.0,
            "n_of_modes": 2, "V": V, "Nd": np.ones((nx, ny)), "me_x": np.ones((1, nx, ny)),
            "me_y": np.ones((1, nx, ny)), "me_z": np.ones((1, nx, ny)), "dg": {"params": params}}


def flat_mat():
    """makeFlatContactModeMat (:319-326)."""
    m = contact_mode_mat()
    m["V"] = np.full_like(m["V"], 0.05)
    m["dg"]["params"].update({"full2D_contactModeCount": 1, "full2D_contactModeFermiModel": "external",
                              "N_rho_y": 7, "L_rho_y": 7 * m["dy"] * 1e-9})
    return m


def masked_mat():
    """makeMaskedContactModeMat (:328-334)."""
    m = contact_mode_mat()
    nd = np.zeros((m["Nx"], m["Ny"]))
    cy = (m["y"] >= -1) & (m["y"] <= 1)
    nd[0, cy] = 1
    nd[-1, cy] = 1
    m["Nd"] = nd
    return m


def reservoir(mat, EfL, EfR, side="source"):
    s = P.Full2DSetup(mat, mat["V"], EfL, EfR)
    rho = s.boundary[side]["rho"]
    R = rho.T.reshape(s.Y.ndof, s.RY.n, s.RX.n)    # [Y][rho_y][rho_x]: MATLAB's R(rho_x, rho_y, Y)
    return s, rho, R, s.reservoir_info.get(side)


def test_mode_extension_is_continuous_at_both_material_edges():
    m = flat_mat()
    m["dg"]["params"].update({"N_rho_y": 3, "L_rho_y": 3e-6 * m["dy"] * 1e-9})
    s, rho, R, _ = reservoir(m, 1, 1)
    edge = R[[0, -1]][:, [0, 2], 2]           # R(3, [1 3], [1 end])
    zero = R[[0, -1]][:, 1, 2]
    assert np.allclose(edge / zero[:, None], 1.0, rtol=0, atol=3e-6)


def test_mode_extension_reaches_zero_at_hamiltonian_ghost_nodes():
    m = flat_mat()
    s, rho, R, info = reservoir(m, 1, 1)
    half = R[[0, -1]][:, [2, 4], 2]
    zero = R[[0, -1]][:, 3, 2]
    expected = 0.25 * (1 + 2 * np.cos(np.pi / (m["Ny"] + 1)))
    assert np.allclose(half / zero[:, None], expected, rtol=0, atol=1e-12)
    assert np.allclose(info["contact"]["modeZeroBoundaryY"],
                       [(m["y"][0] - m["dy"]) * 1e-9, (m["y"][-1] + m["dy"]) * 1e-9], rtol=0, atol=1e-22)


def test_mode_extension_vanishes_at_and_beyond_ghost_nodes():
    m = flat_mat()
    s, rho, R, _ = reservoir(m, 1, 1)
    out = R[[0, -1]][:, [0, 1, 5, 6], 2]
    zero = R[[0, -1]][:, 3, 2]
    assert np.allclose(out / zero[:, None], 0.0, rtol=0, atol=1e-12)


def test_mode_extension_preserves_interior_density_and_energies():
    m = flat_mat()
    s, rho, R, info = reservoir(m, 1, 1)
    h = m["dy"] * 1e-9
    theta = np.pi / (m["Ny"] + 1)
    samples = np.sin(theta * np.arange(1, m["Ny"] + 1))
    wnorm = np.sum(samples ** 2) - 0.5 * (samples[0] ** 2 + samples[-1] ** 2)
    node_index = (s.Y.nodes - m["y"][0] * 1e-9) / h + 1
    expected = np.sin(theta * node_index) ** 2 / wnorm
    sampled = R[:, 3, 2].real * h / info["rhoXModeProfiles"][2, 0]
    t = HBAR ** 2 / (2 * M0 * m["me_y_ch"] * QE * h ** 2)
    assert np.allclose(sampled, expected, rtol=0, atol=1e-12)
    assert np.allclose(info["modeEnergy"], m["V"][0, 0] + 2 * t * (1 - np.cos(theta)), rtol=0, atol=1e-12)
    assert info["fermiLevelUsed"] == pytest.approx(1, abs=1e-14)


def test_contact_mode_reservoir_shape_origin_and_y_variation():
    m = contact_mode_mat()
    s, rho, R, info = reservoir(m, 2.0, 1.8)
    assert rho.shape == (s.nrel, s.Y.ndof) and np.linalg.norm(rho) > 0
    assert info["origin"] == "contact-modes" and info["modeCount"] == 2 and info["contact"]["HNnz"] > 0
    col = np.linalg.norm(rho, axis=0)
    assert col.max() - col.min() > 0
    # continuum normalisation of the modes
    assert np.allclose(info["modeNormalization"]["normalizedIntegral"], 1.0, rtol=0, atol=1e-10)


def test_kz_cutoff_covers_occupied_modes_and_thermal_tail():
    m = contact_mode_mat()
    m["dg"]["params"].update({"full2D_contactModeIntegrateKz": True, "full2D_contactMode_Nkz": 31,
                              "full2D_contactModeKzEnergyWindow": 0.05, "full2D_contactModeKzThermalTail": 8})
    Ef = 2.0
    s, rho, R, info = reservoir(m, Ef, Ef)
    kz = info["kz"]
    adaptive = max(Ef - info["modeEnergy"].min(), 0) + 8 * KB * m["Temp"] / QE
    assert kz["integrate"] and kz["minimumEnergyWindowEV"] == 0.05 and kz["thermalTailKBT"] == 8
    assert kz["adaptiveEnergyWindowEV"] == pytest.approx(adaptive, rel=1e-12)
    assert kz["energyWindowEV"] == pytest.approx(max(0.05, adaptive), rel=1e-12)
    assert kz["nKz"] == 31


def test_fermi_level_solved_from_contact_neutrality():
    m = contact_mode_mat()
    m.update({"N_s": 2e18, "N_d": 2e18, "W_c": 4})
    m["dg"]["params"].update({"full2D_contactModeIntegrateKz": True, "full2D_contactMode_Nkz": 101,
                              "full2D_contactModeFermiModel": "contact-neutrality"})
    s = P.Full2DSetup(m, m["V"], 5.0, -2.0)
    src, drn = s.reservoir_info["source"], s.reservoir_info["drain"]
    assert src["fermiLevel"]["model"] == "contact-neutrality"
    assert src["fermiLevel"]["targetSheetDensityM2"] == pytest.approx(2e18 * 1e6 * 4 * 1e-9, rel=1e-14)
    assert src["fermiLevel"]["relativeResidual"] < 1e-9 and drn["fermiLevel"]["relativeResidual"] < 1e-9
    assert src["fermiLevelUsed"] != 5.0 and drn["fermiLevelUsed"] != -2.0
    assert src["fermiLevelUsed"] == pytest.approx(drn["fermiLevelUsed"], abs=1e-10)
    # the solved level reproduces the target by the closed-form sheet density
    dens, _ = CM.sheet_density(src["fermiLevelUsed"], src["modeEnergy"], src["modeMassXKg"], src["modeMassZKg"],
                               300.0, 1.0)
    assert dens == pytest.approx(2e18 * 1e6 * 4 * 1e-9, rel=1e-9)


def test_external_fermi_level_remains_selectable():
    m = contact_mode_mat()
    m.update({"N_s": 2e18, "N_d": 2e18, "W_c": 4})
    m["dg"]["params"].update({"full2D_contactModeIntegrateKz": True, "full2D_contactMode_Nkz": 31,
                              "full2D_contactModeFermiModel": "external"})
    s, rho, R, info = reservoir(m, 1.25, 1.25)
    assert info["fermiLevel"]["model"] == "external" and info["fermiLevelUsed"] == 1.25


def test_differs_from_material_default_and_explicit_rho_overrides():
    m = contact_mode_mat()
    md = copy.deepcopy(m)
    md["dg"]["params"]["full2D_reservoirModel"] = "material-default"
    a = P.Full2DSetup(m, m["V"], 2.0, 1.8).boundary["source"]["rho"]
    b = P.Full2DSetup(md, md["V"], 2.0, 1.8).boundary["source"]["rho"]
    assert np.linalg.norm(a - b) > 0
    s0 = P.Full2DSetup(m, m["V"], 2.0, 1.8)
    mo = copy.deepcopy(m)
    mo["dg"]["params"].update({"full2D_sourceRho": np.ones((s0.nrel, s0.Y.ndof)),
                               "full2D_drainRho": 2 * np.ones((s0.nrel, s0.Y.ndof))})
    s = P.Full2DSetup(mo, mo["V"], 2.0, 1.8)
    assert np.array_equal(s.boundary["source"]["rho"], np.ones((s.nrel, s.Y.ndof)))
    assert np.array_equal(s.boundary["drain"]["rho"], 2 * np.ones((s.nrel, s.Y.ndof)))


def test_contact_mode_reservoir_respects_contact_mask():
    m = masked_mat()
    s = P.Full2DSetup(m, m["V"], 2.0, 1.8)
    b = s.boundary["source"]
    mask = b["mask"]
    assert mask.any() and not mask.all()
    assert np.linalg.norm(b["rho"][:, mask]) > 0
    assert np.allclose(b["rho"][:, ~mask], 0, rtol=0, atol=1e-12)


def periodic_contact_mat(edge_profile=False):
    """A small periodic-Y slab with constant masses and sheet-density
    neutrality, the setting of the reference's production RTD run."""
    nx, ny = 9, 9
    x, y = np.linspace(0, 8, nx), np.linspace(0, 3, ny)
    X, Y = np.meshgrid(x, y, indexing="ij")
    V = 0.2 + 0.01 * X
    if edge_profile:
        V = V + 0.04 * np.cos(2 * np.pi * Y / 3.0)
    params = {"full2D_coordinate_scale": 1e-9, "N_K_X": 2, "N_K_Y": 3, "N_el_X": 4, "N_el_Y": 2,
              "N_rho_x": 9, "N_rho_y": 12, "L_rho_x": 1e-8, "full2D_Y_boundary": "periodic",
              "full2D_X_flux": "upwind", "full2D_Y_flux": "upwind", "full2D_reservoirModel": "contact-modes",
              "full2D_contactModeFermiModel": "contact-neutrality", "full2D_contactModeIntegrateKz": 1,
              "full2D_contactMode_Nkx": 40, "full2D_contactMode_Nkz": 51,
              "full2D_sourceContactSheetDensity": 6e15, "full2D_drainContactSheetDensity": 6e15}
    return {"Nx": nx, "Ny": ny, "x": x, "y": y, "dx": 1.0, "dy": 0.375, "Temp": 300, "deg_factor": 1,
            "me_x_ch": 0.063, "me_y_ch": 0.063, "me_z_ch": 0.063, "V": V,
            "me_x": np.full((nx, ny), 0.063), "dg": {"params": params}}


def test_periodic_plane_wave_modes_energies_and_density():
    m = periodic_contact_mat()
    s = P.Full2DSetup(m, m["V"], 0.3, 0.3)
    info = s.reservoir_info["source"]
    W = 3e-9
    order = min((m["Ny"] - 2) // 2, (12 - 2) // 4)
    k = 2 * np.pi / W * np.arange(-order, order + 1)
    e = np.sort(m["V"][0, 0] + HBAR ** 2 * k ** 2 / (2 * 0.063 * M0 * QE))
    assert np.allclose(info["modeEnergy"], e, rtol=0, atol=1e-14)
    assert np.allclose(info["modeNormalization"]["normalizedIntegral"], 1.0, rtol=1e-12)
    assert info["fermiLevel"]["relativeResidual"] < 1e-9
    # plane waves: phi(Y + rho_y/2) conj(phi(Y - rho_y/2)) = exp(i k rho_y)/W, the same at every Y node
    rho = s.boundary["source"]["rho"]
    assert np.allclose(rho, rho[:, :1], rtol=0, atol=1e-12 * np.abs(rho).max())


@pytest.mark.parametrize("profile", [False, True])
def test_periodic_reservoir_is_double_cover_periodic(profile):
    """rho(Y + W/2, rho_y + W) = rho(Y, rho_y) (get_PeriodicYStateError_full2D.m)."""
    m = periodic_contact_mat(edge_profile=profile)
    s = P.Full2DSetup(m, m["V"], 0.3, 0.3)
    for side in ("source", "drain"):
        R = s.boundary[side]["rho"].T.reshape(s.Y.ndof, s.RY.n, s.RX.n)
        sh = R[np.roll(np.arange(s.Y.ndof), s.Y.ndof // 2)][:, np.roll(np.arange(s.RY.n), s.RY.n // 2)]
        assert np.linalg.norm(R) > 0
        assert np.linalg.norm(sh - R) <= 1e-12 * np.linalg.norm(R)
    if profile:   # a Y-dependent contact potential diagonalised in the Fourier basis
        info = s.reservoir_info["source"]
        H = info["contact"]["H"]
        assert np.allclose(np.linalg.eigvalsh(H), info["modeEnergy"], rtol=0, atol=1e-12)


def test_sim_res_contact_fermi_levels_follow_the_bias():
    """The reference's production run (periodic Y, contact-mode reservoirs,
    sheet-density neutrality): the drain edge potential is the source edge
    shifted by -bias, so the solved drain Fermi level is the source level
    minus the bias, and both reproduce the 6e15 m^-2 target."""
    from tests._sim_res import load
    mat, cases = load()
    levels = []
    for c in cases:
        s = P.Full2DSetup(mat, c["V"], c["EfL"], c["EfR"])
        src, drn = s.reservoir_info["source"], s.reservoir_info["drain"]
        assert src["fermiLevel"]["relativeResidual"] < 1e-9 and drn["fermiLevel"]["relativeResidual"] < 1e-9
        assert src["modeCount"] == 5 and s.n_total == 121 * 12 * 360 * 6
        assert drn["fermiLevelUsed"] == pytest.approx(src["fermiLevelUsed"] - c["bias"], abs=1e-9)
        levels.append(src["fermiLevelUsed"])
    assert np.ptp(levels) < 1e-12
