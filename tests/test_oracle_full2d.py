"""Pins of the full-2D oracle: the reference's own test assertions
(DG_code/tests/full2DDiffOperatorTest.m, full2DDriftSystemSolveTest.m,
test_full2D_boundaries.m) on their small setups, plus invariants."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import full2d as F


def small_diff_mat():
    """makeSmallDiffMat (full2DDiffOperatorTest.m:272-303)."""
    nx = ny = 5
    x = np.linspace(0, 1, nx)
    y = np.linspace(-0.5, 0.5, ny)
    return {"type": "Full2DDiffTest", "Nx": nx, "Ny": ny, "x": x, "y": y, "dx": x[1] - x[0], "dy": y[1] - y[0],
            "Temp": 300, "deg_factor": 1, "Eg_c": 0.74, "Eg_ox": 8.8, "me_x_ch": 0.041, "me_y_ch": 0.041,
            "V": np.zeros((nx, ny)), "Nd": np.ones((nx, ny)), "me_x": 0.041 * np.ones((1, nx, ny)),
            "me_y": 0.041 * np.ones((1, nx, ny)),
            "dg": {"params": {"full2D_coordinate_scale": 1, "N_K_chi": 3, "N_K_X": 3, "N_K_Y": 3, "N_rho_x": 5,
                              "N_rho_y": 5, "L_rho_x": 1, "L_rho_y": 1, "full2D_thetaLF": 0.03}}}


def small_mat():
    """makeSmallMat (full2DDriftSystemSolveTest.m:561-598)."""
    m = small_diff_mat()
    X, Y = np.meshgrid(m["x"], m["y"], indexing="ij")
    m.update({"type": "Full2DDriftSystemSolveTest", "me_y_ch": 0.052, "V": 0.1 * X + 0.03 * Y ** 2,
              "me_y": 0.052 * np.ones((1, 5, 5))})
    m["dg"]["params"].update({"full2D_fluxType": "rusanov", "full2D_thetaLF": 0.03, "full2D_Q_diff_x": 1,
                              "full2D_Q_diff_y": 0.7, "full2D_Q_drift": 1, "full2D_drift_scale": 1,
                              "full2D_cap_scale": 0.1})
    return m


def det_vec(n):
    k = np.arange(1, n + 1)
    return np.sin(k) + 1j * np.cos(k)


def diff_setup(mat, EfL=0.2, EfR=0.1):
    p = F.init_params_full2d(mat, mat["V"], EfL, EfR)
    b = F.boundary_full2d(mat, p, EfL, EfR, mat["V"])
    parts, rparts, info = F.get_diff_full2d(mat, p, b)
    return p, b, parts, rparts, info, F.assemble_parts(parts, p), F.assemble_rhs(rparts, p)


def test_rusanov_assembly_and_central_difference():
    m = small_diff_mat(); m["dg"]["params"]["full2D_fluxType"] = "rusanov"
    p, b, parts, rparts, info, A, rhs = diff_setup(m)
    assert A.shape == (p["nTotal"], p["nTotal"]) and rhs.shape == (p["nTotal"],)
    assert A.nnz > 0 and np.linalg.norm(rhs) > 0
    assert info["betaLF"][0] > 0 and info["betaLF"][1] > 0
    mc = small_diff_mat(); mc["dg"]["params"]["full2D_fluxType"] = "central"
    _, _, _, _, info_c, Ac, _ = diff_setup(mc)
    assert info_c["betaLF"] == (0.0, 0.0)
    assert sp.linalg.norm(A - Ac) > 0


def test_rho_flux_fallback_and_directional_overrides():
    m = small_diff_mat(); m["dg"]["params"]["rho_flux"] = "central"
    *_, info, A, _ = diff_setup(m)
    assert info["fluxTypeX"] == "central" and info["betaLF"] == (0.0, 0.0)
    m = small_diff_mat(); m["dg"]["params"].update({"rho_flux": "central", "full2D_Y_flux": "upwind"})
    p, b, parts, _, info, Aup, _ = diff_setup(m)
    names = [q[0] for q in parts]
    assert (info["fluxTypeX"], info["fluxTypeY"]) == ("central", "upwind")
    assert "Y interior characteristic upwind" in names and sp.linalg.norm(Aup - A) > 0
    u = det_vec(p["nTotal"])
    assert np.allclose(F.apply_parts(parts, u, p), Aup @ u, atol=1e-10, rtol=0)
    m = small_diff_mat(); m["dg"]["params"].update({"rho_flux": "rusanov", "full2D_Y_flux": "central",
                                                    "full2D_X_flux": "upwind"})
    p, _, parts, _, info, Ax, _ = diff_setup(m)
    names = [q[0] for q in parts]
    assert "X interior characteristic upwind" in names and "X interior Rusanov" not in names
    assert info["betaLF"] == (0.0, 0.0)
    assert np.allclose(F.apply_parts(parts, u, p), Ax @ u, atol=1e-10, rtol=0)
    m = small_diff_mat(); m["dg"]["params"].update({"rho_flux": "rusanov", "full2D_Y_flux": "central"})
    _, _, parts, _, info, _, _ = diff_setup(m)
    names = [q[0] for q in parts]
    assert info["betaLF"][0] > 0 and info["betaLF"][1] == 0
    assert "X interior Rusanov" in names and "Y interior Rusanov" not in names


def test_matrix_free_parts_match_assembly():
    m = small_diff_mat(); m["dg"]["params"]["full2D_fluxType"] = "rusanov"
    p, b, parts, _, _, A, _ = diff_setup(m)
    u = det_vec(p["nTotal"])
    assert np.allclose(F.apply_parts(parts, u, p), A @ u, atol=1e-10, rtol=0)


def test_source_rhs_uses_only_contact_mask():
    m = small_diff_mat()
    p = F.init_params_full2d(m, m["V"], 0.2, 0.1)
    nrel, nyd = p["relative"]["nDof"], p["dg"]["Y"]["nDof"]
    mask = np.zeros(nyd, bool); mask[1:4] = True
    m["dg"]["params"].update({"full2D_fluxType": "rusanov",
                              "full2D_sourceRho": np.repeat(np.arange(1, nrel + 1)[:, None], nyd, axis=1),
                              "full2D_drainRho": np.zeros((nrel, nyd)), "full2D_sourceContactMask": mask,
                              "full2D_drainContactMask": np.zeros(nyd, bool)})
    b = F.boundary_full2d(m, p, 0.2, 0.1, m["V"])
    _, rparts, _ = F.get_diff_full2d(m, p, b)
    rhs = F.assemble_rhs(rparts, p).reshape(p["dg"]["nCenterDof"], nrel).T
    nxd, nl = p["dg"]["X"]["nDof"], p["dg"]["X"]["nLocal"]
    rows = lambda ymask: [x + nxd * iy for iy in np.nonzero(ymask)[0] for x in range(nl)]
    assert np.allclose(rhs[:, rows(~mask)], 0, atol=1e-12)
    assert np.linalg.norm(rhs[:, rows(mask)]) > 0


def test_zero_inflow_omits_reflection_and_specular_is_default():
    m = small_diff_mat(); m["dg"]["params"]["full2D_Y_boundary"] = "zero-inflow"
    _, _, parts_z, _, info_z, Az, _ = diff_setup(m)
    assert info_z["yBoundaryType"] == "zero-inflow"
    assert not any("specular reflection correction" in q[0] for q in parts_z)
    m = small_diff_mat()
    _, _, parts_s, _, info_s, As, _ = diff_setup(m)
    assert info_s["yBoundaryType"] == "specular"
    assert any("specular reflection correction" in q[0] for q in parts_s)
    assert sp.linalg.norm(Az - As) > 0


def test_characteristic_split_identities():
    m = small_diff_mat()
    p = F.init_params_full2d(m, m["V"], 0.2, 0.1)
    A = p["relative"]["Ax1D"].toarray()
    Ap, Am, lam = F.split_characteristic(A)
    assert np.allclose(Ap + Am, A, atol=1e-12)
    assert np.allclose(Ap @ Am, 0, atol=1e-12)
    assert np.all(np.linalg.eigvalsh((Ap + Ap.conj().T) / 2) > -1e-12)   # A+ is positive semi-definite


def test_inflow_known_answer_vectors():
    """testSourceInflowOnlyUsesSourceData / testDrainInflowOnlyUsesDrainData
    (test_full2D_boundaries.m:51-77): exact flux vectors on the toy problem."""
    rel = {"NrhoX": 4, "NrhoY": 1, "Ax1D": sp.diags([[-2.0, -1.0, 1.0, 2.0]], [0]), "Ay1D": sp.csr_matrix((1, 1))}
    src = {"rhoBoundary": np.repeat(np.array([10.0, 20, 30, 40])[:, None], 3, axis=1),
           "normalFlux": F.normal_flux(rel, (-1.0, 0.0))}
    drn = {"rhoBoundary": np.repeat(np.array([100.0, 200, 300, 400])[:, None], 3, axis=1),
           "normalFlux": F.normal_flux(rel, (1.0, 0.0))}
    inside = np.ones((4, 3))
    assert np.allclose(F.apply_inflow_flux(src, inside), np.repeat(np.array([[2.0], [1], [-30], [-80]]), 3, 1),
                       atol=1e-12)
    assert np.allclose(F.apply_inflow_flux(drn, inside), np.repeat(np.array([[-200.0], [-200], [1], [2]]), 3, 1),
                       atol=1e-12)
    assert np.count_nonzero(src["normalFlux"]["inflowMask"]) == 2
    assert np.count_nonzero(drn["normalFlux"]["inflowMask"]) == 2


def test_material_default_reservoir_is_nonzero_and_reflection_is_involution():
    m = small_diff_mat()
    m["dg"]["params"].update({"N_rho_x": 7, "N_rho_y": 5})
    m["V"] = np.zeros((5, 5)); m["V"][-1, :] = 0.05
    p = F.init_params_full2d(m, m["V"], 0.2, 0.1)
    b = F.boundary_full2d(m, p, 0.2, 0.1, m["V"])
    for s in ("source", "drain"):
        assert b[s]["rhoBoundary"].shape == (p["relative"]["nDof"], p["dg"]["Y"]["nDof"])
        assert np.linalg.norm(b[s]["rhoBoundary"]) > 0
    Rx, Ry = F.reflection(p)
    I = sp.identity(p["relative"]["nDof"])
    assert Ry.nnz == p["relative"]["nDof"] and sp.linalg.norm(Ry @ Ry - I) < 1e-12
    # the reservoir column is even in rho_x and rho_y (cosine transform)
    col = b["source"]["rhoBoundary"][:, 0]
    assert np.allclose(Rx @ col, col) and np.allclose(Ry @ col, col)


def test_cap_acts_only_near_relative_boundaries():
    m = small_diff_mat(); m["dg"]["params"].update({"N_rho_x": 21, "N_rho_y": 19})
    p = F.init_params_full2d(m, m["V"], 0.1, 0.0)
    cap = F.cap_full2d(p).reshape(19, 21)
    rx, ry = p["relative"]["rhoX"], p["relative"]["rhoY"]
    from oracle import wigner1d as W
    mx = W.get_cap(rx["cells"], rx["length"], rx["delta"]) != 0
    my = W.get_cap(ry["cells"], ry["length"], ry["delta"]) != 0
    interior = ~(my[:, None] | mx[None, :])
    assert np.count_nonzero(cap) > 0 and np.allclose(cap[interior], 0, atol=1e-14)


def test_drift_constant_potential_without_cap_is_zero_and_zero_separation_vanishes():
    m = small_mat(); m["V"] = 0.37 * np.ones((5, 5)); m["dg"]["params"]["full2D_cap_scale"] = 0
    p = F.init_params_full2d(m, m["V"], 0.2, 0.1)
    G = F.DriftFull2D(m, p, m["V"])
    u = det_vec(p["nTotal"])
    assert np.max(np.abs(G.apply(u))) < 1e-13
    m = small_mat(); m["dg"]["params"]["full2D_potentialDiscretization"] = "collocated"
    p = F.init_params_full2d(m, m["V"], 0.2, 0.1)
    d = F.DriftFull2D(m, p, m["V"]).diagonal()
    rel = p["relative"]
    ix0 = np.argmin(np.abs(rel["rhoX"]["cells"])); iy0 = np.argmin(np.abs(rel["rhoY"]["cells"]))
    ids = ix0 + iy0 * rel["NrhoX"] + np.arange(p["dg"]["nCenterDof"]) * rel["nDof"]
    # zeroRelativeSeparationHasNoPotentialDifference (full2DDriftSystemSolveTest.m:97-113)
    # compares the whole diagonal with 0; on its 5-cell rho grid the CAP of
    # get_CAP.m is nonzero at rho = 0 (0.125 per axis), so the potential part
    # alone is what vanishes there.
    D = F.DriftFull2D(m, p, m["V"])
    cap0 = D.cap_coef * D.cap[ix0 + iy0 * rel["NrhoX"]]
    assert abs(cap0 - 0.025) < 1e-15
    assert np.allclose(d[ids] - cap0, 0, atol=1e-10)


def test_fv_drift_nine_point_relative_stencil_and_independent_assembly():
    m = small_mat(); m["dg"]["params"]["full2D_cap_scale"] = 0
    p = F.init_params_full2d(m, m["V"], 0.2, 0.1)
    D = F.DriftFull2D(m, p, m["V"])
    G = D.assemble().toarray()
    nr, nc = p["relative"]["nDof"], p["dg"]["nCenterDof"]
    blk = G[nr:2 * nr, nr:2 * nr]          # relative block of centre DOF 2 on itself
    r, c = np.nonzero(np.abs(blk) > 0)
    nx = p["relative"]["NrhoX"]
    assert np.count_nonzero(blk - np.diag(np.diag(blk))) > 0
    assert np.max(np.abs(r % nx - c % nx)) <= 1 and np.max(np.abs(r // nx - c // nx)) <= 1
    # independent dense assembly from the formula (1/16) P' diag(dV_q) P (x) W-projection
    Gd = np.zeros((nr * nc, nr * nc), complex)
    for ids, Xq, Yq in D.elements:
        dV = D.delta_v(Xq, Yq)
        for q in range(Xq.size):
            Rq = D.pot_coef / 16.0 * (D.P.T @ sp.diags(dV[:, q]) @ D.P).toarray()
            Cq = np.outer(D.centerProjection[q], D.interpolation[q])   # out local x in local
            for a, ga in enumerate(ids):
                for bb, gb in enumerate(ids):
                    Gd[ga * nr:(ga + 1) * nr, gb * nr:(gb + 1) * nr] += Cq[a, bb] * Rq
    assert np.allclose(G, Gd, atol=1e-12 * np.abs(Gd).max())


def test_small_direct_solve_residual():
    m = small_mat()
    p, A, rhs, rho = F.solve_direct_full2d(m, m["V"], 0.2, 0.1)
    assert np.linalg.norm(A @ rho - rhs) / np.linalg.norm(rhs) < 1e-9
    n, jx, jy = F.observables_full2d(rho, p, m)
    assert n.shape == (p["dg"]["X"]["nDof"], p["dg"]["Y"]["nDof"]) == jx.shape == jy.shape
    assert np.all(np.isfinite(n)) and np.abs(n).max() > 0


@pytest.mark.parametrize("flux", ["central", "rusanov"])
def test_constant_state_has_no_interior_transport_residual(flux):
    """A state constant in X and Y: DG derivative and jump terms vanish in every
    element away from the physical boundary (H1 1 = H2_interior 1 = 0 there);
    a sign or index error in K1/K3/K5/K6 breaks this."""
    m = small_diff_mat()
    m["x"] = np.linspace(0, 1, 9); m["Nx"] = 9
    m["y"] = np.linspace(-0.5, 0.5, 9); m["Ny"] = 9
    m["V"] = np.zeros((9, 9)); m["Nd"] = np.ones((9, 9))
    m["me_x"] = 0.041 * np.ones((1, 9, 9)); m["me_y"] = 0.041 * np.ones((1, 9, 9))
    m["dg"]["params"]["full2D_fluxType"] = flux
    p = F.init_params_full2d(m, m["V"], 0.2, 0.1)
    b = F.boundary_full2d(m, p, 0.2, 0.1, m["V"])
    parts, _, _ = F.get_diff_full2d(m, p, b)
    nr, nc = p["relative"]["nDof"], p["dg"]["nCenterDof"]
    f = np.random.default_rng(0).standard_normal(nr) + 1j * np.random.default_rng(1).standard_normal(nr)
    u = np.tile(f, nc)
    Y = F.apply_parts(parts, u, p).reshape(nc, nr)
    X, Yax = p["dg"]["X"], p["dg"]["Y"]
    interior = [ix + X["nDof"] * iy
                for iy in range(Yax["nLocal"], Yax["nDof"] - Yax["nLocal"])
                for ix in range(X["nLocal"], X["nDof"] - X["nLocal"])]
    assert len(interior) > 0
    assert np.max(np.abs(Y[interior])) < 1e-12 * np.max(np.abs(Y))
    assert np.max(np.abs(Y)) > 0   # the boundary elements do see the inflow penalty


@pytest.mark.parametrize("flux", ["rusanov", "upwind"])
def test_kron_diagonal_and_row_sums_match_assembly(flux):
    """assembleKronDiagonal equals diag(A); assembleKronRowAbsSum bounds the
    row sums of |A| (full2DVariableMassTest.m:82-91 checks the bound)."""
    m = small_mat(); m["dg"]["params"]["full2D_X_flux"] = flux
    m["dg"]["params"]["full2D_Y_flux"] = flux
    p, b, parts, rparts, info, A, rhs = diff_setup(m)
    assert np.allclose(F.kron_diagonal(parts, p), A.diagonal(), rtol=0, atol=1e-12 * abs(A).max())
    rs = np.asarray(abs(A).sum(axis=1)).ravel()
    bound = F.kron_row_abs_sum(parts, p)
    assert np.all(bound >= rs - 1e-12 * rs.max())
    # a single part's bound is exact
    one = [parts[0]]
    A1 = F.assemble_parts(one, p)
    assert np.allclose(F.kron_row_abs_sum(one, p), np.asarray(abs(A1).sum(axis=1)).ravel(), rtol=1e-13, atol=0)


@pytest.mark.parametrize("mode", ["fv-consistent", "collocated"])
def test_drift_diagonal_and_row_bound(mode):
    """fvConsistentDiagonal equals diag of the assembled drift; the row bound
    is >= the row sums of |G| and exact for the collocated diagonal."""
    m = small_mat(); m["dg"]["params"]["full2D_potentialDiscretization"] = mode
    p = F.init_params_full2d(m, m["V"], 0.2, 0.1)
    dr = F.DriftFull2D(m, p, m["V"])
    G = dr.assemble()
    assert np.allclose(dr.diagonal(), G.diagonal(), rtol=0, atol=1e-12 * abs(G).max())
    rs = np.asarray(abs(G).sum(axis=1)).ravel()
    bound = dr.row_abs()
    assert np.all(bound >= rs - 1e-12 * rs.max())
    if mode == "collocated":
        assert np.allclose(bound, rs, rtol=1e-13, atol=0)
    else:
        assert np.max(bound - rs) > 0   # the triangle inequality is strict somewhere


def _line_setup(mode="fv-consistent", **over):
    m = small_mat()
    m["dg"]["params"].update({"full2D_potentialDiscretization": mode, **over})
    p = F.init_params_full2d(m, m["V"], 0.2, 0.1)
    b = F.boundary_full2d(m, p, 0.2, 0.1, m["V"])
    parts, _, info = F.get_diff_full2d(m, p, b)
    drift = F.DriftFull2D(m, p, m["V"])
    return m, p, parts, info, drift, F.LineTransverse(m, p, m["V"], info, drift)


@pytest.mark.parametrize("mode", ["fv-consistent", "collocated"])
def test_line_transverse_inverts_its_factored_operator(mode):
    """apply(r) = M^-1 r with M = (sigma I + Xref)(sigma I + Tref)/sigma
    (get_LineTransversePreconditioner_full2D.m:4-6) assembled independently:
    Xref per mode = s_k H1 + |s_k| H2 (the DG upwind X operator of speed
    s_k = Q_x lambda_k), Tref per mode = K + w_k V + diag(capY + capX_k)."""
    m, p, parts, info, drift, lt = _line_setup(mode)
    nrx, nry, nxd, nyd = lt.sz
    ops = F.one_dimensional_dg_operators(p["dg"]["X"])
    H1, H2 = ops["H1"].toarray(), ops["H2"].toarray() if "H2" in ops else None
    if H2 is None:
        H2 = (ops["H2Interior"] + ops["H2Boundary"]).toarray()
    N = p["nTotal"]
    nT = lt.nT
    # mode-space index (k fastest, rho_y, X, Y)
    Xm = np.zeros((N, N), complex)
    Tm = np.zeros((N, N), complex)
    idx = np.arange(N).reshape(nyd, nxd, nry, nrx)
    for k in range(nrx):
        s = lt.Qx * lt.lam[k]
        Hk = s * H1 + abs(s) * H2
        Bt = lt.K + lt.weight[k] * lt.V + np.diag(lt.capY + lt.capX[k])
        for iyd in range(nyd):
            for ry in range(nry):
                rows = idx[iyd, :, ry, k]
                Xm[np.ix_(rows, rows)] += Hk
        for ixd in range(nxd):
            tt = idx[:, ixd, :, k].reshape(-1)        # (Y, rho_y) -> t = rho_y + nry*Y
            Tm[np.ix_(tt, tt)] += Bt
    sig = lt.sigma
    Mm = (sig * np.eye(N) + Xm) @ (sig * np.eye(N) + Tm) / sig
    G = np.kron(np.eye(N // nrx), lt.U)
    M = G @ Mm @ G.conj().T
    r = np.sin(np.arange(N) * 0.7) + 1j * np.cos(np.arange(N) * 0.3)
    z = lt.apply(M @ r)
    assert np.max(np.abs(z - r)) < 1e-9 * np.max(np.abs(r))
    assert lt.sigma > 0 and lt.factor_bytes > 0


def test_line_transverse_reference_block_is_the_y_transport():
    """K (transverseReference) equals the Y-part of the full operator at rho_x
    cell 1, X DOF 1: the Y parts do not depend on the X strip used to build it."""
    m, p, parts, info, drift, lt = _line_setup("fv-consistent")
    Ay = F.assemble_parts([q for q in parts if q[0].startswith("Y")], p)
    nrx, nry, nxd, nyd = lt.sz
    ids = np.array([nrx * (ry + nry * (nxd * iyd)) for iyd in range(nyd) for ry in range(nry)])
    assert np.allclose(Ay[ids][:, ids].toarray(), lt.K, rtol=0, atol=1e-12 * np.abs(lt.K).max())
    # the rho_x diagonal of P_x^T P_x / 4 on two cells is 1/2 (hence V is doubled)
    Px = sp.diags([np.ones(2), np.ones(2)], [0, -1], shape=(3, 2)).toarray()
    assert np.allclose(np.diag(Px.T @ Px / 4), 0.5)


def test_line_transverse_memory_guard():
    with pytest.raises(MemoryError, match="LineMemoryLimit"):
        _line_setup("fv-consistent", full2D_lineMaxFactorBytes=1000)


def periodic_mat(flux="rusanov", mode="fv-consistent"):
    """makeSmallMat with periodic Y: V periodic in y (period W), L_rho_y = 2W."""
    m = small_mat()
    X, Y = np.meshgrid(m["x"], m["y"], indexing="ij")
    W = m["y"][-1] - m["y"][0]
    m["V"] = 0.1 * X + 0.03 * np.cos(2 * np.pi * (Y - m["y"][0]) / W)
    m["dg"]["params"].pop("L_rho_y")
    m["dg"]["params"].update({"full2D_Y_boundary": "periodic", "N_rho_y": 6, "N_el_Y": 2,
                              "full2D_X_flux": flux, "full2D_Y_flux": flux,
                              "full2D_potentialDiscretization": mode})
    return m


@pytest.mark.parametrize("flux", ["rusanov", "upwind"])
@pytest.mark.parametrize("mode", ["fv-consistent", "collocated"])
def test_periodic_y_commutes_with_half_period_shift(flux, mode):
    """Physical states satisfy rho(Y+W/2, rho_y+W) = rho(Y, rho_y)
    (get_PeriodicYStateError_full2D.m); the periodic operator (transport and
    potential) commutes with that shift, and a Y-constant state has no
    Y-transport residual anywhere."""
    m = periodic_mat(flux, mode)
    p, b, parts, rparts, info, A, rhs = diff_setup(m)
    assert p["relative"]["rhoY"]["length"] == pytest.approx(2 * (m["y"][-1] - m["y"][0]))
    assert not any("Y physical" in q[0] or "specular" in q[0] for q in parts)
    G = F.DriftFull2D(m, p, m["V"]).assemble()
    u = det_vec(p["nTotal"])
    for Op in (A, G):
        lhs = Op @ F.periodic_y_shift(u, p)
        rhs_ = F.periodic_y_shift(Op @ u, p)
        assert np.max(np.abs(lhs - rhs_)) <= 1e-12 * np.max(np.abs(lhs))
    assert np.allclose(F.apply_parts(parts, u, p), A @ u, rtol=0, atol=1e-10 * abs(A).max())
    Ay = F.assemble_parts([q for q in parts if q[0].startswith("Y")], p)
    nr, nc = p["relative"]["nDof"], p["dg"]["nCenterDof"]
    f = det_vec(nr)
    assert np.max(np.abs(Ay @ np.tile(f, nc))) <= 1e-12 * abs(Ay).max() * np.max(np.abs(f))
    assert np.all(F.cap_full2d(p).reshape(p["relative"]["NrhoY"], -1)[:, 0] ==
                  F.cap_full2d(p).reshape(p["relative"]["NrhoY"], -1)[0, 0])   # no rho_y CAP


def test_periodic_y_grid_checks():
    m = periodic_mat(); m["dg"]["params"]["N_rho_y"] = 5
    with pytest.raises(ValueError, match="PeriodicGrid"):
        F.init_params_full2d(m)
    m = periodic_mat(); m["dg"]["params"]["L_rho_y"] = 0.7
    with pytest.raises(ValueError, match="PeriodicRelativeLength"):
        F.init_params_full2d(m)
