"""Host-side setup of the full-2D device operator (real_space_dg_amd.full2d)
against the FP64 oracle, on the CPU: the structured parts handed to the
device, expanded here to sparse matrices, must assemble to the oracle's
operator; the RHS, quadrature factors and observable weights must agree."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import full2d as F
from real_space_dg_amd import _lib
from real_space_dg_amd import full2d as P

from test_oracle_full2d import periodic_mat, small_mat


def configs():
    base = small_mat()
    out = {"rusanov-specular-fv": (base, 0.2, 0.1)}
    m = small_mat()
    m["dg"]["params"].update({"full2D_X_flux": "upwind", "full2D_Y_flux": "upwind", "full2D_Y_boundary": "zero-inflow"})
    out["upwind-zero-inflow"] = (m, 0.2, 0.1)
    m = small_mat()
    m["dg"]["params"].update({"full2D_fluxType": "central", "full2D_potentialDiscretization": "collocated"})
    nd = np.ones((5, 5)); nd[0, :2] = 0; nd[-1, 3:] = 0
    m["Nd"] = nd
    out["central-closed-faces-collocated"] = (m, 0.25, 0.05)
    m = small_mat()
    m["dg"]["params"].update({"N_K_X": 2, "N_K_Y": 2, "N_rho_x": 6, "N_rho_y": 4, "full2D_potentialQuadratureOrder": 5})
    out["p1-even-rho"] = (m, 0.2, 0.15)
    out["periodic-rusanov-fv"] = (periodic_mat("rusanov", "fv-consistent"), 0.2, 0.1)
    out["periodic-upwind-collocated"] = (periodic_mat("upwind", "collocated"), 0.22, 0.12)
    return out


CONFIGS = configs()


def _axis_matrix(part, nl, ne):
    nd = nl * ne
    H = np.zeros((nd, nd))
    if part.center == _lib.WDG2_CENTER_BANDED:
        for e in range(ne):
            r = slice(e * nl, (e + 1) * nl)
            H[r, r] = part.blocks.diag[e]
            H[r, (e * nl - 1) % nd] += part.blocks.lower[e]       # zero at the ends unless periodic
            H[r, ((e + 1) * nl) % nd] += part.blocks.upper[e]
    else:
        rows = slice(0, nl) if part.face == 0 else slice(nd - nl, nd)
        H[rows, 0 if part.face == 0 else nd - 1] = part.lift
    return H


def expand(setup):
    """sum coef kron(C, R) of the device parts, on the CPU."""
    X, Y, RX, RY = setup.X, setup.Y, setup.RX, setup.RY
    A = sp.csr_matrix((setup.n_total, setup.n_total), dtype=complex)
    for part in setup.parts:
        ax = part.axis
        nl, ne = (X.nl, X.ne) if ax == 0 else (Y.nl, Y.ne)
        H = sp.csr_matrix(_axis_matrix(part, nl, ne))
        ntr = Y.ndof if ax == 0 else X.ndof
        T = sp.diags(np.ones(ntr) if part.mask is None else np.asarray(part.mask, float))
        C = sp.kron(T, H) if ax == 0 else sp.kron(H, T)
        R1 = RX if ax == 0 else RY
        if part.rel == _lib.WDG2_REL_CENTRAL:
            n = R1.n
            D = sp.diags([-np.ones(n - 1), np.ones(n - 1)], [-1, 1]).tolil()
            if ax == 1 and setup.periodic:
                D[0, n - 1] = -1
                D[n - 1, 0] = 1
            base = -1j * sp.csr_matrix(D) / (2 * R1.delta)
        elif part.rel == _lib.WDG2_REL_IDENTITY:
            base = sp.identity(R1.n)
        else:
            base = sp.csr_matrix(part.dense)
        R = sp.kron(sp.identity(RY.n), base) if ax == 0 else sp.kron(base, sp.identity(RX.n))
        A = A + part.coef * sp.kron(C, R, format="csr")
    return A.tocsr()


@pytest.mark.parametrize("name", list(CONFIGS))
def test_parts_assemble_to_the_oracle_operator(name):
    m, EfL, EfR = CONFIGS[name]
    s = P.Full2DSetup(m, m["V"], EfL, EfR)
    p = F.init_params_full2d(m, m["V"], EfL, EfR)
    b = F.boundary_full2d(m, p, EfL, EfR, m["V"])
    parts, rparts, _ = F.get_diff_full2d(m, p, b)
    A_or = F.assemble_parts(parts, p)
    A_pr = expand(s)
    scale = abs(A_or).max()
    assert abs(A_pr - A_or).max() <= 1e-12 * scale
    rhs_or = F.assemble_rhs(rparts, p)
    assert np.max(np.abs(s.rhs() - rhs_or)) <= 1e-12 * max(np.max(np.abs(rhs_or)), 1e-300)
    assert np.max(np.abs(rhs_or)) > 0


@pytest.mark.parametrize("name", list(CONFIGS))
def test_potential_factors_match_the_oracle(name):
    m, EfL, EfR = CONFIGS[name]
    s = P.Full2DSetup(m, m["V"], EfL, EfR)
    p = F.init_params_full2d(m, m["V"], EfL, EfR)
    d = F.DriftFull2D(m, p, m["V"])
    assert np.allclose(s.cap_x, F.W.get_cap(p["relative"]["rhoX"]["cells"], p["relative"]["rhoX"]["length"],
                                            p["relative"]["rhoX"]["delta"]), rtol=0, atol=1e-14)
    assert s.pot_coef == pytest.approx(d.pot_coef.imag) and s.cap_coef == pytest.approx(d.cap_coef)
    if s.pot_mode == "fv-consistent":
        assert np.allclose(np.kron(s.interp[1], s.interp[0]), d.interpolation, rtol=0, atol=1e-13)
        assert np.allclose(np.kron(s.proj[1], s.proj[0]), d.centerProjection, rtol=0, atol=1e-12)
        nqx, nqy = s.nq
        for k, (ids, Xq, Yq) in enumerate(d.elements):
            ex, ey = k % s.X.ne, k // s.X.ne
            xq = s.quad[0][ex * nqx:(ex + 1) * nqx]
            yq = s.quad[1][ey * nqy:(ey + 1) * nqy]
            assert np.allclose(np.tile(xq, nqy), Xq, rtol=0, atol=1e-15)
            assert np.allclose(np.repeat(yq, nqx), Yq, rtol=0, atol=1e-15)
        assert np.allclose(s.rel_pts[0], d.vx[:s.RX.n + 1], rtol=0, atol=1e-15)


@pytest.mark.parametrize("name", list(CONFIGS))
def test_observable_weights_reproduce_the_oracle(name):
    m, EfL, EfR = CONFIGS[name]
    s = P.Full2DSetup(m, m["V"], EfL, EfR)
    p = F.init_params_full2d(m, m["V"], EfL, EfR)
    k = np.arange(1, s.n_total + 1)
    rho = np.sin(0.37 * k) + 1j * np.cos(0.11 * k)
    R = rho.reshape(s.ncen, s.nrel)
    wn, wx, wy = s.observable_weights()
    dens = sum(w * R[:, i] for i, w in wn).real.reshape(s.Y.ndof, s.X.ndof).T
    dx = sum(w * R[:, i] for i, w in wx).reshape(s.Y.ndof, s.X.ndof).T
    dy = sum(w * R[:, i] for i, w in wy).reshape(s.Y.ndof, s.X.ndof).T
    cx, cy = s.current_scales()
    n_or, jx_or, jy_or = F.observables_full2d(rho, p, m)
    assert np.allclose(dens, n_or, rtol=1e-13, atol=0)
    assert np.allclose(cx * dx.imag, jx_or, rtol=1e-12, atol=1e-12 * np.abs(jx_or).max())
    assert np.allclose(cy * dy.imag, jy_or, rtol=1e-12, atol=1e-12 * np.abs(jy_or).max())


def test_closed_form_characteristic_split_matches_eig():
    """The product splits -i D by its closed-form eigenpairs; the oracle uses eig."""
    for n in (4, 5, 12, 33):
        delta = 0.37e-9
        Ap, Am, lam = P.characteristic_split(n, delta)
        D = sp.diags([-np.ones(n - 1), np.ones(n - 1)], [-1, 1]) / (2 * delta)
        Op, Om, lo = F.split_characteristic(-1j * D)
        s = 1 / delta
        assert np.allclose(np.sort(lam), np.sort(lo.real), rtol=0, atol=1e-12 * s)
        assert np.allclose(Ap, Op, rtol=0, atol=1e-12 * s) and np.allclose(Am, Om, rtol=0, atol=1e-12 * s)


def test_unsupported_options_raise():
    m = small_mat(); m["dg"]["params"]["full2D_Y_boundary"] = "hard-wall"
    with pytest.raises(NotImplementedError, match="Unsupported"):
        P.Full2DSetup(m, m["V"], 0.2, 0.1)
    m = periodic_mat(); m["V"] = m["V"] + 0.01 * np.arange(5)[None, :]
    with pytest.raises(ValueError, match="NonperiodicPotential"):
        P.Full2DSetup(m, m["V"], 0.2, 0.1)
    m = small_mat(); m["dg"]["params"]["full2D_massModel"] = "position-dependent"
    with pytest.raises(NotImplementedError, match="Unsupported"):
        P.Full2DSetup(m, m["V"], 0.2, 0.1)
