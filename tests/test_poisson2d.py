"""2D electrostatics of the full-2D RTD loop (host): the harmonic-mean
stencil annihilates the discrete constant-flux profile, the Dirichlet and
periodic rows, the direct step reproduces that profile, the Newton step
leaves a fixed point alone, and the DG -> dual-volume projection is exact
for constants and linear fields and conserves the integral."""
import numpy as np
import pytest

from real_space_dg_amd import full2d as P
from real_space_dg_amd import poisson2d as PS

from test_contact_modes import periodic_contact_mat


def slab(nx=9, ny=9, eps=None, periodic=True):
    m = periodic_contact_mat()
    m["dg"]["params"]["full2D_reservoirModel"] = "zero"
    if not periodic:
        m["dg"]["params"]["full2D_Y_boundary"] = "zero-inflow"
    m["eps"] = np.full((nx, ny), 12.9) if eps is None else eps
    m["Nd"] = np.full((nx, ny), 1e24)
    m["Temp"] = 300.0
    return m


def flux_profile(eps_x, dx, bias):
    """phi with a constant discrete flux c_{i+1/2}(phi_{i+1}-phi_i), c the
    harmonic mean, from phi(0) = 0 to phi(L) = bias."""
    c = 2 * eps_x[:-1] * eps_x[1:] / (eps_x[:-1] + eps_x[1:])
    steps = 1.0 / c
    return np.concatenate(([0.0], np.cumsum(steps))) * bias / steps.sum()


def test_stencil_annihilates_the_constant_flux_profile():
    ex = np.array([12.9, 12.9, 12.3, 12.3, 12.3, 12.9, 12.9, 11.0, 12.9])
    m = slab(eps=np.repeat(ex[:, None], 9, axis=1))
    M, types, periodic = PS.eval_poisson2d(m)
    assert periodic and np.all(types[[0, -1], :-1] == 2) and np.all(types[:, -1] == 3)
    phi = np.repeat(flux_profile(ex, 1.0, 0.3)[:, None], 9, axis=1)
    r = (M @ phi.ravel(order="F")).reshape(9, 9, order="F")
    free = types == 1
    assert np.max(np.abs(r[free])) <= 1e-12 * np.max(np.abs(M.data)) * 0.3
    # Dirichlet rows are the identity, periodic rows tie the last column to the first
    assert np.allclose(r[[0, -1], :-1], phi[[0, -1], :-1])
    assert np.allclose(r[:, -1], 0)


def test_direct_step_reproduces_the_neutral_profile_and_newton_keeps_it():
    ex = np.array([12.9, 12.9, 12.3, 12.3, 12.3, 12.9, 12.9, 11.0, 12.9])
    m = slab(eps=np.repeat(ex[:, None], 9, axis=1))
    M, types, _ = PS.eval_poisson2d(m)
    exact = np.repeat(flux_profile(ex, 1.0, 0.3)[:, None], 9, axis=1)
    n = m["Nd"].copy()                      # neutral: zero charge
    phi, info = PS.poisson_step(M, types, n, m, np.zeros((9, 9)), 0.3, "direct")
    assert info["converged"] and np.max(np.abs(phi - exact)) < 1e-12
    phi2, info2 = PS.poisson_step(M, types, n, m, exact, 0.3, "newton-raphson")
    assert info2["physicalRelresBefore"] < 1e-12 and np.max(np.abs(phi2 - exact)) < 1e-12


def test_projection_is_exact_for_linear_fields_and_conservative():
    m = slab()
    s = P.Full2DSetup(m, m["V"], 0.3, 0.3)
    X, Y = np.meshgrid(s.X.nodes, s.Y.nodes, indexing="ij")
    for f, periodic in ((lambda x, y: 3.0 + 0 * x, True), (lambda x, y: 2.0 + 1e8 * x + 0 * y, True),
                        (lambda x, y: 2.0 + 1e8 * x + 3e8 * y, False)):
        nDG = f(X, Y)
        d, info = PS.project_dg_density(nDG, s, periodic)
        xm, ym = np.asarray(m["x"]) * 1e-9, np.asarray(m["y"]) * 1e-9
        G = f(*np.meshgrid(xm, ym, indexing="ij"))
        inner = (slice(1, -1), slice(1, -1))          # symmetric dual volumes: the average is the node value
        assert np.allclose(d[inner], G[inner], rtol=1e-12)
        assert info["relativeConservationError"] < 1e-13
        if periodic:
            assert np.array_equal(d[:, -1], d[:, 0])
    rng = np.random.default_rng(0)
    d, info = PS.project_dg_density(rng.random((s.X.ndof, s.Y.ndof)), s, True)
    assert info["relativeConservationError"] < 1e-13


def test_doping_units_and_errors():
    m = slab()
    m["N_d"] = m["Nd"] / 1e6
    Nd, Na = PS.doping_si(m)
    assert np.allclose(Nd, 1e24) and np.all(Na == 0)
    m["N_d"] = m["Nd"]
    with pytest.raises(ValueError, match="InconsistentDopingUnits"):
        PS.doping_si(m)
    m = slab()
    m["eps"] = m["eps"][:, :-1]
    with pytest.raises(ValueError, match="MissingEps"):
        PS.eval_poisson2d(m)
