"""Results files in the layout of the reference's Sim_Res.mat (MAT v5):
the written variables, field names and array shapes equal those of the
stored file (its inventory is part of tests/golden/sim_res_slices.npz), the
round trip is exact, and a loaded device description builds the same
full-2D setup as the original."""
import json

import numpy as np
import scipy.io as sio

from real_space_dg_amd import full2d as P
from real_space_dg_amd.sim_res import HISTORY_FIELDS, load_sim_res, save_sim_res

import _sim_res as S


def inventory():
    d = np.load(S.FIXTURE, allow_pickle=False)
    return json.loads(str(d["inventory_json"]))


def sweep_result(cases):
    """A two-bias result in the form solve2d_rtd_self_consistent returns."""
    sw = {k: [] for k in ("bias", "current", "density", "potential", "phi", "flag", "relres", "converged",
                          "iterations", "status")}
    hist = []
    for i, c in enumerate(cases[:2]):
        for k, v in (("bias", c["bias"]), ("current", c["j"]), ("density", c["n"]), ("potential", c["V"]),
                     ("phi", c["phi"]), ("flag", 0), ("relres", c["relres"]), ("converged", True),
                     ("iterations", 3 + i), ("status", "fixed-point-converged")):
            sw[k].append(v)
        hist.append([{"iteration": it + 1, "transportFlag": 0, "transportRelres": 1e-6 * (it + 1),
                      "warmStartUsed": it > 0, "initialTrueRelres": 1.0 / (it + 1), "gmresIterations": 500 - it,
                      "densityChange": np.inf if it == 0 else 1e-3 / it, "potentialDefect": 1e-2 / (it + 1),
                      "poissonConverged": True, "poissonPhysicalRelres": 0.5 / (it + 1)} for it in range(3 + i)])
    return {"sweep": sw, "histories": hist, "converged": True}


def test_layout_matches_the_stored_sweep_file(tmp_path):
    mat, cases = S.load()
    inv = inventory()
    f = tmp_path / "Sim_Res.mat"
    save_sim_res(f, sweep_result(cases), mat, requested_bias=[c["bias"] for c in cases[:2]],
                 boundary={"order": np.array(["X-left", "X-right", "Y-bottom", "Y-top"], dtype=object),
                           "yBoundaryType": "periodic"})
    raw = sio.loadmat(f, squeeze_me=False, struct_as_record=True)
    assert list(raw["sweep"].dtype.names) == inv["sweep"]
    # selfConsistency: the stored file's fields in its order (setup details may differ)
    names = list(raw["selfConsistency"].dtype.names)
    assert [n for n in inv["selfConsistency"] if n in names] == names
    assert {"converged", "status", "history", "iterations"} <= set(names)
    nb = 2
    for k, shape in inv["sweep_shapes"].items():
        got = list(np.shape(raw["sweep"][k][0, 0]))
        assert got == shape[:-1] + [nb], (k, got, shape)
    cell = raw["selfConsistency"]["history"][0, 0]
    assert list(cell.shape) == [1, nb]
    rec = cell[0, 0]
    assert list(rec.dtype.names) == inv["history_fields"] == list(HISTORY_FIELDS)
    assert rec.shape == (3, 1) and inv["history_first_shape"][1] == 1
    # the device description keeps every variable the solver reads
    assert {"x", "y", "Vc", "me_x", "Nd", "eps", "Nx", "Ny", "dg"} <= set(raw["mat"].dtype.names)
    assert set(raw["mat"].dtype.names) <= set(inv["mat"])


def test_round_trip_and_loaded_description_builds_the_same_setup(tmp_path):
    mat, cases = S.load()
    res = sweep_result(cases)
    f = tmp_path / "Sim_Res.mat"
    save_sim_res(f, res, mat)
    d = load_sim_res(f)
    sw = d["sweep"]
    assert np.array_equal(sw["density"][:, :, 1], cases[1]["n"])
    assert np.array_equal(sw["current"][:, 0], np.real(cases[0]["j"]))
    assert np.array_equal(sw["potential"][:, :, 0], cases[0]["V"])
    assert np.array_equal(sw["phi"][:, :, 1], cases[1]["phi"])
    assert np.array_equal(sw["bias"].ravel(), [cases[0]["bias"], cases[1]["bias"]])
    assert sw["status"] == ["fixed-point-converged"] * 2
    hist = d["selfConsistency"]["history"]
    assert [len(h) for h in hist] == [3, 4]
    assert hist[1][0]["densityChange"] == np.inf and hist[1][2]["gmresIterations"] == 498
    assert hist[0][1]["warmStartUsed"] == 1.0
    assert d["selfConsistency"]["converged"] == 1.0
    m2 = d["mat"]
    c = cases[0]
    s1 = P.Full2DSetup(mat, c["V"], c["EfL"], c["EfR"])
    s2 = P.Full2DSetup(m2, c["V"], c["EfL"], c["EfR"])
    assert s1.n_total == s2.n_total and s1.periodic == s2.periodic
    assert np.array_equal(s1.X.nodes, s2.X.nodes) and np.array_equal(s1.Y.nodes, s2.Y.nodes)
    assert same(s1.boundary, s2.boundary)


def same(a, b):
    """Recursive equality of nested dicts / sequences / arrays."""
    if isinstance(a, dict):
        return isinstance(b, dict) and a.keys() == b.keys() and all(same(a[k], b[k]) for k in a)
    if isinstance(a, (list, tuple)):
        return len(a) == len(b) and all(same(x, y) for x, y in zip(a, b))
    return np.array_equal(np.asarray(a), np.asarray(b))


def test_unconverged_sweep_is_marked_and_partial(tmp_path):
    mat, cases = S.load()
    res = sweep_result(cases)
    res["sweep"]["converged"][1] = False
    res["sweep"]["status"][1] = "iteration-limit"
    res["converged"] = False
    f = tmp_path / "partial.mat"
    save_sim_res(f, res, mat, requested_bias=[0.0, 0.2, 0.6])
    d = load_sim_res(f)
    assert d["sweep"]["requestedBias"].shape == (1, 3) and d["sweep"]["bias"].shape == (1, 2)
    assert d["selfConsistency"]["converged"] == 0.0 and d["selfConsistency"]["status"] == "iteration-limit"
    assert np.array_equal(d["sweep"]["converged"], [[1.0, 0.0]])


def test_single_bias_and_single_iteration_keep_list_shapes(tmp_path):
    mat, cases = S.load()
    for n_it in (1, 2):
        res = sweep_result(cases[:1])
        res["histories"] = [res["histories"][0][:n_it]]
        f = tmp_path / f"one_{n_it}.mat"
        save_sim_res(f, res, mat)
        d = load_sim_res(f)
        assert d["sweep"]["status"] == ["fixed-point-converged"]
        hist = d["selfConsistency"]["history"]
        assert len(hist) == 1 and len(hist[0]) == n_it and isinstance(hist[0][0], dict)
        assert hist[0][-1]["iteration"] == n_it
