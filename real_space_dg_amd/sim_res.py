"""MAT v5 results files in the layout of the reference's ``Sim_Res.mat``.

The RTD runner saves ``save(..., 'Sim_Res.mat', '-struct', 'results', 'sweep',
'selfConsistency', 'mat', 'boundary')`` (RTD/main2D_rtd.m:55): four top-level
variables, of which ``sweep`` gathers one column per bias
(RTD/solve2D_rtd_self_consistent.m:188-199) and ``selfConsistency`` the
per-bias iteration histories (:216-222).  ``save_sim_res`` writes the result
of ``poisson2d.solve2d_rtd_self_consistent`` in that layout, so MATLAB
post-processing that reads ``Sim_Res.mat`` reads these files unchanged;
``load_sim_res`` reads such a file (the reference's own or one written here)
back into plain Python containers, with ``mat`` in the dict form that
``Full2DSetup`` and ``solve2d_rtd_self_consistent`` take.  Both go through
scipy.io's MAT v5 reader and writer, which execute nothing from the file.

Layout (nb biases, nX X DOF, Nx x Ny material grid, as the reference):
  sweep.requestedBias, bias, flag, relres, converged, iterations: 1 x nb
  sweep.current: nX x nb;  sweep.density: nX x Ny x nb
  sweep.potential, sweep.phi: Nx x Ny x nb;  sweep.status: 1 x nb cell of char
  selfConsistency.history: 1 x nb cell of iterations x 1 struct arrays with
  the fields of RTD/solve2D_rtd_self_consistent.m:115-170
"""
from __future__ import annotations

import numpy as np
import scipy.io as sio

SWEEP_FIELDS = ("requestedBias", "bias", "current", "density", "potential", "phi", "flag", "relres", "converged",
                "iterations", "status")
HISTORY_FIELDS = ("iteration", "transportFlag", "transportRelres", "warmStartUsed", "initialTrueRelres",
                  "gmresIterations", "densityChange", "potentialDefect", "poissonConverged", "poissonPhysicalRelres")
NOTE = ("Built-in: fully ionized electron-only Poisson; approximate Newton uses local Boltzmann screening. "
        "External callbacks remain caller responsibility.")


def _cell(items) -> np.ndarray:
    """1 x n MATLAB cell array."""
    out = np.empty((1, len(items)), dtype=object)
    for i, v in enumerate(items):
        out[0, i] = v
    return out


def _struct_array(rows: list[dict], fields) -> np.ndarray:
    """n x 1 MATLAB struct array with the given fields (missing values: [])."""
    out = np.empty((len(rows), 1), dtype=[(f, object) for f in fields])
    for i, r in enumerate(rows):
        for f in fields:
            v = r.get(f, np.zeros((0, 0)))
            out[f][i, 0] = float(v) if isinstance(v, (bool, np.bool_)) else v
    return out


def sweep_arrays(result: dict, requested_bias=None) -> dict:
    """The ``sweep`` struct of a self-consistent run as MATLAB-shaped arrays.

    Columns follow the biases the loop reached (it stops at the first
    unconverged bias, RTD/solve2D_rtd_self_consistent.m:200-202);
    ``requestedBias`` keeps the whole requested list."""
    sw = result["sweep"]
    nb = len(sw["bias"])
    row = lambda k: np.asarray(sw[k], float).reshape(1, nb)
    out = {
        "requestedBias": np.asarray(sw["bias"] if requested_bias is None else requested_bias, float).reshape(1, -1),
        "bias": row("bias"),
        "current": np.stack([np.real(np.asarray(v)) for v in sw["current"]], axis=-1),
        "density": np.stack([np.real(np.asarray(v)) for v in sw["density"]], axis=-1),
        "potential": np.stack([np.asarray(v, float) for v in sw["potential"]], axis=-1),
        "phi": np.stack([np.asarray(v, float) for v in sw["phi"]], axis=-1),
        "flag": row("flag"),
        "relres": row("relres"),
        "converged": np.asarray(sw["converged"], float).reshape(1, nb),
        "iterations": row("iterations"),
        "status": _cell([str(s) for s in sw["status"]]),
    }
    return out


def save_sim_res(path, result: dict, mat: dict, requested_bias=None, boundary: dict | None = None,
                 poisson_method: str = "direct") -> None:
    """Write ``result`` (``solve2d_rtd_self_consistent``) and the device
    description ``mat`` as a MAT v5 file with the variables of the
    reference's ``Sim_Res.mat``.  ``boundary`` (a dict) is written when given;
    the reference's full boundary-condition struct is not rebuilt here."""
    sw = sweep_arrays(result, requested_bias)
    hist = [_struct_array(h, HISTORY_FIELDS) for h in result["histories"]]
    status = sw["status"][0, -1] if sw["status"].size else "no-bias"
    sc = {"converged": float(bool(result["converged"])
                             and sw["bias"].size == sw["requestedBias"].size),
          "status": status, "history": _cell(hist), "iterations": sw["iterations"],
          "poissonImplementedHere": 1.0,
          "poissonSetup": {"yBoundary": "periodic", "method": poisson_method,
                           "approximateNewton": float(poisson_method != "direct")},
          "poissonCallback": "make_RTD_poisson_step/poissonStep", "note": NOTE}
    out = {"sweep": sw, "selfConsistency": sc, "mat": mat}
    if boundary is not None:
        out["boundary"] = boundary
    sio.savemat(path, out, format="5", long_field_names=True, do_compression=True, oned_as="row")


def load_sim_res(path) -> dict:
    """Read a ``Sim_Res.mat``-layout file: returns ``{"sweep", "selfConsistency",
    "mat", "boundary"}`` (those present) with structs as dicts, cells as
    lists and arrays in MATLAB's shapes (sweep columns = biases)."""
    d = sio.loadmat(path, squeeze_me=False, struct_as_record=False, chars_as_strings=True)
    out = {}
    for k in ("sweep", "selfConsistency", "mat", "boundary"):
        if k in d:
            out[k] = _unwrap(d[k])
    sc = out.get("selfConsistency")
    if isinstance(sc, dict) and "history" in sc:
        # one list of iteration records per bias, also for a single bias or iteration
        hist = sc["history"]
        hist = [hist] if isinstance(hist, dict) else list(hist)
        if hist and isinstance(hist[0], dict) and all(isinstance(h, dict) for h in hist) and \
                len(out.get("sweep", {}).get("status", [None])) == 1:
            hist = [hist]
        sc["history"] = [[h] if isinstance(h, dict) else list(h) for h in hist]
    sw = out.get("sweep")
    if isinstance(sw, dict) and isinstance(sw.get("status"), str):
        sw["status"] = [sw["status"]]
    return out


def _unwrap(v):
    """squeeze_me=False keeps every struct as a 1 x 1 array: unwrap those
    and scalars, keep the shapes of real arrays."""
    if isinstance(v, np.ndarray) and v.dtype == object:
        if v.size == 1 and isinstance(v.flat[0], sio.matlab.mat_struct):
            return _unwrap(v.flat[0])
        return [_unwrap(x) for x in v.ravel(order="F")]
    if isinstance(v, sio.matlab.mat_struct):
        return {f: _unwrap(getattr(v, f)) for f in v._fieldnames}
    if isinstance(v, np.ndarray) and v.dtype.kind == "U":
        return str(v.item()) if v.size == 1 else [str(x) for x in v.ravel(order="F")]
    if isinstance(v, np.ndarray) and v.size == 1 and v.ndim <= 2:
        return v.item()
    return v
