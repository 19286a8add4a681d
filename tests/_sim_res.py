"""The reference's stored full-2D RTD sweep slices (tests/golden/sim_res_slices.npz,
written by tests/golden/make_sim_res_slices.py) as a material description."""
import json
from pathlib import Path

import numpy as np

FIXTURE = Path(__file__).resolve().parent / "golden" / "sim_res_slices.npz"


def load():
    d = np.load(FIXTURE, allow_pickle=False)
    params = json.loads(str(d["params_json"]))
    for k in ("full2D_sourceContactMask", "full2D_drainContactMask"):
        params[k] = np.asarray(params[k])
    mat = dict(json.loads(str(d["scalars_json"])))
    mat.update({k: np.asarray(d[k]) for k in ("x", "y", "Vc", "me_x", "me_y", "me_z", "Nd", "eps")})
    mat["dg"] = {"params": params}
    cases = []
    for i in range(d["bias"].size):
        bias = float(d["bias"][i])
        # RTD/solve2D_rtd_self_consistent.m:106-107
        efl, efr = mat["Vc"][0, 0] + 0.05, mat["Vc"][-1, 0] + 0.05 - bias
        cases.append({"bias": bias, "V": np.asarray(d["potential"][:, :, i]), "EfL": efl, "EfR": efr,
                      "n": np.asarray(d["density"][:, :, i]), "j": np.asarray(d["current"][:, i]),
                      "relres": float(d["relres"][i]), "phi": np.asarray(d["phi"][:, :, i]),
                      "last_potential_defect": float(d["last_potential_defect"][i]),
                      "last_poisson_relres": float(d["last_poisson_relres"][i])})
    return mat, cases
