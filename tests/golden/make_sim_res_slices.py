"""Writes tests/golden/sim_res_slices.npz from the reference's stored
full-2D RTD sweep (2026-10-28_08-47/Sim_Res.mat, MAT v5 data read with
scipy.io.loadmat): the material description and dg.params of the run, and
for three biases the potential V = Vc - phi handed to the transport solve,
the converged density n (X DOF x material y) and current j (X DOF), the
electrostatic potential phi, the solver's relative residual and the last
self-consistency iteration's potential defect and Poisson residual; and the
file's layout (variables, field names, sweep shapes, history fields).  Run from the repository root, in a
container where the reference tree is present:

    python tests/golden/make_sim_res_slices.py
"""
import json
from pathlib import Path

import numpy as np
import scipy.io as sio

SRC = Path("/root/reference/2026-10-28_08-47/Sim_Res.mat")
OUT = Path(__file__).resolve().parent / "sim_res_slices.npz"
BIASES = (0, 10, 30)     # 0.0, 0.2 and 0.6 V


def plain(v):
    if isinstance(v, np.ndarray):
        return v.tolist()
    if isinstance(v, (np.floating, np.integer)):
        return v.item()
    return v


def main():
    d = sio.loadmat(SRC, squeeze_me=True, struct_as_record=False)
    mat, sweep = d["mat"], d["sweep"]
    params = {f: plain(getattr(mat.dg.params, f)) for f in mat.dg.params._fieldnames}
    scal = {f: plain(getattr(mat, f)) for f in ("Nx", "Ny", "Temp", "deg_factor", "me_x_ch", "me_y_ch", "me_z_ch",
                                                "dx", "dy", "Lc1", "Lc2")}
    arrays = {"x": np.asarray(mat.x, float), "y": np.asarray(mat.y, float), "Vc": np.asarray(mat.Vc, float),
              "me_x": np.asarray(mat.me_x, float), "me_y": np.asarray(mat.me_y, float),
              "me_z": np.asarray(mat.me_z, float), "Nd": np.asarray(mat.Nd, float),
              "eps": np.asarray(mat.eps, float)}
    ib = np.array(BIASES)
    # last self-consistency iteration of each bias: unmixed potential defect
    # [V] and the Poisson physical residual at the stored state
    hist = d["selfConsistency"].history
    last = [np.atleast_1d(hist[i])[-1] for i in BIASES]
    arrays["last_potential_defect"] = np.array([float(h.potentialDefect) for h in last])
    arrays["last_poisson_relres"] = np.array([float(h.poissonPhysicalRelres) for h in last])
    # layout of the file itself: top-level variables, their fields (in the
    # stored order), the sweep arrays' shapes and the history record fields
    raw = sio.loadmat(SRC, squeeze_me=False, struct_as_record=True)
    inventory = {k: list(raw[k].dtype.names) for k in ("sweep", "selfConsistency", "mat", "boundary")}
    inventory["sweep_shapes"] = {f: list(np.shape(raw["sweep"][f][0, 0])) for f in raw["sweep"].dtype.names}
    inventory["history_fields"] = list(raw["selfConsistency"]["history"][0, 0][0, 0].dtype.names)
    inventory["history_cell_shape"] = list(raw["selfConsistency"]["history"][0, 0].shape)
    inventory["history_first_shape"] = list(raw["selfConsistency"]["history"][0, 0][0, 0].shape)
    np.savez_compressed(
        OUT, params_json=np.array(json.dumps(params)), scalars_json=np.array(json.dumps(scal)), **arrays,
        inventory_json=np.array(json.dumps(inventory)),
        bias_index=ib, bias=np.asarray(sweep.bias, float)[ib],
        potential=np.asarray(sweep.potential, float)[:, :, ib], phi=np.asarray(sweep.phi, float)[:, :, ib],
        density=np.asarray(sweep.density, float)[:, :, ib],
        current=np.asarray(sweep.current, float)[:, ib], relres=np.asarray(sweep.relres, float)[ib],
        flag=np.asarray(sweep.flag, float)[ib])
    print("wrote", OUT, OUT.stat().st_size, "bytes")


if __name__ == "__main__":
    main()
