"""Batched solves (wdg_solve_batch) agree with one-by-one solves."""
import numpy as np
import pytest

from real_space_dg_amd.devices import rtd_sweep
from real_space_dg_amd.solver import SolveOptions, WignerOperator, solve_batch

pytestmark = pytest.mark.gpu


def _ops(cases):
    ops = []
    for mat, E, Vs, efl, efr in cases:
        op = WignerOperator.from_mat(mat)
        m = mat["me_x"][0, :, 0] * 9.109e-31
        op.set_potential(mat["x"] * 1e-9, E[0, :, 0], efl, efr, m[0], m[-1])
        ops.append(op)
    return ops


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_batch_matches_individual(prec):
    cases = rtd_sweep(6, seed=3, n_elems=(24, 32, 48), nks=(32, 64))
    print("\n[batch] 6 uneven instances", [(c[0]["Nx"], c[0]["dg"]["params"]["N"]) for c in cases], flush=True)
    ops = _ops(cases)
    opts = SolveOptions(tol=1e-10, precision=prec)
    infos = solve_batch(ops, opts)
    m_nodes = [np.full(op.n_chi * op.n_k_chi, 0.067 * 9.109e-31) for op in ops]
    nb = [op.observables(mn)[0] for op, mn in zip(ops, m_nodes)]
    for op, info, mn, n_batch in zip(ops, infos, m_nodes, nb):
        assert info["flag"] == 0 and info["true_relres"] <= 1e-10
        x, i1 = op.solve(opts)
        n1 = op.observables(mn)[0]
        assert i1["iterations"] == info["iterations"]
        assert np.array_equal(n1, n_batch)
    for op in ops:
        op.close()
