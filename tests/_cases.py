"""Shared small workloads and oracle builders for the parity tests."""
import numpy as np
import scipy.sparse as sp

from oracle import wigner1d as W
from real_space_dg_amd.devices import rtd_1d
from real_space_dg_amd.solver import WignerOperator

M0 = 9.109e-31


def build_case(n_elem, degree, nk, bias=0.0, fermi_model="DG_MSA_3D", **kw):
    mat, E, Vs, efl, efr = rtd_1d(n_elem, degree, nk, bias_v=bias, fermi_model=fermi_model, **kw)
    m = mat["me_x"][0, :, 0] * M0
    p = W.init_params(mat, efl, efr, E[0, :, 0], m, 1.0)
    return dict(mat=mat, E=E, Vs=Vs, efl=efl, efr=efr, m=m, p=p)


def oracle_system(c):
    A, rhs = W.get_diff(c["p"])
    G = W.get_drift(c["mat"], c["p"], c["E"][0, :, 0])
    return A.tocsr(), G.tocsr(), rhs


def gpu_operator(c):
    mat = c["mat"]
    op = WignerOperator.from_mat(mat)
    op.set_potential(mat["x"] * 1e-9, c["E"][0, :, 0], c["efl"], c["efr"], c["m"][0], c["m"][-1],
                     temp=mat["Temp"], fermi_model=mat["fermi_model"], me_z=mat["me_z_ch"])
    return op


def line_operator(A, G, n_chi, n, nk):
    """Transport + k-diagonal part of G: the operator the line preconditioner inverts."""
    Gc = G.tocoo()
    j_r = Gc.row % nk
    j_c = Gc.col % nk
    keep = j_r == j_c
    Gd = sp.csr_matrix((Gc.data[keep], (Gc.row[keep], Gc.col[keep])), shape=G.shape)
    return (A + Gd).tocsr()


def rand_c(n, seed=0):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)
