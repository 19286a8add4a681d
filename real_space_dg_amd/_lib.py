"""ctypes binding of libwdg.so (include/wigner_dg.h).

The library is built in-tree by ``real_space_dg_amd.build``.  There is no CPU
fallback: if the shared library is missing or cannot be loaded, every entry
point raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libwdg.so"

WDG_OK = 0
WDG_FERMI_MSA_3D = 0
WDG_FERMI_MSA_2D = 1
WDG_FERMI_TSU_ESAKI = 2
WDG_THETA_AUTO = 0
WDG_THETA_FFT = 1
WDG_THETA_GEMM = 2
WDG_FFT_BLOCK_AUTO = 0
WDG_FFT_BLOCK_256 = 1
WDG_FFT_BLOCK_WIDE = 2
WDG_FFT_BLOCK_NDFT = 3
WDG_PART_TRANSPORT = 1
WDG_PART_POTENTIAL = 2
WDG_PART_ALL = 3
WDG_BASIS_WORKING = 0
WDG_BASIS_BF16 = 1
WDG_BASIS_FP16 = 2
WDG_PREC_FP64 = 0
WDG_PREC_MIXED = 1
WDG2_AXIS_X = 0
WDG2_AXIS_Y = 1
WDG2_CENTER_BANDED = 0
WDG2_CENTER_FACE = 1
WDG2_REL_CENTRAL = 0
WDG2_REL_IDENTITY = 1
WDG2_REL_DENSE = 2
WDG2_POT_FV_CONSISTENT = 0
WDG2_POT_COLLOCATED = 1
WDG2_DV_INTERP = 0
WDG2_DV_TABLE = 1
WDG2_MODES_BOX = 0
WDG2_MODES_FOURIER = 1
WDG2_PRECOND_NONE = 0
WDG2_PRECOND_JACOBI = 1
WDG2_PRECOND_ROWABS = 2
WDG2_PRECOND_LINE = 3

WDG_ERR_ARG = -1
WDG_ERR_HIP = -2
WDG_ERR_MEMORY = -3
WDG_ERR_STATE = -4
WDG_ERR_NUMERIC = -5
WDG_ERR_UNSUPPORTED = -6
WDG_ERR_INITIAL_RHO = -7
WDG_ERR_GMRES_MEMORY = -8
WDG_ERR_LINE_MEMORY = -9

# Status code -> error identifier: the reference's own DG:Full2D:* identifiers
# for the three errors it raises (solve_transport_DG_full2D.m:197-202,519-525,
# get_LineTransversePreconditioner_full2D.m:36-43), wdg:* for this library's.
ERROR_IDS = {
    WDG_ERR_ARG: "wdg:InvalidArgument",
    WDG_ERR_HIP: "wdg:DeviceError",
    WDG_ERR_MEMORY: "wdg:OutOfMemory",
    WDG_ERR_STATE: "wdg:CallOrder",
    WDG_ERR_NUMERIC: "wdg:NumericalFailure",
    WDG_ERR_UNSUPPORTED: "wdg:Unsupported",
    WDG_ERR_INITIAL_RHO: "DG:Full2D:InvalidInitialRho",
    WDG_ERR_GMRES_MEMORY: "DG:Full2D:GmresMemoryLimit",
    WDG_ERR_LINE_MEMORY: "DG:Full2D:LineMemoryLimit",
}


class WdgError(RuntimeError):
    def __init__(self, code: int, msg: str):
        self.code = code
        self.identifier = ERROR_IDS.get(code, "wdg:Error")
        super().__init__(f"{self.identifier} ({code}): {msg}")


class wdg_desc(C.Structure):
    _fields_ = [("n_chi", C.c_int32), ("n_k_chi", C.c_int32), ("n_xi", C.c_int32), ("nk", C.c_int32),
                ("l_chi", C.c_double), ("l_xi", C.c_double)]


class wdg_potential(C.Structure):
    _fields_ = [("n_mat", C.c_int32), ("fermi_model", C.c_int32),
                ("x_mat", C.POINTER(C.c_double)), ("em", C.POINTER(C.c_double)),
                ("m_left", C.c_double), ("m_right", C.c_double), ("ef_left", C.c_double),
                ("ef_right", C.c_double), ("temp", C.c_double), ("deg_factor", C.c_double),
                ("me_z", C.c_double)]


class wdg_solve_opts(C.Structure):
    _fields_ = [("tol", C.c_double), ("restart", C.c_int32), ("max_iter", C.c_int32),
                ("precision", C.c_int32), ("inner_tol", C.c_double), ("max_outer", C.c_int32),
                ("use_precond", C.c_int32), ("use_graph", C.c_int32),
                ("orthogonalization", C.c_int32), ("restart_max", C.c_int32), ("warm_device", C.c_int32),
                ("krylov_basis", C.c_int32)]


class wdg_solve_info(C.Structure):
    _fields_ = [("flag", C.c_int32), ("iterations", C.c_int32), ("outer_iterations", C.c_int32),
                ("restarts", C.c_int32), ("true_relres", C.c_double), ("reported_relres", C.c_double),
                ("initial_true_relres", C.c_double), ("seconds", C.c_double),
                ("reorth_steps", C.c_int32)]


class wdg_transient_opts(C.Structure):
    _fields_ = [("dt", C.c_double), ("nsteps", C.c_int32), ("save_every", C.c_int32),
                ("a_scale", C.c_double), ("b_scale", C.c_double)]


class wdg_cn_opts(C.Structure):
    _fields_ = [("dt", C.c_double), ("nsteps", C.c_int32), ("a_scale", C.c_double), ("b_scale", C.c_double),
                ("tol", C.c_double), ("restart", C.c_int32), ("max_iter", C.c_int32), ("precision", C.c_int32),
                ("krylov_basis", C.c_int32), ("use_precond", C.c_int32), ("use_graph", C.c_int32)]


class wdg_cn_info(C.Structure):
    _fields_ = [("steps", C.c_int32), ("iterations", C.c_int32), ("max_step_iterations", C.c_int32),
                ("flag", C.c_int32), ("max_relres", C.c_double), ("seconds", C.c_double)]


class wdg_poisson_desc(C.Structure):
    _fields_ = [("n_mat", C.c_int32), ("method", C.c_int32), ("x_mat", C.POINTER(C.c_double)),
                ("eps_r", C.POINTER(C.c_double)), ("net_doping", C.POINTER(C.c_double)), ("temp", C.c_double)]


class wdg_poisson_info(C.Structure):
    _fields_ = [("physical_relres_before", C.c_double), ("unmixed_step", C.c_double),
                ("charge_integral", C.c_double), ("dg_integral", C.c_double)]


class wdg2_desc(C.Structure):
    _fields_ = [("n_rho_x", C.c_int32), ("n_rho_y", C.c_int32), ("nl_x", C.c_int32), ("nl_y", C.c_int32),
                ("ne_x", C.c_int32), ("ne_y", C.c_int32), ("delta_x", C.c_double), ("delta_y", C.c_double),
                ("periodic_y", C.c_int32)]


class wdg2_part(C.Structure):
    _fields_ = [("axis", C.c_int32), ("center", C.c_int32), ("rel", C.c_int32), ("face", C.c_int32),
                ("coef", C.c_double), ("c_diag", C.POINTER(C.c_double)), ("c_lower", C.POINTER(C.c_double)),
                ("c_upper", C.POINTER(C.c_double)), ("mask", C.POINTER(C.c_int32)),
                ("r_dense", C.POINTER(C.c_double))]


class wdg2_potential(C.Structure):
    _fields_ = [("mode", C.c_int32), ("nq_x", C.c_int32), ("nq_y", C.c_int32), ("nx_grid", C.c_int32),
                ("ny_grid", C.c_int32)] + \
               [(n, C.POINTER(C.c_double)) for n in ("x_grid", "y_grid", "v", "center_x", "center_y", "rel_x",
                                                     "rel_y", "cap_x", "cap_y", "interp_x", "interp_y", "proj_x",
                                                     "proj_y")] + \
               [("pot_coef", C.c_double), ("cap_coef", C.c_double)]


class wdg2_contact_modes(C.Structure):
    _fields_ = [("n_modes", C.c_int32), ("kind", C.c_int32), ("n_grid", C.c_int32),
                ("grid", C.POINTER(C.c_double)), ("values", C.POINTER(C.c_double)), ("y0", C.c_double),
                ("width", C.c_double), ("energy", C.POINTER(C.c_double)), ("mass_x", C.POINTER(C.c_double)),
                ("mass_z", C.POINTER(C.c_double)), ("fermi_level", C.c_double), ("mask", C.POINTER(C.c_int32)),
                ("char_matrix", C.POINTER(C.c_double)), ("lift", C.POINTER(C.c_double)),
                ("reservoir_out", C.POINTER(C.c_double))]


class wdg2_reservoir_grid(C.Structure):
    _fields_ = [("temp", C.c_double), ("deg_factor", C.c_double), ("n_kx", C.c_int32),
                ("kx", C.POINTER(C.c_double)), ("dkx", C.c_double), ("n_kz", C.c_int32),
                ("kz", C.POINTER(C.c_double)), ("rho_x", C.POINTER(C.c_double)), ("rho_y", C.POINTER(C.c_double)),
                ("y_nodes", C.POINTER(C.c_double)), ("scale_x", C.c_double)]


class wdg2_line_factors(C.Structure):
    _fields_ = [("sigma", C.c_double), ("modes", C.POINTER(C.c_double)), ("forward", C.POINTER(C.c_int32)),
                ("x_inverse", C.POINTER(C.c_double)), ("x_coupling", C.POINTER(C.c_double)),
                ("t_inverse", C.POINTER(C.c_double))]


class wdg2_poisson_desc(C.Structure):
    _fields_ = [("nx", C.c_int32), ("nu", C.c_int32), ("n_entries", C.c_int32),
                ("a_row", C.POINTER(C.c_int32)), ("a_col", C.POINTER(C.c_int32)), ("a_val", C.POINTER(C.c_double)),
                ("node_type", C.POINTER(C.c_int32)), ("net_doping", C.POINTER(C.c_double)), ("temp", C.c_double),
                ("x_dof", C.c_int32), ("y_dof", C.c_int32), ("proj_x", C.POINTER(C.c_double)),
                ("proj_y", C.POINTER(C.c_double)), ("volume", C.POINTER(C.c_double)), ("n_weights", C.c_int32),
                ("w_idx", C.POINTER(C.c_int32)), ("w_val", C.POINTER(C.c_double))]


class wdg2_poisson_info(C.Structure):
    _fields_ = [("physical_relres_before", C.c_double), ("physical_relres_after", C.c_double),
                ("linear_relres", C.c_double), ("unmixed_step", C.c_double)]


# name -> (restype, argtypes)
_P = C.c_void_p
_D = C.POINTER(C.c_double)
SIGNATURES = {
    "wdg_last_error": (C.c_char_p, [_P]),
    "wdg_create": (C.c_int, [C.POINTER(wdg_desc), C.POINTER(_P)]),
    "wdg_destroy": (C.c_int, [_P]),
    "wdg_size": (C.c_int64, [_P]),
    "wdg_set_potential": (C.c_int, [_P, C.POINTER(wdg_potential)]),
    "wdg_rhs": (C.c_int, [_P, _D]),
    "wdg_apply": (C.c_int, [_P, _D, _D, C.c_int32]),
    "wdg_apply_parts": (C.c_int, [_P, _D, _D, C.c_int32, C.c_int32]),
    "wdg_precond": (C.c_int, [_P, _D, _D]),
    "wdg_diagonal": (C.c_int, [_P, _D]),
    "wdg_solve": (C.c_int, [_P, C.POINTER(wdg_solve_opts), _D, _D, C.POINTER(wdg_solve_info)]),
    "wdg_observables": (C.c_int, [_P, _D, _D, _D, _D, _D]),
    "wdg_poisson_setup": (C.c_int, [_P, C.POINTER(wdg_poisson_desc)]),
    "wdg_poisson_step": (C.c_int, [_P, _D, C.c_double, _D, _D, C.POINTER(wdg_poisson_info)]),
    "wdg_apply_device": (C.c_int, [_P, _P, _P, C.c_int32]),
    "wdg_time_apply": (C.c_int, [_P, C.c_int32, C.c_int32, C.POINTER(C.c_double)]),
    "wdg_precond_device": (C.c_int, [_P, _P, _P, C.c_int32]),
    "wdg_precond_device_noalloc": (C.c_int, [_P, _P, _P, C.c_int32]),
    "wdg_host_setup": (C.c_int, [C.POINTER(wdg_desc), C.POINTER(wdg_potential), _D, _D]),
    "wdg_transient_rk4": (C.c_int, [_P, C.POINTER(wdg_transient_opts), _D, _D, _D, _D, _D, C.c_int32]),
    "wdg_transient_cn": (C.c_int, [_P, C.POINTER(wdg_cn_opts), _D, _D, C.POINTER(wdg_cn_info)]),
    "wdg_solve_resvec": (C.c_int, [_P, _D, C.c_int32, C.POINTER(C.c_int32)]),
    "wdg_solution": (C.c_int, [_P, _D]),
    "wdg_solve_batch": (C.c_int, [C.POINTER(_P), C.c_int32, C.POINTER(wdg_solve_opts), C.POINTER(wdg_solve_info)]),
    "wdg_transient_cn_batch": (C.c_int, [C.POINTER(_P), C.c_int32, C.POINTER(wdg_cn_opts), C.POINTER(_D),
                                         C.POINTER(_D), C.POINTER(wdg_cn_info)]),
    "wdg_set_theta_method": (C.c_int, [_P, C.c_int32]),
    "wdg_theta_method": (C.c_int, [_P]),
    "wdg_set_fft_block": (C.c_int, [_P, C.c_int32]),
    "wdg_set_memory_limits": (C.c_int, [_P, C.c_double, C.c_double]),
    "wdg_fft_block": (C.c_int, [_P]),
    "wdg2_create": (C.c_int, [C.POINTER(wdg2_desc), C.POINTER(wdg2_part), C.c_int32, C.POINTER(_P)]),
    "wdg2_set_potential": (C.c_int, [_P, C.POINTER(wdg2_potential)]),
    "wdg2_set_rhs": (C.c_int, [_P, _D]),
    "wdg2_set_dv_mode": (C.c_int, [_P, C.c_int32]),
    "wdg2_set_contact_rhs": (C.c_int, [_P, C.POINTER(wdg2_reservoir_grid), C.POINTER(wdg2_contact_modes),
                                       C.POINTER(wdg2_contact_modes)]),
    "wdg2_rowabs": (C.c_int, [_P, _D]),
    "wdg2_set_precond": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_int32)]),
    "wdg2_set_line_precond": (C.c_int, [_P, C.POINTER(wdg2_line_factors)]),
    "wdg2_observables": (C.c_int, [_P, _D, C.POINTER(C.c_int32), _D, C.POINTER(C.c_int32), _D]),
    "wdg2_poisson_setup": (C.c_int, [_P, C.POINTER(wdg2_poisson_desc)]),
    "wdg2_poisson_step": (C.c_int, [C.POINTER(_P), C.c_int32, _D, _D, C.c_int32, C.c_double, _D, _D, _D,
                                    C.POINTER(wdg2_poisson_info)]),
    "wdg_set_profiling": (C.c_int, [_P, C.c_uint32]),
    "wdg_profile": (C.c_int, [_P, C.c_char_p, C.c_int32, _D, C.POINTER(C.c_int64), _D, C.c_int32,
                              C.POINTER(C.c_int32)]),
}

_lib = None


def load() -> C.CDLL:
    """Load libwdg.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is not built; run python -m real_space_dg_amd.build")
    lib = C.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | C.RTLD_GLOBAL)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int, ctx=None):
    if rc != WDG_OK:
        lib = load()
        msg = lib.wdg_last_error(ctx).decode(errors="replace")
        raise WdgError(rc, msg)


def dptr(a: np.ndarray):
    """Pointer to a C-contiguous float64 (or complex128 viewed as float64) array."""
    if a is None:
        return None
    assert a.flags.c_contiguous and a.dtype in (np.float64, np.complex128)
    return a.ctypes.data_as(_D)
