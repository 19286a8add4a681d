"""GPU parity of the operator pieces against the FP64 oracle (small sizes)."""
import numpy as np
import pytest

from tests._cases import build_case, gpu_operator, line_operator, oracle_system, rand_c

pytestmark = pytest.mark.gpu

# (elements, degree, nk, bias, inflow model).  The negative bias makes the
# potential rise towards x = L, so min(EM) != EM(end) and the x >= L clamp of
# V_transient (to min(EM)) is exercised; DG_MSA_2D is the k_z-integrated inflow.
# Degrees 1-4 run the Theta kernels' compile-time transform plans (n <= 5;
# nk = 64 -> P = 128), degree 5 and nk = 96 the run-time plans.
CASES = [(16, 2, 64, 0.0, "DG_MSA_3D"), (8, 3, 64, 0.2, "DG_MSA_3D"), (12, 1, 96, 0.1, "DG_MSA_3D"),
         (6, 5, 64, 0.1, "DG_MSA_3D"), (8, 2, 64, -0.15, "DG_MSA_3D"), (10, 3, 64, 0.1, "DG_MSA_2D"),
         (8, 4, 64, 0.1, "DG_MSA_3D")]


@pytest.fixture(scope="module", params=CASES, ids=lambda c: "ne%d_p%d_nk%d_b%.2f_%s" % c)
def case(request):
    ne, deg, nk, bias, fermi = request.param
    print(f"\n[case] ne={ne} degree={deg} nk={nk} bias={bias} inflow={fermi}", flush=True)
    c = build_case(ne, deg, nk, bias, fermi_model=fermi)
    c["A"], c["G"], c["rhs"] = oracle_system(c)
    c["op"] = gpu_operator(c)
    yield c
    c["op"].close()


def test_rhs(case):
    r = case["op"].rhs()
    assert np.linalg.norm(r - case["rhs"]) <= 1e-12 * np.linalg.norm(case["rhs"])


METHODS = ["fft", "gemm"]


def test_auto_picks_fft(case):
    case["op"].set_theta_method("auto")
    assert case["op"].theta_method == "fft"


@pytest.mark.parametrize("method", METHODS)
def test_apply_fp64(case, method):
    case["op"].set_theta_method(method)
    u = rand_c(case["A"].shape[0], 1)
    y_ref = (case["A"] + case["G"]) @ u
    y = case["op"].apply(u, "fp64")
    err = np.linalg.norm(y - y_ref) / np.linalg.norm(y_ref)
    print(method, "apply fp64 rel err", err)
    assert err < 1e-13


@pytest.mark.parametrize("method", METHODS)
def test_apply_theta_part(case, method):
    case["op"].set_theta_method(method)
    u = rand_c(case["A"].shape[0], 2)
    y_ref = case["G"] @ u
    y = case["op"].apply(u, "fp64", parts="potential")
    err = np.linalg.norm(y - y_ref) / np.linalg.norm(y_ref)
    print(method, "theta fp64 rel err", err)
    assert err < 1e-13
    yt = case["op"].apply(u, "fp64", parts="transport")
    y_ref = case["A"] @ u
    assert np.linalg.norm(yt - y_ref) <= 1e-13 * np.linalg.norm(y_ref)


@pytest.mark.parametrize("method", METHODS)
def test_apply_mixed(case, method):
    case["op"].set_theta_method(method)
    u = rand_c(case["A"].shape[0], 3)
    y_ref = (case["A"] + case["G"]) @ u
    y = case["op"].apply(u, "mixed")
    err = np.linalg.norm(y - y_ref) / np.linalg.norm(y_ref)
    th = np.linalg.norm(case["op"].apply(u, "mixed", parts="potential") - case["G"] @ u) / \
        np.linalg.norm(case["G"] @ u)
    print(method, "apply mixed rel err", err, "theta part", th)
    if method == "gemm":
        # BF16 rounding of phi (unit roundoff 2^-9) in two GEMMs bounds the
        # Theta part; the transport part is FP32.
        assert err < 3e-3
        assert th < 5e-3
    else:
        # FP32 chirp-z transforms: a few FP32 roundings times log2(P)
        assert err < 1e-5
        assert th < 1e-5


@pytest.mark.parametrize("method", METHODS)
def test_observables_rho_u(case, method):
    case["op"].set_theta_method(method)
    p = case["p"]
    u = rand_c(case["A"].shape[0], 5)
    m_nodes = np.full(p["N_chi"] * p["N_K_chi"], 0.067 * 9.109e-31)
    from oracle import wigner1d as W
    n_ref, j_ref, ru_ref = W.observables(p, u, m_nodes)
    n, j, ru = case["op"].observables(m_nodes, rho=u, want_rho_u=True)
    assert np.max(np.abs(ru - ru_ref)) <= 1e-12 * np.max(np.abs(ru_ref))
    assert np.max(np.abs(n - n_ref)) <= 1e-12 * np.max(np.abs(n_ref))
    assert np.max(np.abs(j - j_ref)) <= 1e-10 * np.max(np.abs(j_ref))


def test_diagonal(case):
    d = case["op"].diagonal()
    d_ref = (case["A"] + case["G"]).diagonal()
    assert np.linalg.norm(d - d_ref) <= 1e-12 * np.linalg.norm(d_ref)


def test_precond_inverts_line_operator(case):
    p = case["p"]
    P = line_operator(case["A"], case["G"], p["N_chi"], p["N_K_chi"], p["Nk"])
    r = rand_c(P.shape[0], 4)
    z = case["op"].precond(r)
    res = np.linalg.norm(P @ z - r) / np.linalg.norm(r)
    print("precond residual", res)
    assert res < 1e-11


def test_precond_fp32_matches_fp64(case):
    """The FP32 line preconditioner (stored block inverses: D^-1 in BF16,
    the coupling columns in FP32) against the FP64 sweeps, both through
    wdg_precond_device in device units.  BF16 rounds each stored entry to
    2^-9 relative, which bounds the difference (the mixed solver's FP64
    refinement, tested elsewhere, keeps the solution at FP64 accuracy)."""
    import torch
    from real_space_dg_amd import _lib as L
    from real_space_dg_amd.solver import check
    op = case["op"]
    r = rand_c(op.N, 5)
    r64 = torch.tensor(r, dtype=torch.complex128, device="cuda")
    r32 = r64.to(torch.complex64)
    z64 = torch.empty_like(r64)
    z32 = torch.empty_like(r32)
    torch.cuda.synchronize()
    check(op._lib.wdg_precond_device(op._h, r64.data_ptr(), z64.data_ptr(), L.WDG_PREC_FP64), op._h)
    check(op._lib.wdg_precond_device(op._h, r32.data_ptr(), z32.data_ptr(), L.WDG_PREC_MIXED), op._h)
    a, b = z32.cpu().numpy().astype(np.complex128), z64.cpu().numpy()
    err = np.linalg.norm(a - b) / np.linalg.norm(b)
    print("precond fp32 vs fp64", err)
    assert err < 4e-3
