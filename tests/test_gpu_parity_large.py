"""GPU parity of the operator apply at the sizes the benchmarks run, against
the oracle's matrix-free FP64 operator (oracle.wigner1d.MatrixFreeOperator:
transport of DG_code/core/get_Diff.m, Theta[V] as the two dense phi products of
DG_code/core/get_Drift_accel_trial_permutedShell.m:95-151).

Each shape selects a different chirp-z plan of the fused Theta kernel:
  512 x 4 x 1024  (the 0.2 V RTD of the headline solve)   P = 2048, compile-time plan
  256 x 4 x 512   (the self-consistent RTD)               P = 1024, compile-time plan
  128 x 4 x 768   (a bias-sweep instance)                 P = 1536, run-time plan
  128 x 4 x 256   (a bias-sweep instance)                 P = 512,  compile-time plan
and the FP32 apply is run in all three kernel forms: the chirp-z kernel in both
workgroup shapes (256 threads looping over the nodes, and the ~1024-thread
wide block with its aliased LDS buffers) and the near-DFT kernel (length-nk
transforms, nk = 1024/512/768/256 -> compile-time/compile-time/run-time/
compile-time plan, two Taylor terms of exp(i s ubar vbar) for nk >= 512 and
three for nk = 256).
Bounds: 1e-13 relative for the FP64 apply, 1e-12 for its Theta part alone,
1e-5 for the FP32 (mixed inner) apply and its Theta part.
"""
import numpy as np
import pytest

from oracle import wigner1d as W
from tests._cases import build_case, gpu_operator, rand_c

pytestmark = pytest.mark.gpu

SHAPES = [(512, 3, 1024, 0.2, 2048), (256, 3, 512, 0.1, 1024), (128, 3, 768, 0.3, 1536),
          (128, 3, 256, 0.35, 512)]


@pytest.fixture(scope="module", params=SHAPES, ids=lambda c: "ne%d_p%d_nk%d_P%d" % (c[0], c[1], c[2], c[4]))
def big(request):
    ne, deg, nk, bias, P = request.param
    print(f"\n[case] ne={ne} degree={deg} nk={nk} bias={bias} (P={P})", flush=True)
    c = build_case(ne, deg, nk, bias)
    c["mf"] = W.MatrixFreeOperator.build(c["mat"], c["p"], c["E"][0, :, 0])
    c["bbar"] = c["mf"].edge_coefficients()
    c["op"] = gpu_operator(c)
    c["P"] = P
    u = rand_c(c["op"].N, 11)
    c["u"] = u
    c["y_tr"] = c["mf"].apply_transport(u)
    c["y_th"] = c["mf"].apply_drift(u, c["bbar"])
    yield c
    c["op"].close()


def _errs(c, prec):
    """Relative error of the whole apply and of its potential part alone (the
    transport dominates the norm of A u by ~1e4 at these sizes, so the
    potential part is compared on its own, wdg_apply_parts)."""
    op, u = c["op"], c["u"]
    y = op.apply(u, prec)
    y_ref = c["y_tr"] + c["y_th"]
    full = np.linalg.norm(y - y_ref) / np.linalg.norm(y_ref)
    yt = op.apply(u, prec, parts="potential")
    theta = np.linalg.norm(yt - c["y_th"]) / np.linalg.norm(c["y_th"])
    ytr = op.apply(u, prec, parts="transport")
    tr = np.linalg.norm(ytr - c["y_tr"]) / np.linalg.norm(c["y_tr"])
    return full, theta, tr


def test_fp64_apply_matches_oracle(big):
    op = big["op"]
    op.set_theta_method("fft")
    full, theta, tr = _errs(big, "fp64")
    print(f"fp64 apply rel err {full:.2e}, Theta part {theta:.2e}, transport part {tr:.2e}")
    assert full < 1e-13
    assert theta < 1e-12
    assert tr < 1e-13


@pytest.mark.parametrize("block", ["256", "wide", "ndft"])
def test_fp32_apply_matches_oracle(big, block):
    op = big["op"]
    op.set_theta_method("fft")
    op.set_fft_block(block)
    threads = op.fft_block
    tt = big["P"] // 8
    if block == "ndft":   # one group of nk/8 threads per DG node
        assert threads == 4 * (big["p"]["Nk"] // 8)
    else:
        assert threads == (tt * max(1, 256 // tt) if block == "256" else tt * (1024 // tt))
    full, theta, tr = _errs(big, "mixed")
    op.set_fft_block("auto")
    print(f"fp32 apply ({threads} threads) rel err {full:.2e}, Theta part {theta:.2e}, transport part {tr:.2e}")
    assert full < 1e-5
    assert theta < 1e-5
    assert tr < 1e-5
