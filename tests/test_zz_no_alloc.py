"""wdg_precond_device_noalloc (no device allocation: its line-sweep
workspace is the library's file-scope scratch array) against
wdg_precond_device, bit for bit, in both precisions, on two calls in a row
with the second on a smaller grid (the array keeps the first call's
contents)."""
import numpy as np
import pytest

from real_space_dg_amd import _lib as L
from real_space_dg_amd.solver import check
from tests._cases import build_case, gpu_operator, rand_c

pytestmark = pytest.mark.gpu


def _pair(op, prec, seed):
    import torch
    dt = torch.complex128 if prec == L.WDG_PREC_FP64 else torch.complex64
    r = torch.tensor(rand_c(op.N, seed), dtype=torch.complex128, device="cuda").to(dt)
    z_ref = torch.empty_like(r)
    z = torch.full_like(r, float("nan"))
    torch.cuda.synchronize()
    check(op._lib.wdg_precond_device(op._h, r.data_ptr(), z_ref.data_ptr(), prec), op._h)
    check(op._lib.wdg_precond_device_noalloc(op._h, r.data_ptr(), z.data_ptr(), prec), op._h)
    torch.cuda.synchronize()
    return z.cpu().numpy(), z_ref.cpu().numpy()


@pytest.mark.parametrize("prec", [L.WDG_PREC_FP64, L.WDG_PREC_MIXED], ids=["fp64", "mixed"])
def test_noalloc_precond_matches_the_allocating_entry_point(prec):
    big = gpu_operator(build_case(64, 3, 256, 0.2))
    small = gpu_operator(build_case(12, 2, 64, 0.1))
    try:
        for op, seed in ((big, 1), (small, 2), (big, 3)):
            print(f"[noalloc] N = {op.N}, precision {prec}", flush=True)
            z, z_ref = _pair(op, prec, seed)
            assert np.isfinite(z).all()
            assert np.array_equal(z, z_ref)
    finally:
        big.close()
        small.close()


def test_noalloc_requires_a_potential():
    from real_space_dg_amd.solver import WignerOperator
    op = WignerOperator(8, 3, 64, 60e-9, 60e-9)
    try:
        rc = op._lib.wdg_precond_device_noalloc(op._h, 1, 1, L.WDG_PREC_FP64)
        assert rc == L.WDG_ERR_STATE
    finally:
        op.close()
