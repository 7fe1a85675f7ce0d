"""The allocation-free field solve (vpfp_field_noalloc) against the plan-based
vpfp_field it varies, through the C ABI.

Each case makes two calls in a row on the shared scratch, the second with a
smaller grid, and checks both against vpfp_field on the same f: FP64 and
FP32 storage, power-of-two nx (the FFT field kernel with the full twiddle
table, and nx = 2048 with the two-level table) and other nx (the direct
DFT), with and without a driver.  The twiddles are formed on the device, so
agreement is to rounding: 1e-12 of max |E|.  nx above the scratch's limit
is refused with ValueError.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _field_pair(nx, nv, precision, seed, with_driver):
    import torch

    from vlapy_amd import ops
    from vlapy_amd._lib import Plan
    from vlapy_amd.scan import instance_grids

    dev = torch.device("cuda:0")
    x, kx, okx, v, kv, dv = instance_grids(nx, 0.3, nv, 6.4)
    plan = Plan(nx, nv, precision, kx, okx, v, kv, dv, 0.0, 0.1, "dg")
    rng = np.random.default_rng(seed)
    dens = 1.0 + 0.2 * rng.standard_normal(nx) / np.sqrt(nx)
    f0 = dens[:, None] * (np.exp(-v ** 2 / 2.0) / np.sqrt(2 * np.pi))[None, :] * (1 + 0.01 * rng.random((nx, nv)))
    f = torch.as_tensor(f0, dtype=ops.f_dtype(plan), device=dev).contiguous()
    drv = torch.as_tensor(0.01 * rng.standard_normal(nx), device=dev) if with_driver else None
    e_plan = torch.zeros(nx, dtype=torch.float64, device=dev)
    e_free = torch.full((nx,), np.nan, dtype=torch.float64, device=dev)
    okx_d = torch.as_tensor(okx, device=dev)
    ops.field(plan, f, drv, e_plan)
    ops.field_noalloc(nx, nv, precision, dv, okx_d, f, drv, e_free)
    torch.cuda.synchronize()
    a, b = e_plan.cpu().numpy(), e_free.cpu().numpy()
    assert np.all(np.isfinite(b))
    return float(np.abs(a - b).max() / np.abs(a).max())


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("big,small", [((128, 512), (64, 256)), ((2048, 64), (256, 128)), ((48, 256), (24, 64)),
                                       ((4096, 16), (6, 32))])
def test_noalloc_field_equals_plan_field_on_two_calls(precision, big, small):
    for i, (nx, nv) in enumerate((big, small)):
        print("field_noalloc", precision, "nx", nx, "nv", nv, flush=True)
        err = _field_pair(nx, nv, precision, 10 * nx + i, with_driver=(i == 1))
        print("  max |E_noalloc - E_plan| / max |E_plan| = %.2e" % err, flush=True)
        assert err <= 1e-12, (nx, nv, err)


def test_noalloc_field_refuses_grids_beyond_its_scratch():
    import torch

    from vlapy_amd import _lib, ops

    dev = torch.device("cuda:0")
    nx, nv = _lib.NOALLOC_MAX_NX + 1, 4
    okx = torch.zeros(nx, dtype=torch.float64, device=dev)
    f = torch.ones(nx, nv, dtype=torch.float64, device=dev)
    e = torch.zeros(nx, dtype=torch.float64, device=dev)
    with pytest.raises(ValueError):
        ops.field_noalloc(nx, nv, "fp64", 0.1, okx, f, None, e)
