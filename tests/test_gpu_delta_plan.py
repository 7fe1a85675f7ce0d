"""Delta-f storage on a single plan (vpfp_plan_set_delta) through the fused
step loop vpfp_step_n with storage moments.

Workload: nx = 32, nv = 256, v in [-6.4, 6.4], k = 0.3, density
1 + 0.05 cos(kx) on a unit Maxwellian, Dougherty nu = 1e-3, dt = 0.05, 200
steps.  FP64: the delta-f plan reproduces the plain plan (field history and
moments within 1e-11 of their largest values, f within 1e-12), since the
algorithm is the same.  FP32: the field history is closer to the FP64 one
with delta-f storage than with f stored.  Entry points whose result would
change meaning refuse while delta-f is on.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NX, NV, K, AMP, NU, DT, STEPS = 32, 256, 0.3, 0.05, 1e-3, 0.05, 200


def _run(precision, delta, row_q=None):
    import torch

    from vlapy_amd import ops
    from vlapy_amd._lib import NMOMENTS, Plan
    from vlapy_amd.scan import instance_grids

    x, kx, okx, v, kv, dv = instance_grids(NX, K, NV, 6.4)
    plan = Plan(NX, NV, precision, kx, okx, v, kv, dv, NU, DT, "dg")
    if row_q is not None:
        plan.set_tuning(row_q=row_q)
    dev = torch.device("cuda:0")
    m = np.exp(-v ** 2 / 2.0) / np.sqrt(2.0 * np.pi)
    f0 = (1.0 + AMP * np.cos(K * x))[:, None] * m[None, :]
    if delta:
        plan.set_delta(True)
        f0 = f0 - m[None, :]
    f = torch.as_tensor(f0, dtype=ops.f_dtype(plan), device=dev).contiguous()
    e = torch.zeros(NX, dtype=torch.float64, device=dev)
    eh = torch.zeros(STEPS, NX, dtype=torch.float64, device=dev)
    mo = torch.zeros(STEPS, NMOMENTS, NX, dtype=torch.float64, device=dev)
    ops.step_n(plan, f, e, STEPS, None, eh, mo, None)
    if delta:
        plan.background(f, 1)
    torch.cuda.synchronize()
    return plan, eh.cpu().numpy(), mo.cpu().numpy(), f.double().cpu().numpy()


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("row_q", [0, 4, 8, 16])
def test_plan_delta_f_fp64_reproduces_plain_plan(row_q):
    """Every row-kernel shape: 0 (row-count rule), 4, 8, and 16 (the Q = 8
    kernel under delta-f storage; Q = 16 does not exist at nv = 256)."""
    _, eh, mo, f = _run("fp64", False, row_q)
    _, ehd, mod, fd = _run("fp64", True, row_q)
    print("fp64 delta-f vs plain: E %.2e moments %.2e f %.2e" % (_rel(ehd, eh), _rel(mod[:, :6], mo[:, :6]),
                                                                _rel(fd, f)))
    assert _rel(ehd, eh) < 1e-11
    # the v^4 and v^5 moments weight the velocity tails, where an absolute
    # rounding difference of ~1e-17 in f is ~1e-8 of f itself, by up to
    # 6.4^5 ~ 1e4: 5e-11 for them (measured 1.3e-11 at Q = 8), 1e-11 otherwise
    for q in range(8):
        assert _rel(mod[:, q], mo[:, q]) < (5e-11 if q in (4, 5) else 1e-11), q
    assert _rel(fd, f) < 1e-12


def test_plan_delta_f_fp32_is_closer_to_fp64():
    _, eh64, _, f64 = _run("fp64", False)
    _, eh32, _, f32 = _run("fp32", False)
    _, ehd, _, fd = _run("fp32", True)
    err_plain, err_delta = _rel(eh32, eh64), _rel(ehd, eh64)
    print("fp32 E error vs fp64: plain %.2e delta-f %.2e; f: plain %.2e delta-f %.2e"
          % (err_plain, err_delta, _rel(f32, f64), _rel(fd, f64)))
    assert err_delta < err_plain
    assert _rel(fd, f64) < _rel(f32, f64)


def test_plan_delta_f_refuses_entry_points_that_read_f_itself():
    import torch

    from vlapy_amd import ops

    plan, _, _, _ = _run("fp64", True)
    f = torch.zeros(NX, NV, dtype=torch.float64, device="cuda:0")
    n = torch.zeros(NX, dtype=torch.float64, device="cuda:0")
    with pytest.raises(NotImplementedError):
        ops.compute_charges(plan, f, n)
    plan.set_delta(False)
    ops.compute_charges(plan, f, n)
    torch.cuda.synchronize()
