"""FP32 delta-f storage on the coarse grid of the closure configuration,
measured against the fine grid.

32 held-out runs (draw_instances(32, seed=12): k ~ U[0.25, 0.45], amplitude
10^U[-4, -1], nu 10^U[-5, -2]; nx = 64, dt = 0.05, 4000 steps) on the fine
grid (nv = 4096, FP64) and on the coarse grid (nv = 256, FP32) with f and with
f - M stored.  Storing the deviation removes the FP32 rounding floor of the
coarse grid: asserted that every fitted run's damping rate is within 1 % of
the fine grid's, that the <E^2> history error is at most half the plain
coarse run's, and that the density drifts less than the plain coarse run's.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_coarse_delta_f_matches_fine_grid():
    from vlapy_amd import closure as CL

    inst = CL.draw_instances(32, 12)
    steps = 4000
    t = CL.DT * (np.arange(steps) + 1)
    print("delta-f coarse: fine FP64 rollout", flush=True)
    ref = CL.rollout(inst, CL.NV_FINE, "fp64", steps)
    plain = CL.rollout(inst, CL.NV, "fp32", steps)
    drift_plain = CL.rollout.density_drift.max()
    delta = CL.rollout(inst, CL.NV, "fp32", steps, delta=True)
    drift_delta = CL.rollout.density_drift.max()
    assert np.all(np.isfinite(delta))
    res = {}
    for name, e2 in (("plain", plain), ("delta-f", delta)):
        h = CL.history_errors(ref, e2)
        d = CL.damping_errors(inst, ref, e2, t)
        ok = ~np.isnan(d)
        res[name] = (np.median(h), np.median(d[ok]), np.mean(d[ok] <= 0.01), int(ok.sum()))
        print(name, "history err median %.4f damping err median %.5f within 1%% %.3f (%d fitted)" % res[name],
              flush=True)
    print("density drift max: plain %.2e delta-f %.2e" % (drift_plain, drift_delta), flush=True)
    assert res["delta-f"][3] >= 20
    assert res["delta-f"][2] == 1.0
    assert res["delta-f"][0] <= 0.5 * res["plain"][0]
    assert drift_delta <= drift_plain
