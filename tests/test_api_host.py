"""Host-side behaviour of the vlapy.core-compatible package (no GPU needed):
factory-time errors for unknown names, grids, driver, diagnostics and the
inner-loop bookkeeping, each against the oracle."""

import json
import os

import numpy as np
import pytest

from oracle import vpfp_oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def stuff():
    g = O.Grid.make(8, 16, 2 * np.pi / 0.3, 6.0)
    return {"kx": g.kx, "x": g.x, "v": g.v, "kv": g.kv, "dv": g.dv, "one_over_kx": g.one_over_kx,
            "nx": 8, "nv": 16, "nu": 0.0, "dt": 0.1, "driver_function": lambda t: np.zeros(8)}


def test_unknown_names_raise_at_factory_time():
    from vlapy_amd import outer_loop
    from vlapy_amd.core import collisions, field, step, vlasov, vlasov_poisson

    s = stuff()
    with pytest.raises(NotImplementedError):
        vlasov.get_vdfdx(s, "sl")
    with pytest.raises(NotImplementedError):
        vlasov.get_edfdv(s, "cd2")
    with pytest.raises(NotImplementedError):
        field.get_field_solver(s, "fem")
    with pytest.raises(NotImplementedError):
        vlasov_poisson.get_time_integrator("rk4", None, None, None, s)
    with pytest.raises(NotImplementedError):
        collisions.get_matrix_solver(8, 16, "lu")
    with pytest.raises(NotImplementedError):
        collisions.get_batched_array_maker(s["v"], 16, 8, 1e-2, 0.1, s["dv"], operator="fp")
    with pytest.raises(NotImplementedError):
        step.get_collision_step(s, {"nu": -1.0, "fokker-planck": {"type": "lb", "solver": "batched_tridiagonal"}})
    with pytest.raises(NotImplementedError):
        step.get_collision_step(s, {"nu": 1e-2, "fokker-planck": {"type": "lb", "solver": "sor"}})
    with pytest.raises(NotImplementedError):
        step.get_f_update({"space": "some"})
    with pytest.raises(NotImplementedError):
        outer_loop.get_inner_loop_stepper({"backend": {"core": "numpy"}}, s, 4)


def test_zero_collision_step_is_identity():
    from vlapy_amd.core import step

    f = np.random.default_rng(0).random((8, 16))
    fp = step.get_collision_step(stuff(), {"nu": 0.0, "fokker-planck": {"type": "dg", "solver": "naive"}})
    assert fp(f) is f


def test_grids_match_oracle_bitwise():
    from vlapy_amd import initializers as I

    for vmax, nv in ((6.4, 512), (6.0, 1024), (8.0, 8192)):
        a, b = I.initialize_velocity_quantities(vmax, nv), O.velocity_grid(vmax, nv)
        assert a[0] == b[0] and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    for nx, k in ((64, 0.3), (1024, 0.35), (96, 0.25)):
        a, b = I.initialize_spatial_quantities(0.0, 2 * np.pi / k, nx), O.spatial_grid(0.0, 2 * np.pi / k, nx)
        assert all(np.array_equal(p, q) for p, q in zip(a, b))
    np.testing.assert_allclose(I.initialize_distribution(4, 256, 6.0), O.maxwellian(4, 256, 6.0), rtol=1e-15)


def test_driver_matches_oracle():
    from vlapy_amd import field_driver

    x = np.linspace(0.1, 20.0, 32)
    pulses = {"a": {"k0": 0.3, "w0": 1.16, "t_L": 6, "t_R": 20, "t_wL": 2.5, "t_wR": 2.5, "a0": 1e-7},
              "b": {"k0": 0.6, "w0": 1.5, "t_L": 1, "t_R": 9, "t_wL": 1.5, "t_wR": 0.5, "a0": 1e-3}}
    fn = field_driver.get_driver_function(x, pulses)
    for t in (0.0, 3.3, 12.0, 40.0):
        np.testing.assert_allclose(fn(t), O.driver_field(x, pulses, t), rtol=1e-14, atol=1e-20)
    assert np.all(field_driver.get_driver_function(x, {})(1.0) == 0.0)


def test_z_function_roots():
    from vlapy_amd.diagnostics import z_function

    tab = json.load(open(os.path.join(GOLD, "canosa_1973.json")))
    for k, wr, wi in zip(tab["k0"], tab["w_real"], tab["w_imag"]):
        w = z_function.get_roots_to_electrostatic_dispersion(wp_e=1.0, vth_e=1.0, k0=k)
        np.testing.assert_almost_equal(w, wr + 1j * wi, decimal=4)


def test_diagnostics_match_oracle():
    from vlapy_amd.diagnostics import low_level_helpers as llh

    t = np.arange(1, 2001) * 0.05
    x = np.linspace(0, 2 * np.pi / 0.3, 64, endpoint=False)
    amp = 1e-3 * np.exp(-0.0126 * t) * np.cos(1.16 * t + 0.2)
    e = amp[:, None] * np.sin(0.3 * x)[None, :]
    g1 = llh.get_standing_wave_rate_and_frequency(np.mean(e ** 2, axis=1), t, t_min=5.0)
    g2 = O.damping_rate_and_frequency(np.mean(e ** 2, axis=1), t, t_min=5.0)
    np.testing.assert_allclose(g1, g2, rtol=1e-12)
    assert g1[0] == pytest.approx(-0.0126, rel=2e-3) and g1[1] == pytest.approx(1.16, rel=2e-3)
    assert llh.get_damping_rate(e, t=t) == pytest.approx(O.damping_rate_reference(e, t), rel=1e-12)
    with pytest.raises(ValueError):
        llh.get_standing_wave_rate_and_frequency(np.mean(e ** 2, axis=1), t, t_min=6.0, t_max=3.0)


def test_post_inner_loop_update_matches_oracle():
    from vlapy_amd import outer_loop

    rng = np.random.default_rng(1)
    ts = {"mean_cum_de2_previous": 0.3, "series": {k: rng.random(5) for k in ("mean_T", "mean_e2", "mean_de2")}}
    ref = {"mean_cum_de2_previous": 0.3, "series": {k: v.copy() for k, v in ts["series"].items()}}
    outer_loop.post_inner_loop_update(ts)
    s = ref["series"]
    cum = 0.3 + np.cumsum(s["mean_de2"])
    np.testing.assert_array_equal(ts["series"]["mean_cum_de2"], cum)
    np.testing.assert_array_equal(ts["series"]["mean_t_plus_e2_minus_cum_de2"], s["mean_T"] + (s["mean_e2"] - cum))
    assert ts["mean_cum_de2_previous"] == cum[-1]


def test_arrays_for_inner_loop_schema():
    from vlapy_amd import outer_loop

    s = stuff()
    s["f"], s["e"] = np.ones((8, 16)), np.zeros(8)
    a = outer_loop.get_arrays_for_inner_loop(s, 5, {"space": "all"})
    assert a["stored_f"].shape == (5, 8, 16) and a["fields"]["vN"].shape == (5, 8)
    b = outer_loop.get_arrays_for_inner_loop(s, 5, {"time": "first-last", "space": ["k0", "k1"]})
    assert b["stored_f"].shape == (5, 2, 16) and b["stored_f"].dtype == np.complex64
    assert set(a["series"]) == set(O.SERIES_KEYS) and set(a["fields"]) == set(O.FIELD_KEYS)


def test_bench_refuses_several_gpus():
    """bench.py measures one GPU: --gpus N > 1 exits with status 2 and says
    why, before touching a GPU."""
    import subprocess
    import sys

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    res = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                         text=True, timeout=300)
    assert res.returncode == 2
    assert "one GPU" in res.stderr


def test_plan_background_validates_f_before_the_device_call():
    """Plan.background writes nx * nv elements of the plan's storage type:
    a wrong shape, dtype or layout, or a host tensor, raises ValueError
    before any C call (the handle here is never used)."""
    import types

    import torch

    from vlapy_amd._lib import Plan

    fake = types.SimpleNamespace(nx=4, nv=8, precision="fp64", handle=None)
    for bad in (torch.zeros(4, 8, dtype=torch.float32), torch.zeros(4, 16, dtype=torch.float64),
                torch.zeros(8, 4, dtype=torch.float64).t(), torch.zeros(4, 8, dtype=torch.float64)):
        with pytest.raises(ValueError):
            Plan.background(fake, bad, 1)


def test_fused_inner_loop_driver_batch_equals_per_step_calls():
    """The fused inner loop evaluates driver_function(t + dt) for a whole
    batch at once when the function broadcasts over a column of times, and
    falls back to one call per step otherwise; both give the per-step values
    the reference's leapfrog step uses (vlasov_poisson.py:55)."""
    from vlapy_amd import field_driver
    from vlapy_amd.outer_loop import _FusedInnerLoop

    g = O.Grid.make(16, 8, 2 * np.pi / 0.3, 6.0)
    pulses = {"p": {"k0": 0.3, "w0": 1.16, "t_L": 2.0, "t_R": 20.0, "t_wL": 2.5, "t_wR": 2.5, "a0": 1e-2}}
    t = 0.1 * np.arange(40)

    class Fake:
        dt, nx = 0.1, 16
        _driver_at = _FusedInnerLoop._driver_at

    for fn in (field_driver.get_driver_function(g.x, pulses),
               lambda tt: O.driver_field(g.x, pulses, tt),
               lambda tt: np.full(16, float(tt)) * 1e-3,      # does not broadcast: per-step fallback
               lambda tt: np.zeros(16)):                       # time-independent
        f = Fake()
        f.driver_function = fn
        got = _FusedInnerLoop._driver(f, t)
        want = np.stack([np.broadcast_to(np.asarray(fn(ti + 0.1), dtype=np.float64), (16,)) for ti in t])
        assert got.shape == (40, 16)
        np.testing.assert_allclose(got, want, rtol=1e-14, atol=1e-300)


def test_scheme_field_times_follow_the_core_integrators():
    """The fused PEFRL / 6th-order batches evaluate the driver at the times
    the per-operator integrators use: PEFRL's running sum of the drift
    lengths, the 6th-order scheme's t + a_k dt."""
    from vlapy_amd.core import vlasov_poisson as VP
    from vlapy_amd.outer_loop import _scheme_field_times

    seen = []

    def drift(f, dt=None, **kw):
        return f

    def kick(f, e=None, dt=None, **kw):
        return f

    def field(driver_field=None, f=None):
        return driver_field

    t = np.array([0.0, 0.3, 1.7])
    for name in ("pefrl", "h-sixth"):
        seen.clear()
        step = VP.get_time_integrator(name, drift, kick, field,
                                      {"dt": 0.1, "driver_function": lambda tt: seen.append(tt) or 0.0})
        want = []
        for ti in t:
            seen.clear()
            step(0.0, np.zeros(4), ti)
            want.append(list(seen))
        np.testing.assert_array_equal(_scheme_field_times(name, t, 0.1), np.array(want))
