"""The vlapy.core-compatible closures on the GPU, called the way the
reference's code and tests call them, against the reference's own golden
vectors (tests/golden/vlapy_core_small.npz, written from vlapy.core) and the
oracle.  FP64 bar: 1e-12 relative per call; 1e-11 over a few composed steps.
"""

import copy
import os

import numpy as np
import pytest

from oracle import vpfp_oracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
PULSES = {"p": {"k0": 0.35, "w0": 1.2, "t_L": 0.0, "t_R": 20.0, "t_wL": 2.5, "t_wR": 2.5, "a0": 0.05}}


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.fixture(scope="module")
def z():
    import torch

    assert torch.cuda.is_available()
    return dict(np.load(os.path.join(GOLD, "vlapy_core_small.npz")))


def stuff_from(z, nu=0.0):
    from vlapy_amd import field_driver

    x = z["x"]
    return {"kx": z["kx"], "x": x, "v": z["v"], "kv": z["kv"], "dv": float(z["dv"]),
            "one_over_kx": z["one_over_kx"], "nx": x.size, "nv": z["v"].size, "dt": float(z["dt"]), "nu": nu,
            "driver_function": field_driver.get_driver_function(x, PULSES), "rules_to_store_f": {"space": "all"}}


def test_pushers_and_field_vs_reference_golden(z):
    from vlapy_amd.core import field, vlasov

    s = stuff_from(z)
    dt, t0 = float(z["dt"]), float(z["t0"])
    assert rel(vlasov.get_vdfdx(s, "exponential")(z["f0"], dt), z["vdfdx_f"]) < 1e-12
    assert rel(vlasov.get_edfdv(s, "exponential")(z["f0"], z["e0"], dt), z["edfdv_f"]) < 1e-12
    fs = field.get_field_solver(s, "spectral")
    assert rel(fs(s["driver_function"](t0), z["f0"]), z["field_e"]) < 1e-12


@pytest.mark.parametrize("name", ["leapfrog", "pefrl", "h-sixth"])
def test_integrators_vs_reference_golden(z, name):
    from vlapy_amd.core import field, vlasov, vlasov_poisson

    s = stuff_from(z)
    vp = vlasov_poisson.get_time_integrator(name, vlasov.get_vdfdx(s), vlasov.get_edfdv(s),
                                            field.get_field_solver(s), s)
    e, f = vp(e=z["e0"], f=z["f0"], t=float(z["t0"]))
    assert rel(e, z[name + "_e"]) < 1e-11
    assert rel(f, z[name + "_f"]) < 1e-12


@pytest.mark.parametrize("op", ["lb", "dg"])
def test_collision_step_vs_reference_golden(z, op):
    from vlapy_amd.core import step

    fc = z["coll_" + op + "_in"]
    s = {"f": fc, "v": z["v"], "nv": z["v"].size, "nx": 3, "nu": 3e-2, "dt": 0.2, "dv": float(z["dv"])}
    for solver in ("batched_tridiagonal", "naive"):
        fp = step.get_collision_step(s, {"fokker-planck": {"type": op, "solver": solver}, "nu": 3e-2})
        assert rel(fp(fc.copy()), z["coll_" + op + "_out"]) < 1e-12


@pytest.mark.parametrize("op", ["lb", "dg"])
def test_matrix_maker_and_solvers(z, op):
    from vlapy_amd.core import collisions

    fc = z["coll_" + op + "_in"]
    v, dv = z["v"], float(z["dv"])
    mk = collisions.get_batched_array_maker(v, v.size, 3, 3e-2, 0.2, dv, operator=op)
    a, b, c = mk(fc)
    wa, wb, wc = O.collision_diagonals(fc, v, 3e-2, 0.2, dv, op)
    assert rel(a, wa) < 1e-12 and rel(b, wb) < 1e-12 and rel(c, wc) < 1e-12
    x = collisions.get_matrix_solver(3, v.size, "batched_tridiagonal")(a, b, c, fc)
    assert rel(x, z["coll_" + op + "_out"]) < 1e-12
    f_in = fc.copy()
    out = collisions.get_matrix_solver(3, v.size, "naive")(a, b, c, f_in)
    assert out is f_in and rel(f_in, z["coll_" + op + "_out"]) < 1e-12


def test_timestep_storage_vs_reference_golden(z):
    from vlapy_amd.core import step

    s = stuff_from(z, nu=1e-2)
    ap = {"vlasov-poisson": {"time": "leapfrog", "vdfdx": "exponential", "edfdv": "exponential",
                             "poisson": "spectral"},
          "fokker-planck": {"type": "dg", "solver": "batched_tridiagonal"}, "nu": 1e-2}
    ts = step.get_timestep(ap, s)
    nt, nx, nv = 3, s["nx"], s["nv"]
    temp = {"time_batch": float(z["t0"]) + float(z["dt"]) * np.arange(nt), "driver_array_batch": z["ts_driver"],
            "e": z["e0"].copy(), "f": z["f0"].copy(), "stored_f": np.zeros((nt, nx, nv)),
            "series": {k: np.zeros(nt) for k in O.SERIES_KEYS},
            "fields": {k: np.zeros((nt, nx)) for k in O.FIELD_KEYS}}
    for i in range(nt):
        temp, _ = ts(temp, i)
    assert rel(temp["f"], z["ts_f"]) < 1e-11
    assert rel(temp["e"], z["ts_e"]) < 1e-11
    for k in O.FIELD_KEYS:
        ref = z["ts_fields_" + k]
        p = {"n": 0, "j": 1, "T": 2, "q": 3, "fv4": 4, "vN": 5}.get(k)
        scale = np.abs(ref).max() if p is None else O.trapz_v(np.abs(z["ts_f"]) * np.abs(s["v"]) ** p, s["dv"]).max()
        assert np.abs(temp["fields"][k] - ref).max() <= 1e-11 * scale, k
    for k in O.SERIES_KEYS:
        np.testing.assert_allclose(temp["series"][k], z["ts_series_" + k], rtol=1e-10, err_msg=k)


def test_solve_for_field_reference_closed_forms():
    """tests/test_fieldsolver.py:27-54 of the reference, through the drop-in API (nx=96)."""
    from vlapy_amd.core import field

    nx, kp = 96, 0.25
    xmax = 2 * np.pi / kp
    dx = xmax / nx
    x = np.linspace(dx / 2, xmax - dx / 2, nx)
    kx = np.fft.fftfreq(x.size, d=dx) * 2.0 * np.pi
    okx = np.zeros_like(kx)
    okx[1:] = 1.0 / kx[1:]
    dens = [1.0 + np.sin(kp * x), 1.0 + np.cos(2 * kp * x), 1.0 + np.sin(2 * kp * x) + np.cos(8 * kp * x)]
    want = [np.cos(kp * x) / kp, -np.sin(2 * kp * x) / 2.0 / kp,
            np.cos(2 * kp * x) / 2.0 / kp - np.sin(8 * kp * x) / 8.0 / kp]
    for n, w in zip(dens, want):
        np.testing.assert_almost_equal(w, field.solve_for_field(1.0 - n, okx), decimal=12)


@pytest.mark.parametrize("op", ["lb", "dg"])
def test_reference_collision_suite_on_gpu(op):
    """tests/test_collisions.py:39-213 of the reference through the drop-in API."""
    from vlapy_amd.core import step

    def run(vshift):
        nx, nv, nu, dt, vmax = 2, 1024, 1e-2, 0.1, 6.0
        dv = 2 * vmax / nv
        v = np.linspace(-vmax + dv / 2.0, vmax - dv / 2.0, nv)
        f = np.repeat(np.exp(-((v - vshift) ** 2.0) / 2.0)[None, :], nx, axis=0)
        f = f / O.trapz_v(f, dv)[:, None]
        s = {"f": f, "v": v, "nv": nv, "nx": nx, "nu": nu, "dt": dt, "dv": dv}
        fp = step.get_collision_step(all_params={"fokker-planck": {"type": op, "solver": "batched_tridiagonal"},
                                                 "nu": nu}, stuff_for_time_loop=s)
        fo = copy.deepcopy(f)
        for _ in range(16):
            fo = fp(fo)
        return f, fo, v, dv

    f, fo, v, dv = run(0.0)
    np.testing.assert_almost_equal(f, fo, decimal=4)
    np.testing.assert_almost_equal(O.trapz_v(f * v, dv), O.trapz_v(fo * v, dv), decimal=4)
    f, fo, v, dv = run(0.5)
    np.testing.assert_almost_equal(O.trapz_v(fo * v ** 2, dv), O.trapz_v(f * v ** 2, dv), decimal=4)
    np.testing.assert_almost_equal(O.trapz_v(fo, dv), O.trapz_v(f, dv), decimal=4)
    f, fo, v, dv = run(1.5)
    if op == "lb":
        assert np.all(O.trapz_v(fo * v, dv) < O.trapz_v(f * v, dv))
    else:
        np.testing.assert_almost_equal(O.trapz_v(f * v, dv), O.trapz_v(fo * v, dv), decimal=4)


def test_device_resident_mode_matches_numpy_mode(z):
    import torch

    from vlapy_amd.core import field, vlasov

    s = stuff_from(z)
    vd, ed, fs = vlasov.get_vdfdx(s), vlasov.get_edfdv(s), field.get_field_solver(s)
    f_t = torch.as_tensor(z["f0"], device="cuda")
    e_t = torch.as_tensor(z["e0"], device="cuda")
    out_t = ed(vd(f_t, 0.1), e_t, 0.05)
    assert isinstance(out_t, torch.Tensor) and out_t.is_cuda
    out_n = ed(vd(z["f0"], 0.1), z["e0"], 0.05)
    assert np.array_equal(out_t.cpu().numpy(), out_n)
    e_out = fs(0.0, out_t)
    assert isinstance(e_out, torch.Tensor)
    assert rel(e_out.cpu().numpy(), O.field_solve(0.0, out_n, s["dv"], s["one_over_kx"])) < 1e-12


@pytest.mark.parametrize("integrator,rule", [("leapfrog", {"space": "all"}),
                                              ("leapfrog", {"time": "first-last", "space": ["k0", "k1"]}),
                                              ("pefrl", {"space": "all"}),
                                              ("h-sixth", {"time": "first-last", "space": ["k0", "k1"]})])
def test_inner_loop_matches_oracle(integrator, rule):
    """outer_loop with the HIP backend against the oracle's inner loop: each
    batch is one fused device loop (vpfp_step_n_store for leapfrog,
    vpfp_step_n_scheme for PEFRL and the 6th-order splitting)."""
    from vlapy_amd import field_driver, initializers, outer_loop

    nx, nv, k, nt = 16, 128, 0.3, 6
    ap = initializers.make_default_params_dictionary()
    ap.update({"nx": nx, "nv": nv, "nu": 5e-3, "xmax": 2 * np.pi / k, "tmax": 2.0, "nt": 21})
    ap["fokker-planck"]["type"] = "dg"
    ap["vlasov-poisson"]["time"] = integrator

    class D:
        rules_to_store_f = rule

    pulses = {"p": {"k0": k, "w0": 1.16, "t_L": 0.0, "t_R": 20.0, "t_wL": 2.5, "t_wR": 2.5, "a0": 1e-2}}
    stuff = outer_loop.get_everything_ready_for_outer_loop(D, ap, pulses, nt)
    stuff["f"] = stuff["f"] * (1 + 0.05 * np.cos(k * stuff["x"]))[:, None]
    tarr, darr = stuff["t"][:nt], stuff["driver"][:nt]
    fused = outer_loop.get_inner_loop_stepper(ap, stuff, nt)
    ts_a = fused(tarr, darr, outer_loop.get_arrays_for_inner_loop(stuff, nt, rule))
    g = O.Grid(stuff["x"], stuff["kx"], stuff["one_over_kx"], stuff["v"], stuff["kv"], stuff["dv"])
    ts_o = O.new_temp_storage(stuff["e"], stuff["f"], nt)
    ts_o = O.inner_loop(g, ts_o, tarr, darr, stuff["dt"], 5e-3, "dg", integrator, stuff["driver_function"])
    assert rel(ts_a["f"], ts_o["f"]) < 1e-11
    assert rel(ts_a["fields"]["e"], ts_o["fields"]["e"]) < 1e-10
    for key in ("mean_n", "mean_T", "mean_e2", "mean_de2", "mean_f2", "mean_cum_de2"):
        np.testing.assert_allclose(ts_a["series"][key], ts_o["series"][key], rtol=1e-10, err_msg=key)
    if rule["space"] == "all":
        assert rel(ts_a["stored_f"], ts_o["stored_f"]) < 1e-11
    else:
        want = np.fft.fft(ts_o["f"], axis=0)[:2]
        np.testing.assert_allclose(ts_a["stored_f"][-1], want.astype(np.complex64), rtol=1e-5, atol=1e-6)


def test_inner_loop_two_batches_match_oracle():
    """Two consecutive batches through the fused inner loop (its output
    arrays are page-locked buffers reused by the next batch, as the
    reference's storage step overwrites its arrays in place): after each
    batch f, E, the series, the cumulative driver energy and the k0/k1 store
    equal the oracle's inner loop run batch by batch."""
    from vlapy_amd import initializers, outer_loop

    nx, nv, k, nt = 16, 128, 0.3, 5
    ap = initializers.make_default_params_dictionary()
    ap.update({"nx": nx, "nv": nv, "nu": 5e-3, "xmax": 2 * np.pi / k, "tmax": 2.0, "nt": 21})
    ap["fokker-planck"]["type"] = "dg"
    rule = {"time": "all", "space": ["k0", "k1"]}

    class D:
        rules_to_store_f = rule

    pulses = {"p": {"k0": k, "w0": 1.16, "t_L": 0.0, "t_R": 20.0, "t_wL": 2.5, "t_wR": 2.5, "a0": 1e-2}}
    stuff = outer_loop.get_everything_ready_for_outer_loop(D, ap, pulses, 2 * nt)
    stuff["f"] = stuff["f"] * (1 + 0.05 * np.cos(k * stuff["x"]))[:, None]
    fused = outer_loop.get_inner_loop_stepper(ap, stuff, nt)
    ts_a = outer_loop.get_arrays_for_inner_loop(stuff, nt, rule)
    g = O.Grid(stuff["x"], stuff["kx"], stuff["one_over_kx"], stuff["v"], stuff["kv"], stuff["dv"])
    ts_o = O.new_temp_storage(stuff["e"], stuff["f"], nt)
    for b in range(2):
        sl = slice(b * nt, (b + 1) * nt)
        tarr, darr = stuff["t"][sl], stuff["driver"][sl]
        ts_a = fused(tarr, darr, ts_a)
        ts_o = O.inner_loop(g, ts_o, tarr, darr, stuff["dt"], 5e-3, "dg", "leapfrog", stuff["driver_function"])
        print("batch", b, flush=True)
        assert rel(ts_a["f"], ts_o["f"]) < 1e-11
        assert rel(ts_a["fields"]["e"], ts_o["fields"]["e"]) < 1e-10
        for key in ("mean_n", "mean_T", "mean_e2", "mean_de2", "mean_cum_de2"):
            np.testing.assert_allclose(ts_a["series"][key], ts_o["series"][key], rtol=1e-10, err_msg=key)
        assert ts_a["mean_cum_de2_previous"] == pytest.approx(ts_o["mean_cum_de2_previous"], rel=1e-10)
        want = np.fft.fft(ts_o["stored_f"], axis=1)[:, :2]
        np.testing.assert_allclose(ts_a["stored_f"], want.astype(np.complex64), rtol=1e-5, atol=1e-6)
