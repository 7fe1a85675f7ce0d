"""CPU FP64 restatement of VlaPy's 1D-1V Vlasov-Poisson-Fokker-Planck time step.

TEST INFRASTRUCTURE ONLY.  This module is the checker for the HIP path: only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and there only to compute expected values or to
time the CPU baseline.  The product package ``vlapy_amd`` never imports it and
has no CPU fallback.

Every function cites the reference file:line whose arithmetic it restates
(paths relative to the VlaPy source tree).  The code is written from the
mathematics of those lines; it keeps the same argument meaning, array layout
((nx, nv) float64, v fastest) and the same order of floating-point operations
where that order matters for the 1e-12 parity bar (phase arguments, matrix
coefficients).

How each part is pinned (tests/test_oracle_*.py):
  * advection (vdfdx/edfdv): exact translation of band-limited data, the
    Nyquist-bin cosine law that follows from taking the real part, and a
    brute-force O(N^2) DFT evaluation on tiny grids;
  * field solve: the reference's own closed-form fields
    (tests/test_fieldsolver.py:37-54) and trapz closed forms;
  * collisions: dense ``numpy.linalg.solve`` of the same tridiagonal system
    (the reference's "naive" solver), the reference's invariant suite
    (tests/test_collisions.py:39-213), exact discrete conservation identities,
    and the notebook known-answer values (notebooks/test_fokker_planck.ipynb
    cell 5: dn = 8.55e-9, dp = -1.11e-16, dE = 3.09e-7);
  * integrators: convergence order (2, 4, 6) and the Canosa Landau roots
    (tests/extras/canosa_values.py);
  * compositions and the storage epilogue: golden vectors written by
    tests/golden/make_golden.py from the reference itself.
"""

import numpy as np

# ----------------------------------------------------------------------------
# grids and initial state
# ----------------------------------------------------------------------------


def velocity_grid(vmax, nv):
    """Cell-centred v grid and its angular wavenumbers.

    Follows vlapy/initializers.py:56-69: dv = 2 vmax / nv,
    v = linspace(-vmax + dv/2, vmax - dv/2, nv), kv = fftfreq(nv, dv) * 2 * pi.
    """
    dv = 2 * vmax / nv
    v = np.linspace(-vmax + dv / 2.0, vmax - dv / 2.0, nv)
    kv = np.fft.fftfreq(v.size, d=dv) * 2.0 * np.pi
    return dv, v, kv


def spatial_grid(xmin, xmax, nx):
    """Cell-centred periodic x grid, wavenumbers and 1/kx with 0 at DC.

    Follows vlapy/initializers.py:72-89.
    """
    dx = (xmax - xmin) / nx
    x = np.linspace(xmin + dx / 2.0, xmax - dx / 2.0, nx)
    kx = np.fft.fftfreq(x.size, d=dx) * 2.0 * np.pi
    one_over_kx = np.zeros_like(kx)
    one_over_kx[1:] = 1.0 / kx[1:]
    return dx, x, kx, one_over_kx


def trapz_v(g, dv):
    """Trapezoid rule along the last axis with spacing dv.

    Same value as ``np.trapz(g, dx=dv, axis=-1)`` used throughout the reference
    (vlapy/core/field.py:36, collisions.py:57,117,125, step.py:166-223):
    dv * (sum_j g_j - (g_0 + g_{n-1}) / 2).
    """
    g = np.asarray(g)
    return dv * (g.sum(axis=-1) - 0.5 * (g[..., 0] + g[..., -1]))


def maxwellian(nx, nv, vmax):
    """Trapz-normalised Maxwellian in every x cell (vlapy/initializers.py:29-53)."""
    dv, v, _ = velocity_grid(vmax, nv)
    row = np.exp(-(v ** 2.0) / 2.0)
    f = np.repeat(row[None, :], nx, axis=0)
    return f / trapz_v(f, dv)[:, None]


# ----------------------------------------------------------------------------
# advection pushers (vlapy/core/vlasov.py)
# ----------------------------------------------------------------------------


def vdfdx(f, dt, kx, v):
    """x-advection: f <- Re ifft_x(exp(-i kx v dt) fft_x f).

    Follows vlapy/core/vlasov.py:94-108 (step_vdfdx_exponential).  The phase
    argument is formed as (kx * dt) * v, as numpy evaluates
    ``-1j * kx[:, None] * dt * v`` left to right.
    """
    theta = (kx[:, None] * dt) * v[None, :]
    phase = np.cos(theta) - 1j * np.sin(theta)
    return np.real(np.fft.ifft(phase * np.fft.fft(f, axis=0), axis=0))


def edfdv(f, e, dt, kv):
    """v-advection: f <- Re ifft_v(exp(-i kv e dt) fft_v f).

    Follows vlapy/core/vlasov.py:123-138 (step_edfdv_exponential); the phase
    argument is (kv * dt) * e[:, None].
    """
    theta = (kv[None, :] * dt) * e[:, None]
    phase = np.cos(theta) - 1j * np.sin(theta)
    return np.real(np.fft.ifft(phase * np.fft.fft(f, axis=1), axis=1))


# ----------------------------------------------------------------------------
# field solve (vlapy/core/field.py)
# ----------------------------------------------------------------------------


def compute_charges(f, dv):
    """Electron density n(x) = trapz_v f (vlapy/core/field.py:27-36)."""
    return trapz_v(f, dv)


def solve_for_field(charge_density, one_over_kx):
    """E = Re ifft(i (1/kx) fft(1 - n)) (vlapy/core/field.py:39-63)."""
    net = 1.0 - charge_density
    return np.real(np.fft.ifft(1j * one_over_kx * np.fft.fft(net)))


def field_solve(driver_field, f, dv, one_over_kx):
    """Total field = driver + Poisson field of f (vlapy/core/field.py:66-88)."""
    return driver_field + solve_for_field(compute_charges(f, dv), one_over_kx)


# ----------------------------------------------------------------------------
# collisions (vlapy/core/collisions.py)
# ----------------------------------------------------------------------------


def collision_moments(f, v, dv, operator):
    """Per-x drift velocity and (un-normalised) temperature of the operator.

    LB (collisions.py:57-65): vbar = 0, T = trapz(f v^2).
    DG (collisions.py:117-136): vbar = trapz(f v), T = trapz(f (v - vbar)^2).
    Neither is divided by the density.
    """
    if operator == "lb":
        vbar = np.zeros(f.shape[0])
        temp = trapz_v(f * v[None, :] ** 2.0, dv)
    elif operator == "dg":
        vbar = trapz_v(f * v[None, :], dv)
        temp = trapz_v(f * (v[None, :] - vbar[:, None]) ** 2.0, dv)
    else:
        raise NotImplementedError(operator)
    return vbar, temp


def collision_diagonals(f, v, nu, dt, dv, operator):
    """Sub-, main- and super-diagonals of the implicit collision matrix.

    Follows collisions.py:67-79 (LB) and :138-156 (DG):
      a_j = nu dt (-T/dv^2 + (v_j - vbar)/(2 dv)),       j = 0..nv-2
      b   = 1 + nu dt (2 T / dv^2)
      c_j = nu dt (-T/dv^2 - (v_{j+1} - vbar)/(2 dv)),   j = 0..nv-2
    with vbar = 0 for LB.  a[j] multiplies x_j in row j+1; c[j] multiplies
    x_{j+1} in row j.
    """
    nx, nv = f.shape
    vbar, temp = collision_moments(f, v, dv, operator)
    ndt = nu * dt
    t_term = -temp[:, None] / dv ** 2.0
    a = ndt * (t_term + (v[None, :-1] - vbar[:, None]) / 2.0 / dv)
    b = 1.0 + ndt * np.ones((nx, nv)) * (2.0 * temp[:, None] / dv ** 2.0)
    c = ndt * (t_term - (v[None, 1:] - vbar[:, None]) / 2.0 / dv)
    return a, b, c


def thomas(a, b, c, d):
    """Batched Thomas algorithm without pivoting, Dirichlet-zero ends.

    Follows collisions.py:232-263 (_batched_tridiag_solver_): forward
    elimination m = a_{i-1}/b'_{i-1}, b'_i -= m c_{i-1}, d'_i -= m d'_{i-1};
    back substitution x_{n-1} = d'/b', x_i = (d'_i - c_i x_{i+1}) / b'_i.
    Inputs are not modified.
    """
    bp = np.array(b, dtype=np.float64, copy=True)
    dp = np.array(d, dtype=np.float64, copy=True)
    n = bp.shape[1]
    for i in range(1, n):
        m = a[:, i - 1] / bp[:, i - 1]
        bp[:, i] = bp[:, i] - m * c[:, i - 1]
        dp[:, i] = dp[:, i] - m * dp[:, i - 1]
    x = np.empty_like(dp)
    x[:, -1] = dp[:, -1] / bp[:, -1]
    for i in range(n - 2, -1, -1):
        x[:, i] = (dp[:, i] - c[:, i] * x[:, i + 1]) / bp[:, i]
    return x


def dense_tridiag_solve(a, b, c, d):
    """Per-x dense solve of the same system (collisions.py:163-219, "naive")."""
    out = np.empty_like(d)
    for ix in range(d.shape[0]):
        m = np.diag(a[ix], -1) + np.diag(b[ix], 0) + np.diag(c[ix], 1)
        out[ix] = np.linalg.solve(m, d[ix])
    return out


def collide(f, v, nu, dt, dv, operator="lb"):
    """One implicit Fokker-Planck step (vlapy/core/step.py:70-113).

    nu == 0 returns f unchanged; nu < 0 raises NotImplementedError.
    """
    if nu == 0.0:
        return f
    if nu < 0.0:
        raise NotImplementedError("nu < 0")
    a, b, c = collision_diagonals(f, v, nu, dt, dv, operator)
    return thomas(a, b, c, f)


# ----------------------------------------------------------------------------
# driver (vlapy/field_driver.py)
# ----------------------------------------------------------------------------


def driver_field(x, pulses, t):
    """Sum of tanh-enveloped travelling waves (vlapy/field_driver.py:24-50)."""
    total = np.zeros(x.size)
    for p in pulses.values():
        env = 0.5 * (np.tanh((t - p["t_L"]) / p["t_wL"]) - np.tanh((t - p["t_R"]) / p["t_wR"]))
        total += env * p["k0"] * p["a0"] * np.sin(p["k0"] * x - p["w0"] * t)
    return total


# ----------------------------------------------------------------------------
# time integrators (vlapy/core/vlasov_poisson.py)
# ----------------------------------------------------------------------------


class Grid:
    """Captured grid quantities (the reference's ``stuff_for_time_loop``)."""

    def __init__(self, x, kx, one_over_kx, v, kv, dv):
        self.x, self.kx, self.one_over_kx = x, kx, one_over_kx
        self.v, self.kv, self.dv = v, kv, dv

    @classmethod
    def make(cls, nx, nv, xmax, vmax, xmin=0.0):
        _, x, kx, okx = spatial_grid(xmin, xmax, nx)
        dv, v, kv = velocity_grid(vmax, nv)
        return cls(x, kx, okx, v, kv, dv)


def _field(g, driver, f):
    return field_solve(driver, f, g.dv, g.one_over_kx)


def leapfrog_step(g, e, f, t, dt, driver=lambda t: 0.0):
    """v-x-v Strang split with the field solved after the drift.

    Follows vlasov_poisson.py:53-56: kick(e_in, dt/2), drift(dt),
    E = field(driver(t + dt), f), kick(E, dt/2).  Returns (E, f).
    """
    f = edfdv(f, e, 0.5 * dt, g.kv)
    f = vdfdx(f, dt, g.kx, g.v)
    e = _field(g, driver(t + dt), f)
    f = edfdv(f, e, 0.5 * dt, g.kv)
    return e, f


PEFRL_XSI = 0.1786178958448091
PEFRL_LAMBDA = -0.2123418310626054
PEFRL_CHI = -0.6626458266981849e-1


def pefrl_step(g, e, f, t, dt, driver=lambda t: 0.0):
    """4th-order PEFRL split (vlasov_poisson.py:99-154).

    Five drifts with weights (xsi, chi, 1-2(chi+xsi), chi, xsi), each followed
    by a field solve at the cumulative drift time, and four kicks with weights
    ((1-2 lambda)/2, lambda, lambda, (1-2 lambda)/2).
    """
    xw = [PEFRL_XSI, PEFRL_CHI, 1.0 - 2.0 * (PEFRL_CHI + PEFRL_XSI), PEFRL_CHI, PEFRL_XSI]
    vw = [0.5 * (1.0 - 2.0 * PEFRL_LAMBDA), PEFRL_LAMBDA, PEFRL_LAMBDA, 0.5 * (1.0 - 2.0 * PEFRL_LAMBDA)]
    tcum = t
    for i in range(5):
        ddt = xw[i] * dt
        f = vdfdx(f, ddt, g.kx, g.v)
        tcum = tcum + ddt
        e = _field(g, driver(tcum), f)
        if i < 4:
            f = edfdv(f, e, vw[i] * dt, g.kv)
    return e, f


H6 = dict(
    a1=0.168735950563437422448196,
    a2=0.377851589220928303880766,
    a3=-0.093175079568731452657924,
    b1=0.049086460976116245491441,
    b2=0.264177609888976700200146,
    b3=0.186735929134907054308413,
    c1=-0.000069728715055305084099,
    c2=-0.000625704827430047189169,
    c3=-0.002213085124045325561636,
    d2=-2.916600457689847816445691e-6,
    d3=3.048480261700038788680723e-5,
    e3=4.985549387875068121593988e-7,
)


def hsixth_step(g, e, f, t, dt, driver=lambda t: 0.0):
    """6th-order Casas et al. split (vlasov_poisson.py:191-232).

    Kick weights D1, D2, D3, D3, D2, D1 with dt-dependent corrections and
    drift weights a1, a2, a3, a2, a1.  The driver is evaluated at
    t + a_k dt (not cumulative), as the reference does.
    """
    h = H6
    d1 = h["b1"] + 2.0 * h["c1"] * dt ** 2.0
    d2 = h["b2"] + 2.0 * h["c2"] * dt ** 2.0 + 4.0 * h["d2"] * dt ** 4.0
    d3 = h["b3"] + 2.0 * h["c3"] * dt ** 2.0 + 4.0 * h["d3"] * dt ** 4.0 - 8.0 * h["e3"] * dt ** 6.0
    kicks = [d1, d2, d3, d3, d2, d1]
    drifts = [h["a1"], h["a2"], h["a3"], h["a2"], h["a1"]]
    f = edfdv(f, e, kicks[0] * dt, g.kv)
    for i in range(5):
        f = vdfdx(f, drifts[i] * dt, g.kx, g.v)
        e = _field(g, driver(t + drifts[i] * dt), f)
        f = edfdv(f, e, kicks[i + 1] * dt, g.kv)
    return e, f


INTEGRATORS = {"leapfrog": leapfrog_step, "pefrl": pefrl_step, "h-sixth": hsixth_step}


# ----------------------------------------------------------------------------
# storage epilogue and inner loop (vlapy/core/step.py, vlapy/outer_loop.py)
# ----------------------------------------------------------------------------

FIELD_KEYS = ("e", "driver", "n", "j", "T", "q", "fv4", "vN")
SERIES_KEYS = ("mean_n", "mean_j", "mean_T", "mean_e2", "mean_de2", "mean_f2", "mean_flogf")


def step_moments(f, v, dv):
    """Per-x velocity moments stored every step (step.py:164-171)."""
    return {
        "n": trapz_v(f, dv),
        "j": trapz_v(f * v, dv),
        "T": trapz_v(f * v ** 2, dv),
        "q": trapz_v(f * v ** 3, dv),
        "fv4": trapz_v(f * v ** 4, dv),
        "vN": trapz_v(f * v ** 5, dv),
    }


def step_series(moments, e, de, f, dv):
    """Box-averaged series stored every step (step.py:200-224).

    mean_flogf is NaN as soon as any f < 0 (log of a negative number), and
    f == 0 contributes 0 * -inf = NaN, as in the reference.
    """
    with np.errstate(divide="ignore", invalid="ignore"):
        flogf = trapz_v(f * np.log(f), dv)
    return {
        "mean_n": np.mean(moments["n"]),
        "mean_j": np.mean(moments["j"]),
        "mean_T": np.mean(moments["T"]),
        "mean_e2": np.mean(e ** 2.0),
        "mean_de2": np.mean(de ** 2.0),
        "mean_f2": np.mean(trapz_v(np.real(f) ** 2.0, dv)),
        "mean_flogf": np.mean(flogf),
    }


def new_temp_storage(e, f, steps):
    """Storage dict for one batch of steps (outer_loop.py:149-215, 'all' f rule)."""
    nx = e.shape[0]
    return {
        "time_batch": np.zeros(steps),
        "e": np.array(e),
        "f": np.array(f),
        "stored_f": np.zeros((steps,) + f.shape),
        "mean_cum_de2_previous": 0.0,
        "series": {k: np.zeros(steps) for k in SERIES_KEYS},
        "fields": {k: np.zeros((steps, nx)) for k in FIELD_KEYS},
    }


def inner_loop(g, temp_storage, time_array, driver_array, dt, nu, fp_type="lb",
               integrator="leapfrog", driver=lambda t: 0.0):
    """The inner step loop and its post-update (outer_loop.py:248-281, 218-245).

    Each step: (e, f) = vp_step(e, f, t_i); f = collide(f); store f, e,
    per-x moments and series (step.py:302-326).
    """
    vp = INTEGRATORS[integrator]
    temp_storage["time_batch"] = time_array
    temp_storage["driver_array_batch"] = driver_array
    for i in range(len(time_array)):
        e, f = vp(g, temp_storage["e"], temp_storage["f"], time_array[i], dt, driver)
        f = collide(f, g.v, nu, dt, g.dv, fp_type)
        de = driver_array[i]
        temp_storage["stored_f"][i] = f
        temp_storage["e"], temp_storage["f"] = e, f
        mom = step_moments(f, g.v, g.dv)
        fields = temp_storage["fields"]
        fields["e"][i], fields["driver"][i] = e, de
        for k, val in mom.items():
            fields[k][i] = val
        ser = step_series(mom, e, de, f, g.dv)
        for k, val in ser.items():
            temp_storage["series"][k][i] = val
    s = temp_storage["series"]
    s["mean_cum_de2"] = temp_storage["mean_cum_de2_previous"] + np.cumsum(s["mean_de2"])
    s["mean_t_plus_e2_minus_cum_de2"] = s["mean_T"] + (s["mean_e2"] - s["mean_cum_de2"])
    s["mean_t_plus_e2_plus_cum_de2"] = s["mean_T"] + s["mean_e2"] + s["mean_de2"]
    temp_storage["mean_cum_de2_previous"] = s["mean_cum_de2"][-1]
    return temp_storage


# ----------------------------------------------------------------------------
# workloads named by the benchmark configurations
# ----------------------------------------------------------------------------


def perturbed_maxwellian(x, v, k, amp, harmonics=None):
    """(1 + amp cos(k x) [+ sum_m amp cos(m k x + phi_m)]) exp(-v^2/2)/sqrt(2 pi)."""
    if harmonics is None:
        dens = 1.0 + amp * np.cos(k * x)
    else:
        dens = 1.0 + sum(amp * np.cos(m * k * x + ph) for m, ph in harmonics)
    return dens[:, None] * (np.exp(-(v ** 2) / 2.0) / np.sqrt(2.0 * np.pi))[None, :]


def run_steps(g, e, f, nsteps, dt, nu=0.0, fp_type="dg", integrator="leapfrog"):
    """Plain driver-free loop returning the final state and E history."""
    vp = INTEGRATORS[integrator]
    e_hist = np.zeros((nsteps, e.shape[0]))
    for i in range(nsteps):
        e, f = vp(g, e, f, i * dt, dt)
        f = collide(f, g.v, nu, dt, g.dv, fp_type)
        e_hist[i] = e
    return e, f, e_hist


# ----------------------------------------------------------------------------
# diagnostics (vlapy/diagnostics)
# ----------------------------------------------------------------------------


def plasma_dispersion(z):
    """Z(z) = i sqrt(pi) w(z) (vlapy/diagnostics/z_function.py:6-13)."""
    from scipy import special

    return special.wofz(z) * np.sqrt(np.pi) * 1j


def epw_root(k0, wp_e=1.0, vth_e=1.0, factor=2.0, guess=None):
    """Complex EPW frequency from 1 - chi Z'(zeta) = 0 (z_function.py:25-50).

    Newton iteration on zeta = omega / (k vth sqrt(2)), with
    Z'(z) = -2 (1 + z Z(z)) and chi = (wp / (vth k))^2 / 2.
    """
    from scipy import optimize

    if guess is None:
        guess = np.sqrt(wp_e ** 2.0 + 3 * (k0 * vth_e) ** 2.0)
    chi = (wp_e / (vth_e * k0)) ** 2.0 / factor

    def eps(z):
        return 1.0 - chi * (-2.0 * (1.0 + z * plasma_dispersion(z)))

    root = optimize.newton(eps, guess)
    return root * k0 * vth_e * np.sqrt(factor)


def damping_rate_reference(e_hist, t):
    """Mean d/dt log|E_k1| over the last 75 % (low_level_helpers.py:69-87)."""
    t_ind = t.size // 4
    ek = np.fft.fft(e_hist, axis=1)
    mag = np.abs(ek[:, 1])[t_ind:]
    return np.mean(np.gradient(np.log(mag), t[2] - t[1]))


def damping_rate_and_frequency(mean_e2, t, t_min=None, t_max=None):
    """Standing-wave estimator: fit the local maxima of <E^2>(t).

    For an initial-value (standing) mode, <E^2> oscillates at 2 omega with an
    envelope exp(2 gamma t).  gamma is half the least-squares slope of
    log(peak values) against the parabola-refined peak times; omega = pi over
    the mean peak spacing.  The same estimator is applied to oracle and GPU
    histories.
    """
    y = np.asarray(mean_e2)
    sel = np.ones(y.size, dtype=bool)
    if t_min is not None:
        sel &= t >= t_min
    if t_max is not None:
        sel &= t <= t_max
    idx = np.nonzero(sel)[0]
    lo, hi = idx[0], idx[-1]
    peaks_t, peaks_y = [], []
    for i in range(max(lo, 1), min(hi, y.size - 2) + 1):
        if y[i] > y[i - 1] and y[i] >= y[i + 1]:
            # parabola through the three log samples
            l0, l1, l2 = np.log(y[i - 1]), np.log(y[i]), np.log(y[i + 1])
            den = l0 - 2 * l1 + l2
            off = 0.5 * (l0 - l2) / den if den != 0 else 0.0
            dtt = t[i + 1] - t[i]
            peaks_t.append(t[i] + off * dtt)
            peaks_y.append(l1 - 0.25 * (l0 - l2) * off)
    peaks_t, peaks_y = np.array(peaks_t), np.array(peaks_y)
    if peaks_t.size < 3:
        raise ValueError("need at least 3 peaks of <E^2> to estimate gamma and omega")
    slope = np.polyfit(peaks_t, peaks_y, 1)[0]
    gamma = 0.5 * slope
    omega = np.pi / np.mean(np.diff(peaks_t))
    return gamma, omega
