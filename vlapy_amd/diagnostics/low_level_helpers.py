"""Estimators on the stored field history (vlapy/diagnostics/low_level_helpers.py).

The reference functions take an xarray DataArray with a "time" coordinate;
these accept that or a plain (nt, nx) array plus the time axis `t`.
"""

import numpy as np


def _unpack(efield_arr, t=None):
    if t is not None:
        return np.asarray(efield_arr), np.asarray(t)
    return np.asarray(efield_arr.data), np.asarray(efield_arr.coords["time"].data)


def get_nth_mode(efield_arr, mode_number, t=None):
    """Real-space reconstruction of one Fourier mode (low_level_helpers.py:27-40)."""
    e, _ = _unpack(efield_arr, t) if t is not None else (np.asarray(getattr(efield_arr, "data", efield_arr)), None)
    ek = np.fft.fft(e, axis=1, norm="ortho")
    rec = np.zeros(e.shape, dtype=np.complex128)
    rec[:, mode_number] = ek[:, mode_number]
    return 2 * np.fft.ifft(rec, axis=1, norm="ortho")


def get_w_ax(efield_arr, t=None):
    _, tax = _unpack(efield_arr, t)
    return np.fft.fftfreq(tax.size, d=tax[1] - tax[0]) * 2 * np.pi


def get_e_ss(efield_arr, t=None):
    """Amplitude and phase of mode 1 at the last time (low_level_helpers.py:56-66)."""
    e, _ = _unpack(efield_arr, t) if t is not None else (np.asarray(getattr(efield_arr, "data", efield_arr)), None)
    ek1 = 2 * np.fft.fft(e, axis=1, norm="ortho")[-1, 1]
    return np.abs(ek1), np.angle(ek1)


def get_damping_rate(efield_arr, t=None):
    """Mean d/dt log|E_k1| over the last 75 % of the history (low_level_helpers.py:69-87).
    Meant for a driven travelling mode."""
    e, tax = _unpack(efield_arr, t)
    t_ind = tax.size // 4
    mag = np.abs(np.fft.fft(e, axis=1)[:, 1])[t_ind:]
    return np.mean(np.gradient(np.log(mag), tax[2] - tax[1]))


def get_oscillation_frequency(efield_arr, t=None):
    """Argmax of |E(k1, omega)| over the last 75 % (low_level_helpers.py:147-163)."""
    e, tax = _unpack(efield_arr, t)
    t_ind = tax.size // 4
    ek1w = np.abs(np.fft.fft2(e[t_ind:, :])[:, 1])
    wax = np.fft.fftfreq(tax.size, d=tax[1] - tax[0]) * 2 * np.pi
    return wax[ek1w.argmax()]


def get_standing_wave_rate_and_frequency(mean_e2, t, t_min=None, t_max=None):
    """Damping rate and real frequency of an initial-value (standing) mode.

    <E^2>(t) of a standing wave oscillates at 2 omega under an envelope
    exp(2 gamma t): gamma is half the least-squares slope of the log of the
    local maxima (times refined by a parabola through the three log samples
    around each maximum) and omega is pi over their mean spacing.
    """
    y = np.asarray(mean_e2, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    keep = np.ones(y.size, dtype=bool)
    if t_min is not None:
        keep &= t >= t_min
    if t_max is not None:
        keep &= t <= t_max
    idx = np.nonzero(keep)[0]
    if idx.size < 3:
        raise ValueError("the window [t_min, t_max] holds fewer than three samples")
    tp, lp = [], []
    for i in range(max(idx[0], 1), min(idx[-1], y.size - 2) + 1):
        if y[i] > y[i - 1] and y[i] >= y[i + 1]:
            l0, l1, l2 = np.log(y[i - 1]), np.log(y[i]), np.log(y[i + 1])
            den = l0 - 2.0 * l1 + l2
            off = 0.5 * (l0 - l2) / den if den != 0 else 0.0
            tp.append(t[i] + off * (t[i + 1] - t[i]))
            lp.append(l1 - 0.25 * (l0 - l2) * off)
    tp, lp = np.asarray(tp), np.asarray(lp)
    if tp.size < 3:
        raise ValueError("need at least three maxima of <E^2> to estimate gamma and omega")
    gamma = 0.5 * np.polyfit(tp, lp, 1)[0]
    omega = np.pi / np.mean(np.diff(tp))
    return gamma, omega
