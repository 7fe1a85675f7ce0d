"""Grids, initial state and parameter dictionaries (vlapy/initializers.py),
without the mlflow parameter logging."""

import numpy as np

from .diagnostics import z_function


def initialize_distribution(nx, nv, vmax=6.0):
    """Trapz-normalised Maxwellian in every x cell (initializers.py:29-53)."""
    dv = 2.0 * vmax / nv
    vax = np.linspace(-vmax + dv / 2.0, vmax - dv / 2.0, nv)
    f = np.repeat(np.exp(-(vax ** 2.0) / 2.0)[None, :], nx, axis=0)
    norm = dv * (f.sum(axis=1) - 0.5 * (f[:, 0] + f[:, -1]))
    return f / norm[:, None]


def initialize_velocity_quantities(vmax, nv):
    """dv, cell-centred v and kv = fftfreq * 2 pi (initializers.py:56-69)."""
    dv = 2 * vmax / nv
    v = np.linspace(-vmax + dv / 2.0, vmax - dv / 2.0, nv)
    kv = np.fft.fftfreq(v.size, d=dv) * 2.0 * np.pi
    return dv, v, kv


def initialize_spatial_quantities(xmin, xmax, nx):
    """dx, cell-centred x, kx and 1/kx with 0 at DC (initializers.py:72-89)."""
    dx = (xmax - xmin) / nx
    x = np.linspace(xmin + dx / 2.0, xmax - dx / 2.0, nx)
    kx = np.fft.fftfreq(x.size, d=dx) * 2.0 * np.pi
    one_over_kx = np.zeros_like(kx)
    one_over_kx[1:] = 1.0 / kx[1:]
    return dx, x, kx, one_over_kx


def make_default_params_dictionary():
    """Default parameters (initializers.py:118-149) with the HIP backend."""
    return {
        "nx": 32,
        "xmin": 0.0,
        "nv": 512,
        "vmax": 6.4,
        "nt": 500,
        "tmax": 80,
        "fokker-planck": {"type": "lb", "solver": "batched_tridiagonal"},
        "vlasov-poisson": {"time": "leapfrog", "vdfdx": "exponential", "edfdv": "exponential", "poisson": "spectral"},
        "backend": {"core": "hip", "max_GB_for_device": int(1), "precision": "fp64"},
        "a0": 4e-7,
    }


def specify_collisions_to_dict(log_nu_over_nu_ld, all_params_dict):
    """nu = |nu_ld| 10^log_nu_over_nu_ld, or 0 (initializers.py:152-167)."""
    if log_nu_over_nu_ld is None:
        all_params_dict["nu"] = 0.0
    else:
        all_params_dict["nu"] = np.abs(all_params_dict["nu_ld"]) * 10 ** log_nu_over_nu_ld
    return all_params_dict


def specify_epw_params_to_dict(k0, all_params_dict):
    """EPW frequency and Landau rate from the dispersion relation; box = 2 pi/k0
    (initializers.py:170-190)."""
    root = z_function.get_roots_to_electrostatic_dispersion(wp_e=1.0, vth_e=1.0, k0=k0)
    all_params_dict["w_epw"] = np.real(root)
    all_params_dict["nu_ld"] = np.imag(root)
    all_params_dict["xmax"] = 2.0 * np.pi / k0
    all_params_dict["v_ph"] = all_params_dict["w_epw"] / k0
    return all_params_dict
