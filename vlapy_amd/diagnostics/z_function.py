"""Plasma dispersion function and the EPW root (vlapy/diagnostics/z_function.py)."""

import numpy as np
from scipy import optimize, special


def plasma_dispersion(value):
    """Z(z) = i sqrt(pi) w(z), w the Faddeeva function."""
    return special.wofz(value) * np.sqrt(np.pi) * 1j


def plasma_dispersion_prime(value):
    """Z'(z) = -2 (1 + z Z(z))."""
    return -2.0 * (1.0 + value * plasma_dispersion(value))


def get_roots_to_electrostatic_dispersion(wp_e, vth_e, k0, maxwellian_convention_factor=2.0,
                                          initial_root_guess=None):
    """Newton root of 1 - chi Z'(zeta) = 0, returned as omega = zeta k vth sqrt(factor)."""
    if initial_root_guess is None:
        initial_root_guess = np.sqrt(wp_e ** 2.0 + 3 * (k0 * vth_e) ** 2.0)
    chi_e = np.power(wp_e / (vth_e * k0), 2.0) / maxwellian_convention_factor

    def eps(z):
        return 1.0 - chi_e * plasma_dispersion_prime(z)

    root = optimize.newton(eps, initial_root_guess)
    return root * k0 * vth_e * np.sqrt(maxwellian_convention_factor)
