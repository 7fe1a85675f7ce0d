"""Ponderomotive driver (vlapy/field_driver.py): a sum of tanh-enveloped
travelling waves, evaluated on the host (nx values per field solve)."""

import numpy as np


def get_driver_function(x, pulse_dictionary, np=np):
    def driver_function(current_time):
        total_field = np.zeros(x.size)
        for pulse in pulse_dictionary.values():
            envelope = 0.5 * (np.tanh((current_time - pulse["t_L"]) / pulse["t_wL"])
                              - np.tanh((current_time - pulse["t_R"]) / pulse["t_wR"]))
            total_field = total_field + envelope * pulse["k0"] * pulse["a0"] * np.sin(
                pulse["k0"] * x - pulse["w0"] * current_time)
        return total_field

    return driver_function


def make_driver_array(function, x_axis, time_axis):
    out = np.zeros(time_axis.shape + x_axis.shape)
    for i in range(time_axis.size):
        out[i] = function(time_axis[i])
    return out


def get_driver_array_using_function(x, t, pulse_dictionary, this_np=np):
    return make_driver_array(get_driver_function(x, pulse_dictionary, this_np), x, t)
