"""Chemical-potential schedules for simulated annealing (chromo/util/mu_schedules.py).

A schedule maps the save point i of N to the factor ``mu_adjust_factor`` that
``polymer_in_field`` hands to ``mc_sim`` for that save point; the binding dE
scales a negative chemical potential by f and a positive one by 2 - f
(chromo/polymers.pyx:1522-1537).  ``Schedule`` wraps a function as the
reference does (``schedule.function(i, N)``, ``schedule.name``, ``to_file``).
The schedule formulas are the reference's, so that an annealing run follows the
same factors.
"""
from __future__ import annotations

import numpy as np


class Schedule:
    """A named schedule: ``function(i, N)`` gives the factor at save point i of N."""

    def __init__(self, fxn):
        self.function = fxn
        self.name = fxn.__name__

    def to_file(self, path):
        with open(path, "w") as f:
            f.write(self.name)


def _delayed_linear(lo, hi, frac):
    """lo until save point floor(N frac), then linear to hi at N."""
    def f(i, N):
        delay = np.floor(N * frac)
        if i <= delay:
            return lo
        return (hi - lo) * (float(i - delay) / float(N - delay)) + lo
    return f


def _delayed_tanh(frac):
    def f(i, N):
        delay = np.floor(N * frac)
        if i <= delay:
            return np.tanh(-1)
        return np.tanh(3 * float(i - delay) / float(N - delay) - 1)
    return f


def _stepped_linear(start_value, end_value=1.0, frac=0.2):
    """start_value for the first fifth of the run, end_value for the last fifth,
    linear between."""
    def f(i, N):
        start = np.floor(N * frac)
        end = np.floor(N * frac)
        if i <= start:
            return float(start_value)
        if i >= (N - end):
            return float(end_value)
        return (end_value - start_value) * (float(i - start) / float(N - (start + end))) + start_value
    return f


def linear_1(i, N):
    """Linear from -1 to 1."""
    return 2 * (i / N) - 1


def linear_2(i, N):
    """-1 for the first quarter, then linear to 1."""
    return _delayed_linear(-1, 1, 1 / 4)(i, N)


def linear_3(i, N):
    """-1 for the first sixth, then linear to 1."""
    return _delayed_linear(-1, 1, 1 / 6)(i, N)


def linear_4(i, N):
    """-1 for the first eighth, then linear to 1."""
    return _delayed_linear(-1, 1, 1 / 8)(i, N)


def tanh_1(i, N):
    """tanh(5 i / N - 3): from about -1 to 1."""
    return np.tanh(5 * float(i) / N - 3)


def tanh_2(i, N):
    """tanh(3 i / N - 1): from about -0.76 to 0.96."""
    return np.tanh(3 * float(i) / N - 1)


def tanh_3(i, N):
    """tanh(-1) for the first sixth, then tanh_2 over the rest."""
    return _delayed_tanh(1 / 6)(i, N)


def tanh_4(i, N):
    """tanh(-1) for the first eighth, then tanh_2 over the rest."""
    return _delayed_tanh(1 / 8)(i, N)


def log_1(i, N):
    return np.log(i / (N / 6.25) + 1) - 1


def log_2(i, N):
    return np.log(i / (N / 4) + 0.4) - 0.5


def log_3(i, N):
    return np.log(i / (N / 2.5) + 0.5) - 0.098


def sawtooth_1(i, N):
    """Logarithmic rise with a drop of 0.15 every 20 save points; 1 at the end."""
    steps = np.floor(i / 20)
    if (N - i) > 2:
        return np.log(i / (N / 10) + 0.5) - 0.04 - 0.15 * steps
    return 1


def linear_2_for_negative_cp(i, N):
    """4 for the first quarter, then linear to 1 (for negative chemical potentials)."""
    return _delayed_linear(4.0, 1.0, 1 / 4)(i, N)


def linear_step_for_negative_cp(i, N):
    """4, linear to 1 over the middle three fifths, then 1."""
    return _stepped_linear(4.0)(i, N)


def linear_step_for_negative_cp_mild(i, N):
    """2, linear to 1 over the middle three fifths, then 1."""
    return _stepped_linear(2.0)(i, N)
