"""mentpy's black-box gradients and the exact shift rule (float64).

``fd_gradient`` follows `mentpy/gradients/_finite_difference.py:9-24`
(central by default, h = 1e-5); ``psr_gradient`` follows
`mentpy/gradients/_parameter_shift.py:9-25` (shift 1.5, divided by 2*shift);
``exact_shift_gradient`` is ``[f(x+s e_i) - f(x-s e_i)] / (2 sin s)``, exact for
costs of the form ``a + b cos(x_i) + c sin(x_i)`` in every coordinate, which
every post-selected fidelity loss is (each amplitude is affine in e^{-i x_i}).
"""
from __future__ import annotations

import numpy as np


def fd_gradient(f, x, h=1e-5, type="central"):
    x = np.asarray(x, dtype=np.float64)
    g = np.zeros(len(x))
    for i in range(len(x)):
        e = np.eye(len(x))[i]
        if type == "central":
            g[i] = (f(x + h * e) - f(x - h * e)) / (2 * h)
        elif type == "forward":
            g[i] = (f(x + h * e) - f(x)) / h
        else:
            g[i] = (f(x) - f(x - h * e)) / h
    return g


def psr_gradient(cost, x, shift=1.5):
    x = np.asarray(x, dtype=np.float64)
    g = np.zeros(len(x))
    for i in range(len(x)):
        e = np.eye(len(x))[i]
        g[i] = (cost(x + shift * e) - cost(x - shift * e)) / (2 * shift)
    return g


def exact_shift_gradient(cost, x, shift=np.pi / 2):
    x = np.asarray(x, dtype=np.float64)
    g = np.zeros(len(x))
    for i in range(len(x)):
        e = np.eye(len(x))[i]
        g[i] = (cost(x + shift * e) - cost(x - shift * e)) / (2 * np.sin(shift))
    return g
