"""Gradient entry point used by the optimizers (`mentpy/gradients/grad.py:13-56`).

``get_gradient(cost, x, method=...)`` keeps mentpy's methods and defaults --
``"parameter-shift"`` (shift 1.5, divided by 2*shift, as in
`mentpy/gradients/_parameter_shift.py:9-25`) and ``"finite-differences"``
(central, h = 1e-5, `_finite_difference.py:9-24`) -- which evaluate the cost
2n times, and adds ``"adjoint"``: one forward and one backward sweep on the
GPU, for costs that provide ``value_and_grad`` (see ``AdjointCost``).  When
the cost is an ``AdjointCost`` and no method is given, the adjoint is used.
"""
from __future__ import annotations

import numpy as np

__all__ = ["get_gradient", "get_hessian", "fd_gradient", "psr_gradient", "AdjointCost"]


def fd_gradient(f, x, h=1e-5, type="central"):
    if type not in ("central", "forward", "backward"):
        raise UserWarning(f"Expected type to be 'central', 'forward', or 'backward' but {type} was given")
    x = np.asarray(x, dtype=np.float64)
    grad = np.zeros(len(x))
    for i in range(len(x)):
        e = np.zeros(len(x))
        e[i] = h
        if type == "central":
            grad[i] = (f(x + e) - f(x - e)) / (2 * h)
        elif type == "forward":
            grad[i] = (f(x + e) - f(x)) / h
        else:
            grad[i] = (f(x) - f(x - e)) / h
    return grad


def psr_gradient(cost, x, shift=1.5):
    x = np.asarray(x, dtype=np.float64)
    grad = np.zeros(len(x))
    for i in range(len(x)):
        e = np.zeros(len(x))
        e[i] = shift
        grad[i] = (cost(x + e) - cost(x - e)) / (2 * shift)
    return grad


def fd_hessian(f, x, h=1e-5):
    x = np.asarray(x, dtype=np.float64)
    n = len(x)
    hess = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            ei = np.zeros(n)
            ej = np.zeros(n)
            ei[i] = h
            ej[j] = h
            hess[i, j] = (f(x + ei + ej) - f(x + ei - ej) - f(x - ei + ej) + f(x - ei - ej)) / (4 * h ** 2)
    return hess


def psr_hessian(cost, x, shift=1.5):
    return fd_hessian(cost, x, h=shift)


class AdjointCost:
    """MB-QML infidelity cost of a hip-sv simulator with an adjoint gradient.

    ``cost = AdjointCost(sim, inputs, targets)``; ``cost(x)`` is the mean
    infidelity (a float, like mentpy's tutorial cost) and
    ``cost.value_and_grad(x)`` returns ``(value, gradient)`` from one adjoint
    sweep on the GPU.  ``x`` may also be a batch ``[B, n]`` of parameter sets.
    """

    def __init__(self, sim, inputs, targets, mode: str = "auto"):
        self.sim = sim
        self.inputs = inputs
        self.targets = targets
        self.mode = mode

    def _run(self, x, grad):
        from .adjoint import mbqml_infidelity
        x = np.asarray(x, dtype=np.float64)
        loss, g = mbqml_infidelity(self.sim, x.reshape(-1, x.shape[-1]), self.inputs,
                                   self.targets, mode=self.mode, grad=grad)
        loss = loss.double().cpu().numpy()
        g = None if g is None else g.double().cpu().numpy()
        if x.ndim == 1:
            return float(loss[0]), (None if g is None else g[0])
        return loss, g

    def __call__(self, x):
        return self._run(x, False)[0]

    def value_and_grad(self, x):
        return self._run(x, True)


def get_gradient(cost, x, method=None, *args, **kwargs):
    """Gradient of ``cost`` at ``x``; default ``"parameter-shift"`` as in mentpy,
    or ``"adjoint"`` when ``cost`` is an ``AdjointCost``."""
    if method is None:
        method = "adjoint" if isinstance(cost, AdjointCost) else "parameter-shift"
    if method in ("parameter-shift", "psr", "parametershift"):
        return psr_gradient(cost, x, *args, **kwargs)
    if method in ("finite-differences", "fd", "finitedifferences"):
        return fd_gradient(cost, x, *args, **kwargs)
    if method == "adjoint":
        if not hasattr(cost, "value_and_grad"):
            raise TypeError("method='adjoint' needs a cost with value_and_grad (AdjointCost)")
        return cost.value_and_grad(x)[1]
    raise UserWarning(f"Expected method to be 'parameter-shift', 'finite-difference' or "
                      f"'adjoint' but {method} was given")


def get_hessian(cost, x, method="parameter-shift", *args, **kwargs):
    if method in ("parameter-shift", "psr", "parametershift"):
        return psr_hessian(cost, x, *args, **kwargs)
    if method in ("finite-differences", "fd", "finitedifferences"):
        return fd_hessian(cost, x, *args, **kwargs)
    raise UserWarning(f"Expected method to be 'parameter-shift' or 'finite-difference' but "
                      f"{method} was given")
