"""Adam and SGD with mentpy's update rules (`mentpy/optimizers/adam.py:12-113`,
`mentpy/optimizers/sgd.py:12-103`).

``step(f, x, i, **kwargs)`` calls ``get_gradient(f, x, **kwargs)`` as mentpy
does; when ``f`` is an ``AdjointCost`` the gradient comes from one adjoint
sweep on the GPU instead of 2n cost evaluations.
"""
from __future__ import annotations

import numpy as np

from .gradients.grad import get_gradient

__all__ = ["AdamOptimizer", "SGDOptimizer"]


class AdamOptimizer:
    def __init__(self, step_size=0.1, b1=0.9, b2=0.999, eps=10 ** -8):
        self.step_size, self.b1, self.b2, self.eps = step_size, b1, b2, eps
        self.m = None
        self.v = None

    def step(self, f, x, i, **kwargs):
        g = get_gradient(f, x, **kwargs)
        if self.m is None:
            self.m = np.zeros(len(x))
        if self.v is None:
            self.v = np.zeros(len(x))
        self.m = self.b1 * self.m + (1 - self.b1) * g
        self.v = self.b2 * self.v + (1 - self.b2) * g ** 2
        m_hat = self.m / (1 - self.b1 ** (i + 1))
        v_hat = self.v / (1 - self.b2 ** (i + 1))
        return x - self.step_size * m_hat / (np.sqrt(v_hat) + self.eps)

    def optimize(self, f, x0, num_iters=100, callback=None, verbose=False, **kwargs):
        x = x0
        for i in range(num_iters):
            x = self.step(f, x, i, **kwargs)
            if callback is not None:
                callback(x, i)
            if verbose:
                print(f"Iteration {i+1}/{num_iters} - x: {x}")
        return x

    def optimize_and_gradient_norm(self, f, x0, num_iters=100, callback=None, verbose=False, **kwargs):
        x = x0
        norm = np.zeros(num_iters)
        m = np.zeros(len(x0))
        v = np.zeros(len(x0))
        for i in range(num_iters):
            g = get_gradient(f, x, **kwargs)
            m = (1 - self.b1) * g + self.b1 * m
            v = (1 - self.b2) * g ** 2 + self.b2 * v
            x = x - self.step_size * (m / (1 - self.b1 ** (i + 1))) / (
                np.sqrt(v / (1 - self.b2 ** (i + 1))) + self.eps)
            norm[i] = np.linalg.norm(g)
            if callback is not None:
                callback(x, i)
        return x, norm

    def update_step_size(self, x, i, factor=0.99):
        self.step_size *= factor

    def reset(self):
        self.m = None
        self.v = None


class SGDOptimizer:
    def __init__(self, step_size=0.1, momentum=0.0, nesterov=False):
        self.step_size, self.momentum, self.nesterov = step_size, momentum, nesterov
        self.v = None

    def step(self, f, x, i, **kwargs):
        g = get_gradient(f, x, **kwargs)
        if self.v is None:
            self.v = np.zeros(len(x))
        self.v = self.momentum * self.v - self.step_size * g
        if self.nesterov:
            return x + self.momentum * self.v - self.step_size * g
        return x + self.v

    def optimize(self, f, x0, num_iters=100, callback=None, verbose=False, **kwargs):
        x = x0
        for i in range(num_iters):
            x = self.step(f, x, i, **kwargs)
            if callback is not None:
                callback(x, i)
            if verbose:
                print(f"Iteration {i+1}/{num_iters}")
        return x

    def optimize_and_gradient_norm(self, f, x0, num_iters=100, callback=None, verbose=False, **kwargs):
        v = np.zeros(len(x0))
        x = x0
        norm = []
        for i in range(num_iters):
            g = get_gradient(f, x, **kwargs)
            norm.append(np.linalg.norm(g))
            v = self.momentum * v - self.step_size * g
            x = x + self.momentum * v - self.step_size * g if self.nesterov else x + v
            if callback is not None:
                callback(x, i)
        return x, norm

    def update_step_size(self, x, i, factor=0.99):
        self.step_size *= factor

    def reset(self, *args, **kwargs):
        self.v = None
