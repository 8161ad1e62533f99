"""Adam and SGD with mentpy's update rules (`mentpy/optimizers/adam.py:12-113`,
`mentpy/optimizers/sgd.py:12-103`).

``step(f, x, i, **kwargs)`` calls ``get_gradient(f, x, **kwargs)`` as mentpy
does; when ``f`` is an ``AdjointCost`` the gradient comes from one adjoint
sweep on the GPU instead of 2n cost evaluations.
"""
from __future__ import annotations

import numpy as np

from .gradients.grad import get_gradient

__all__ = ["AdamOptimizer", "SGDOptimizer", "DeviceAdam", "DeviceSGD", "MBQMLTrainer"]


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


# ------------------------------------------------------------ on the device
class DeviceAdam:
    """Adam on a device tensor of parameters (any shape, float32), updated in
    place by one HIP kernel per step with the same formulas as ``AdamOptimizer``."""

    def __init__(self, params, step_size=0.1, b1=0.9, b2=0.999, eps=10 ** -8):
        import torch
        if params.dtype != torch.float32 or not params.is_contiguous():
            raise TypeError("params must be a contiguous float32 device tensor")
        self.params = params
        self.step_size, self.b1, self.b2, self.eps = step_size, b1, b2, eps
        self.m = torch.zeros_like(params)
        self.v = torch.zeros_like(params)
        self.t = torch.zeros(1, dtype=torch.int32, device=params.device)

    def step(self, grad):
        import ctypes as C
        import torch
        from . import _native as nat
        g = grad.to(torch.float32).contiguous()
        stream = C.c_void_p(torch.cuda.current_stream(self.params.device).cuda_stream)
        nat.check(nat.lib().mbqc_adam_step_f32(
            C.c_void_p(self.params.data_ptr()), C.c_void_p(g.data_ptr()), C.c_void_p(self.m.data_ptr()),
            C.c_void_p(self.v.data_ptr()), self.params.numel(), self.step_size, self.b1, self.b2,
            self.eps, C.c_void_p(self.t.data_ptr()), stream))


class DeviceSGD:
    """SGD with momentum (optionally Nesterov) on a device tensor, as ``SGDOptimizer``."""

    def __init__(self, params, step_size=0.1, momentum=0.0, nesterov=False):
        import torch
        if params.dtype != torch.float32 or not params.is_contiguous():
            raise TypeError("params must be a contiguous float32 device tensor")
        self.params = params
        self.step_size, self.momentum, self.nesterov = step_size, momentum, nesterov
        self.v = torch.zeros_like(params)

    def step(self, grad):
        import ctypes as C
        import torch
        from . import _native as nat
        g = grad.to(torch.float32).contiguous()
        stream = C.c_void_p(torch.cuda.current_stream(self.params.device).cuda_stream)
        nat.check(nat.lib().mbqc_sgd_step_f32(
            C.c_void_p(self.params.data_ptr()), C.c_void_p(g.data_ptr()), C.c_void_p(self.v.data_ptr()),
            self.params.numel(), self.step_size, self.momentum, int(self.nesterov), stream))


class MBQMLTrainer:
    """MB-QML training of ``n_sets`` independent parameter sets entirely on the
    GPU: each iteration is one adjoint loss-and-gradient step (``MBQMLStep``)
    and one device Adam/SGD update, captured together in a HIP graph.

    ``trainer.run(num_iters)`` returns the loss history ``[num_iters, n_sets]``;
    ``trainer.params`` holds the parameters (``[n_sets, n_angles]``).
    """

    def __init__(self, sim, inputs, targets, init_params, optimizer: str = "adam", mode: str = "auto",
                 **opt_kwargs):
        import torch
        from .gradients.adjoint import MBQMLStep
        p0 = torch.as_tensor(init_params, dtype=torch.float32)
        if p0.dim() == 1:
            p0 = p0.unsqueeze(0)
        self.step_obj = MBQMLStep(sim, inputs, targets, n_sets=p0.shape[0], mode=mode)
        self.params = self.step_obj.angles
        self.params.copy_(p0.to(self.params.device))
        if optimizer == "adam":
            self.opt = DeviceAdam(self.params, **opt_kwargs)
        elif optimizer == "sgd":
            self.opt = DeviceSGD(self.params, **opt_kwargs)
        else:
            raise ValueError("optimizer must be 'adam' or 'sgd'")
        self._graph = None

    def _iteration(self):
        loss, g = self.step_obj.run()
        self.opt.step(g)
        return loss

    def capture(self):
        """Capture one iteration in a HIP graph (the parameters advance by one
        optimizer step during capture warm-up; use a fresh trainer to time).
        Live windows of 11..16 qubits allocate the adjoint's workspace on every
        call, which graph capture refuses (NotImplementedError): run those
        without a graph."""
        import torch
        self._iteration()
        torch.cuda.synchronize(self.params.device)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._iteration()
        torch.cuda.synchronize(self.params.device)
        self._graph = g
        return g

    def run(self, num_iters: int, graph: bool = True):
        import torch
        hist = torch.empty((num_iters, self.params.shape[0]), dtype=torch.float32,
                           device=self.params.device)
        for i in range(num_iters):
            if graph and self._graph is not None:
                self._graph.replay()
            else:
                self._iteration()
            hist[i].copy_(self.step_obj.loss)
        return hist
