"""The allocation-free Gram loss entry point against the workspace one.

``mbqc_gram_loss_seed_c64_noalloc`` runs the kernels of
``mbqc_gram_loss_seed_c64`` (the MB-QML infidelity loss and adjoint seed from
the transfer matrices and the data Gram matrix) with its loss partials in the
library's static device array instead of a caller workspace.  Two calls in a
row (the second smaller, so stale partials of the first would show) must
equal the workspace entry point bit for bit and a complex128 evaluation to
the complex64 bar; sizes beyond the array are rejected before any device
work.
"""
import ctypes as C

import numpy as np
import pytest


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def test_noalloc_rejects_sizes_beyond_the_scratch():
    from mentpy_amd import _native as nat
    L = nat.lib()
    rc = L.mbqc_gram_loss_seed_c64_noalloc(None, None, 20000, 256, 4, None, None, None)
    assert rc == nat.MBQC_ERR_VALUE
    assert "static scratch" in L.mbqc_last_error().decode()


@pytest.mark.gpu
def test_noalloc_gram_loss_equals_workspace_entry_point():
    torch = _torch()
    from mentpy_amd import _native as nat
    L = nat.lib()
    rng = np.random.default_rng(77)
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    for B, K, S in ((4000, 256, 64), (37, 256, 64), (300, 48, 7), (16384, 256, 256)):
        t = rng.normal(size=(B, K)) + 1j * rng.normal(size=(B, K))
        t /= np.linalg.norm(t, axis=1, keepdims=True)
        D = rng.normal(size=(S, K)) + 1j * rng.normal(size=(S, K))
        M = D.conj().T @ D / K
        td = torch.as_tensor(t, dtype=torch.complex64, device="cuda")
        MTd = torch.as_tensor(M.T.copy(), dtype=torch.complex64, device="cuda")
        ws = torch.empty(int(L.mbqc_gram_workspace_bytes(B, K)) // 4, dtype=torch.float32,
                         device="cuda")
        outs = []
        for fn, extra in ((L.mbqc_gram_loss_seed_c64, (C.c_void_p(ws.data_ptr()),)),
                          (L.mbqc_gram_loss_seed_c64_noalloc, ())):
            loss = torch.full((B,), float("nan"), dtype=torch.float32, device="cuda")
            seed = torch.full((B, K), float("nan"), dtype=torch.complex64, device="cuda")
            nat.check(fn(C.c_void_p(td.data_ptr()), C.c_void_p(MTd.data_ptr()), B, K, S,
                         C.c_void_p(loss.data_ptr()), C.c_void_p(seed.data_ptr()), *extra,
                         stream))
            torch.cuda.synchronize()
            outs.append((loss.cpu().numpy(), seed.cpu().numpy()))
        (l0, s0), (l1, s1) = outs
        assert np.array_equal(l0, l1) and np.array_equal(s0, s1), (B, K, S)
        t32 = t.astype(np.complex64).astype(np.complex128)
        M32 = M.astype(np.complex64).astype(np.complex128)
        z = t32 @ M32.T
        want_loss = 1 - np.real(np.sum(t32.conj() * z, 1)) / S
        want_seed = -2.0 / S * z
        err_l = np.abs(l1 - want_loss).max() / max(1.0, np.abs(want_loss).max())
        err_s = np.abs(s1 - want_seed).max() / np.abs(want_seed).max()
        print(f"noalloc B={B} K={K}: loss err {err_l:.1e}, seed err {err_s:.1e}", flush=True)
        assert err_l < 1e-5 and err_s < 1e-5
