"""GPU parity of the adjoint angle gradient for live windows of 6..16 slots
(one workgroup per simulation in the forward pass, one per partial-sum row in
the backward pass; states in LDS up to 10 slots, in a stream-ordered global
workspace above).

References: the exact shift rule (s = pi/2) of the complex128 oracle cost,
exact for these costs (a + b cos t + c sin t in every angle), and the
register-resident adjoint of the same pattern on its minimal window.
Tolerance: 2e-4 absolute on gradient entries of order 1e-2..1, 2e-5 on the
loss (complex64 sweeps over up to 30 steps of 2^10 amplitudes).
"""
import numpy as np
import pytest

from conftest import haar_state

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _dataset(rows, S, rng):
    from mentpy_amd.operators.gates import haar_unitary
    X = np.stack([haar_state(rows, rng) for _ in range(S)])
    if rows <= 8:
        return X, X @ haar_unitary(2 ** rows, rng).T
    return X, np.stack([haar_state(rows, rng) for _ in range(S)])


def _oracle_cost(g, X, Y):
    from oracle.vectorized import sv_batch

    def cost(th):
        out = sv_batch(g, np.asarray(th)[None], X)[0]
        f = np.sum(np.conj(Y) * out, -1)
        return 1.0 - np.mean(np.abs(f) ** 2)
    return cost


@pytest.mark.parametrize("rows,cols,S,mode,B", [
    (6, 3, 8, "direct", 3), (6, 3, 64, "transfer", 2), (7, 3, 5, "direct", 2),
    (7, 3, 130, "transfer", 2), (8, 3, 6, "direct", 2), (9, 3, 3, "direct", 2),
    (10, 3, 4, "direct", 2), (11, 3, 3, "direct", 2), (12, 3, 2, "direct", 1),
    (16, 3, 1, "direct", 1)])
def test_lds_adjoint_matches_exact_shift(rows, cols, S, mode, B, rng):
    """Column counts 3..128 cover the three partial-row layouts (one row of
    C <= 64 columns, rows of 64 columns, one row per column)."""
    _torch()
    import mentpy_amd as mp
    from mentpy_amd.gradients import MBQMLStep
    from oracle.gradients import exact_shift_gradient
    g = mp.templates.grid_cluster(rows, cols)
    sim = mp.PatternSimulator(g, backend="hip-sv")
    assert sim.simulator.plan.info["n_slots"] == rows
    X, Y = _dataset(rows, S, rng)
    th = rng.uniform(0, 2 * np.pi, (B, len(g.trainable_nodes))).astype(np.float32)
    step = MBQMLStep(sim, X, Y, n_sets=B, mode=mode)
    assert step.segments == 1
    loss, grad = step.run(th)
    loss, grad = loss.cpu().numpy(), grad.cpu().numpy()
    cost = _oracle_cost(g, X, Y)
    for b in range(B):
        ref_g = exact_shift_gradient(cost, th[b].astype(np.float64))
        ref_l = cost(th[b].astype(np.float64))
        err = np.abs(grad[b] - ref_g).max()
        print(f"{rows}x{cols} S={S} {mode} b={b}: loss err {abs(loss[b] - ref_l):.2e} "
              f"grad err {err:.2e}", flush=True)
        assert abs(loss[b] - ref_l) < 2e-5
        assert err < 2e-4


@pytest.mark.parametrize("rows,cols,S,B", [(6, 4, 128, 512), (6, 5, 64, 1024)])
def test_lds_adjoint_checkpoint_pairs(rows, cols, S, B, rng):
    """From 65,536 columns one checkpoint is kept per two blocks (3 blocks on
    the 6x4 grid, 4 on the 6x5 grid): the gradient still matches the exact
    shift rule."""
    _torch()
    import mentpy_amd as mp
    from mentpy_amd.gradients import MBQMLStep
    from oracle.gradients import exact_shift_gradient
    g = mp.templates.grid_cluster(rows, cols)
    sim = mp.PatternSimulator(g, backend="hip-sv")
    X, Y = _dataset(rows, S, rng)
    th = rng.uniform(0, 2 * np.pi, (B, len(g.trainable_nodes))).astype(np.float32)
    step = MBQMLStep(sim, X, Y, n_sets=B, mode="direct")
    loss, grad = step.run(th)
    cost = _oracle_cost(g, X, Y)
    for b in (0, B - 1):
        ref = exact_shift_gradient(cost, th[b].astype(np.float64))
        err = np.abs(grad[b].cpu().numpy() - ref).max()
        print(f"{rows}x{cols} B={B} set {b}: grad err {err:.2e}", flush=True)
        assert err < 2e-4
        assert abs(loss[b].item() - cost(th[b].astype(np.float64))) < 2e-5


@pytest.mark.parametrize("rows,cols,w,S,mode", [(5, 4, 7, 8, "direct"), (5, 5, 8, 40, "transfer"),
                                                (4, 6, 8, 6, "direct")])
def test_lds_adjoint_wide_window_matches_register_path(rows, cols, w, S, mode, rng):
    """A grid compiled on a window above its minimum (6 or 7 slots, pair
    coefficients from the tables) runs the LDS sweeps; its loss and gradient
    equal those of the register sweeps on the minimal window."""
    _torch()
    import mentpy_amd as mp
    from mentpy_amd.gradients import MBQMLStep
    g = mp.templates.grid_cluster(rows, cols)
    wide = mp.PatternSimulator(g, backend="hip-sv", window_size=w)
    assert wide.simulator.plan.info["n_slots"] == w - 1
    base = mp.PatternSimulator(g, backend="hip-sv")
    X, Y = _dataset(rows, S, rng)
    B = 4
    th = rng.uniform(0, 2 * np.pi, (B, len(g.trainable_nodes))).astype(np.float32)
    l1, g1 = MBQMLStep(wide, X, Y, n_sets=B, mode=mode).run(th)
    l1, g1 = l1.clone(), g1.clone()
    l2, g2 = MBQMLStep(base, X, Y, n_sets=B, mode=mode).run(th)
    lerr = (l1 - l2).abs().max().item()
    gerr = (g1 - g2).abs().max().item()
    print(f"{rows}x{cols} w={w} {mode}: loss diff {lerr:.2e} grad diff {gerr:.2e}", flush=True)
    assert lerr < 2e-6 and gerr < 2e-5


def test_lds_adjoint_refusals(rng):
    """Time segments stay with the register sweeps (at most 5 slots); windows
    above 16 slots have no adjoint."""
    _torch()
    import mentpy_amd as mp
    from mentpy_amd import _native as nat
    from mentpy_amd.gradients import MBQMLStep
    g = mp.templates.grid_cluster(6, 4)
    sim = mp.PatternSimulator(g, backend="hip-sv")
    assert nat.lib().mbqc_sv_segments_auto(sim.simulator.plan.handle, 4) == 1
    X, Y = _dataset(6, 80, rng)
    with pytest.raises((NotImplementedError, ValueError)):
        MBQMLStep(sim, X, Y, n_sets=4, mode="transfer", segments=2).run(
            np.zeros((4, len(g.trainable_nodes)), np.float32))
    g17 = mp.templates.grid_cluster(17, 3)
    sim17 = mp.PatternSimulator(g17, backend="hip-sv")
    X, Y = _dataset(17, 1, rng)
    with pytest.raises(NotImplementedError):
        MBQMLStep(sim17, X, Y, n_sets=1, mode="direct").run(
            np.zeros((1, len(g17.trainable_nodes)), np.float32))
