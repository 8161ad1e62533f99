"""GPU parity of the adjoint angle gradient.

References: the exact shift rule (s = pi/2) of the complex128 oracle cost
(exact for these costs, which are a + b cos t + c sin t in every angle), the
GPU central finite-difference loop (h = 1e-3), and the oracle's central
finite differences (h = 1e-5, mentpy's default) on the 4x16 grid.
Tolerance: 2e-4 absolute on gradient entries of order 1e-2..1 (complex64
forward/backward over up to 60 steps; FD in complex64 with h = 1e-3 carries
~1e-4 of cancellation error).
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
    U = haar_unitary(2 ** rows, rng)
    X = np.stack([haar_state(rows, rng) for _ in range(S)])
    return X, X @ U.T


def _oracle_cost(g, X, Y):
    from oracle.vectorized import sv_batch

    def cost(th):
        out = sv_batch(g, np.asarray(th)[None], X)[0]
        f = np.sum(np.conj(Y) * out, -1)
        return 1.0 - np.mean(np.abs(f) ** 2)
    return cost


@pytest.mark.parametrize("rows,cols,S,mode", [
    (1, 5, 3, "direct"), (2, 4, 3, "direct"), (2, 4, 9, "transfer"), (3, 5, 16, "transfer"),
    (3, 5, 4, "direct"), (4, 6, 20, "transfer"), (5, 4, 8, "direct"), (2, 4, 9, "gram"),
    (3, 5, 40, "gram"), (4, 6, 300, "gram")])
def test_adjoint_matches_exact_shift(rows, cols, S, mode, rng):
    torch = _torch()
    import mentpy_amd as mp
    from mentpy_amd.gradients import mbqml_infidelity
    from oracle.gradients import exact_shift_gradient
    g = mp.templates.grid_cluster(rows, cols) if rows > 1 else mp.templates.linear_cluster(cols)
    sim = mp.PatternSimulator(g, backend="hip-sv")
    X, Y = _dataset(rows, S, rng)
    B = 3
    th = rng.uniform(0, 2 * np.pi, (B, len(g.trainable_nodes))).astype(np.float32)
    loss, grad = mbqml_infidelity(sim, th, X, Y, mode=mode)
    loss, grad = loss.cpu().numpy(), grad.cpu().numpy()
    for b in range(B):
        cost = _oracle_cost(g, X, Y)
        ref_g = exact_shift_gradient(cost, th[b].astype(np.float64))
        ref_l = cost(th[b].astype(np.float64))
        err = np.abs(grad[b] - ref_g).max()
        print(f"{rows}x{cols} S={S} {mode} b={b}: loss err {abs(loss[b]-ref_l):.2e} grad err {err:.2e}",
              flush=True)
        assert abs(loss[b] - ref_l) < 2e-5
        assert err < 2e-4


@pytest.mark.parametrize("fd_mode", ["direct", "gram"])
def test_adjoint_matches_gpu_fd_loop(fd_mode, rng):
    """Both GPU finite-difference loops (inputs propagated with the fused
    overlap loss; basis columns with the Gram loss) agree with the adjoint."""
    torch = _torch()
    import mentpy_amd as mp
    from mentpy_amd.gradients import fd_loop_infidelity_grad, mbqml_infidelity
    g = mp.templates.grid_cluster(3, 6)
    sim = mp.PatternSimulator(g, backend="hip-sv")
    X, Y = _dataset(3, 24, rng)
    th = rng.uniform(0, 2 * np.pi, (8, len(g.trainable_nodes))).astype(np.float32)
    _, ga = mbqml_infidelity(sim, th, X, Y)
    gf = fd_loop_infidelity_grad(sim, th, X, Y, h=1e-3, mode=fd_mode)
    err = (ga - gf).abs().max().item()
    print(f"adjoint vs GPU FD ({fd_mode}, h=1e-3): {err:.2e}", flush=True)
    assert err < 5e-3


def test_overlap_loss_kernel_matches_oracle(rng):
    """The fused loss/seed kernel of the direct mode, for output sizes below,
    at and above one wave per row, against a complex128 evaluation."""
    torch = _torch()
    import ctypes as C
    from mentpy_amd import _native as nat
    for B, S, O in ((3, 5, 1), (4, 37, 16), (2, 300, 64), (2, 9, 128)):
        out = rng.normal(size=(B, S, O)) + 1j * rng.normal(size=(B, S, O))
        Y = rng.normal(size=(S, O)) + 1j * rng.normal(size=(S, O))
        od = torch.as_tensor(out, dtype=torch.complex64, device="cuda")
        yd = torch.as_tensor(Y, dtype=torch.complex64, device="cuda")
        loss = torch.empty(B, dtype=torch.float32, device="cuda")
        seed = torch.empty_like(od)
        nat.check(nat.lib().mbqc_overlap_loss_seed_c64(
            C.c_void_p(od.data_ptr()), C.c_void_p(yd.data_ptr()), B, S, O,
            C.c_void_p(loss.data_ptr()), C.c_void_p(seed.data_ptr()),
            C.c_void_p(torch.cuda.current_stream().cuda_stream)))
        torch.cuda.synchronize()
        out32 = out.astype(np.complex64).astype(np.complex128)
        Y32 = Y.astype(np.complex64).astype(np.complex128)
        f = np.einsum("so,bso->bs", Y32.conj(), out32)
        want_loss = 1 - (np.abs(f) ** 2).mean(1)
        want_seed = -2.0 / S * f[:, :, None] * Y32[None]
        scale = max(1.0, np.abs(want_loss).max())
        assert np.abs(loss.cpu().numpy() - want_loss).max() < 1e-5 * scale * O, (B, S, O)
        assert np.abs(seed.cpu().numpy() - want_seed).max() < 1e-5 * np.abs(want_seed).max(), (B, S, O)


def test_adjoint_4x16_mbqml_vs_oracle_fd(rng):
    """The MB-QML step configuration (4x16 grid, 60 angles, 256 Haar inputs),
    checked on 2 parameter sets against the oracle's central FD (h = 1e-5)."""
    torch = _torch()
    import mentpy_amd as mp
    from mentpy_amd.gradients import mbqml_infidelity
    from oracle.gradients import fd_gradient
    g = mp.templates.grid_cluster(4, 16)
    sim = mp.PatternSimulator(g, backend="hip-sv")
    X, Y = _dataset(4, 256, rng)
    th = rng.uniform(0, 2 * np.pi, (2, 60)).astype(np.float32)
    loss_t, g_t = mbqml_infidelity(sim, th, X, Y, mode="transfer")
    loss_d, g_d = mbqml_infidelity(sim, th, X, Y, mode="direct")
    loss_g, g_g = mbqml_infidelity(sim, th, X, Y, mode="gram")
    assert (g_t - g_d).abs().max().item() < 1e-4
    assert (g_g - g_d).abs().max().item() < 1e-4
    assert (loss_t - loss_d).abs().max().item() < 1e-5
    assert (loss_g - loss_d).abs().max().item() < 1e-5
    cost = _oracle_cost(g, X, Y)
    ref = fd_gradient(cost, th[0].astype(np.float64), h=1e-5)
    err = np.abs(g_t[0].cpu().numpy() - ref).max()
    print(f"4x16 adjoint vs oracle FD(h=1e-5): {err:.2e}", flush=True)
    assert err < 2e-4


def test_mbqml_step_graph_replay(rng):
    """HIP-graph capture of a whole loss+gradient step, replayed with new angles
    and eager work in between; every replay equals the eager step."""
    torch = _torch()
    import mentpy_amd as mp
    from mentpy_amd.gradients import MBQMLStep
    g = mp.templates.grid_cluster(4, 7)
    sim = mp.PatternSimulator(g, backend="hip-sv")
    X, Y = _dataset(4, 64, rng)
    for mode in ("gram", "direct"):
        step = MBQMLStep(sim, X, Y, n_sets=8, mode=mode)
        graph = step.capture()
        for rep in range(3):
            th = torch.tensor(rng.uniform(0, 2 * np.pi, (8, len(g.trainable_nodes))),
                              dtype=torch.float32, device="cuda")
            step.angles.copy_(th)
            graph.replay()
            torch.cuda.synchronize()
            l_graph = step.loss.clone()
            g_graph = (step.part.sum(1) if step.P > 1 else step.part[:, 0]).clone()
            l_eager, g_eager = step.run(th)
            torch.cuda.synchronize()
            assert torch.allclose(l_graph, l_eager, atol=1e-6), (mode, rep)
            assert torch.allclose(g_graph, g_eager, atol=1e-6), (mode, rep)


def test_autograd_function_custom_loss(rng):
    """Gradient of an arbitrary (phase-invariant) loss through autograd."""
    torch = _torch()
    import mentpy_amd as mp
    from mentpy_amd.gradients import pattern_sv
    from oracle.gradients import exact_shift_gradient
    from oracle.vectorized import sv_batch
    g = mp.templates.grid_cluster(2, 5)
    sim = mp.PatternSimulator(g, backend="hip-sv")
    cols = np.stack([haar_state(2, rng) for _ in range(3)])
    w = rng.uniform(0.5, 1.5, 4)
    th = torch.tensor(rng.uniform(0, 2 * np.pi, (1, len(g.trainable_nodes))), dtype=torch.float32,
                      device="cuda", requires_grad=True)
    out = pattern_sv(sim, th, torch.as_tensor(cols, dtype=torch.complex64, device="cuda"))
    wt = torch.as_tensor(w, dtype=torch.float32, device="cuda")
    loss = ((out.abs() ** 2) * wt).sum()
    loss.backward()

    def cost(x):
        o = sv_batch(g, x[None], cols)[0]
        return float(((np.abs(o) ** 2) * w).sum())
    ref = exact_shift_gradient(cost, th.detach().cpu().numpy()[0].astype(np.float64))
    err = np.abs(th.grad.cpu().numpy()[0] - ref).max()
    print(f"autograd custom loss grad err {err:.2e}", flush=True)
    assert err < 2e-4


def test_two_pattern_sv_calls_before_one_backward(rng):
    """Two applications summed into one loss before backward(): each keeps
    its own angles and checkpoints, so both terms differentiate correctly."""
    torch = _torch()
    import mentpy_amd as mp
    from mentpy_amd.gradients import pattern_sv
    from oracle.gradients import exact_shift_gradient
    from oracle.vectorized import sv_batch
    g = mp.templates.grid_cluster(3, 5)
    sim = mp.PatternSimulator(g, backend="hip-sv")
    n = len(g.trainable_nodes)
    c1 = np.stack([haar_state(3, rng) for _ in range(2)])
    c2 = np.stack([haar_state(3, rng) for _ in range(5)])
    w1, w2 = rng.uniform(0.5, 1.5, 8), rng.uniform(0.5, 1.5, 8)
    a1 = torch.tensor(rng.uniform(0, 2 * np.pi, (1, n)), dtype=torch.float32, device="cuda",
                      requires_grad=True)
    a2 = torch.tensor(rng.uniform(0, 2 * np.pi, (1, n)), dtype=torch.float32, device="cuda",
                      requires_grad=True)
    dev = lambda x: torch.as_tensor(x, dtype=torch.complex64, device="cuda")
    o1 = pattern_sv(sim, a1, dev(c1))
    o2 = pattern_sv(sim, a2, dev(c2))
    wt1 = torch.as_tensor(w1, dtype=torch.float32, device="cuda")
    wt2 = torch.as_tensor(w2, dtype=torch.float32, device="cuda")
    loss = ((o1.abs() ** 2) * wt1).sum() + ((o2.abs() ** 2) * wt2).sum()
    loss.backward()

    def cost(x, cols, w):
        o = sv_batch(g, x[None], cols)[0]
        return float(((np.abs(o) ** 2) * w).sum())
    for a, cols, w in ((a1, c1, w1), (a2, c2, w2)):
        x0 = a.detach().cpu().numpy()[0].astype(np.float64)
        ref = exact_shift_gradient(lambda x: cost(x, cols, w), x0)
        err = np.abs(a.grad.cpu().numpy()[0] - ref).max()
        print(f"two pending forwards: grad err {err:.2e}", flush=True)
        assert err < 2e-4


@pytest.mark.parametrize("mode", ["gram", "transfer", "direct"])
def test_mbqml_infidelity_new_data_same_shape(mode, rng):
    """Calls with different data of the same shape (fresh numpy arrays, and a
    device tensor changed in place) use the new data, not a cached step's."""
    torch = _torch()
    import mentpy_amd as mp
    from mentpy_amd.gradients import mbqml_infidelity
    from oracle.gradients import exact_shift_gradient
    g = mp.templates.grid_cluster(2, 5)
    sim = mp.PatternSimulator(g, backend="hip-sv")
    th = rng.uniform(0, 2 * np.pi, (2, len(g.trainable_nodes))).astype(np.float32)
    S = 3 if mode == "direct" else 12
    sets = [_dataset(2, S, rng) for _ in range(3)]

    def check(X, Y, loss, grad):
        cost = _oracle_cost(g, X, Y)
        t0 = th[1].astype(np.float64)
        assert abs(loss[1].item() - cost(t0)) < 2e-5
        err = np.abs(grad[1].cpu().numpy() - exact_shift_gradient(cost, t0)).max()
        assert err < 2e-4, err
    for X, Y in sets[:2]:                      # numpy arrays, new values each call
        loss, grad = mbqml_infidelity(sim, th, X, Y, mode=mode)
        check(X, Y, loss, grad)
    Xd = torch.as_tensor(sets[0][0], dtype=torch.complex64, device="cuda")
    Yd = torch.as_tensor(sets[0][1], dtype=torch.complex64, device="cuda")
    loss, grad = mbqml_infidelity(sim, th, Xd, Yd, mode=mode)
    check(*sets[0], loss, grad)
    Xd.copy_(torch.as_tensor(sets[2][0], dtype=torch.complex64))   # in place: same tensor
    Yd.copy_(torch.as_tensor(sets[2][1], dtype=torch.complex64))
    loss, grad = mbqml_infidelity(sim, th, Xd, Yd, mode=mode)
    check(*sets[2], loss, grad)


def test_adam_with_adjoint_cost_decreases_loss(rng):
    torch = _torch()
    import mentpy_amd as mp
    from mentpy_amd.gradients import AdjointCost
    g = mp.templates.grid_cluster(2, 6)
    g[3] = mp.Ment("X")
    sim = mp.PatternSimulator(g, backend="hip-sv")
    X, Y = _dataset(2, 12, rng)
    cost = AdjointCost(sim, X, Y)
    x0 = rng.uniform(0, 2 * np.pi, len(g.trainable_nodes))
    opt = mp.optimizers.AdamOptimizer(step_size=0.08)
    c0 = cost(x0)
    x = opt.optimize(cost, x0, num_iters=40)
    c1 = cost(x)
    print(f"Adam: cost {c0:.4f} -> {c1:.4f}", flush=True)
    assert c1 < c0


def test_device_adam_matches_host_adam(rng):
    """Device Adam/SGD updates equal mentpy's host formulas (adam.py:54-66, sgd.py:49-59)."""
    torch = _torch()
    from mentpy_amd.optimizers import DeviceAdam, DeviceSGD
    x0 = rng.uniform(0, 6, (4, 9))
    grads = [rng.normal(size=(4, 9)) for _ in range(5)]
    x = torch.tensor(x0, dtype=torch.float32, device="cuda")
    opt = DeviceAdam(x, step_size=0.08)
    m = np.zeros_like(x0)
    v = np.zeros_like(x0)
    xr = x0.copy()
    for i, g in enumerate(grads):
        opt.step(torch.tensor(g, dtype=torch.float32, device="cuda"))
        m = 0.9 * m + 0.1 * g
        v = 0.999 * v + 0.001 * g ** 2
        xr = xr - 0.08 * (m / (1 - 0.9 ** (i + 1))) / (np.sqrt(v / (1 - 0.999 ** (i + 1))) + 1e-8)
    assert np.abs(x.cpu().numpy() - xr).max() < 1e-5
    y = torch.tensor(x0, dtype=torch.float32, device="cuda")
    sgd = DeviceSGD(y, step_size=0.1, momentum=0.5, nesterov=True)
    vv = np.zeros_like(x0)
    yr = x0.copy()
    for g in grads:
        sgd.step(torch.tensor(g, dtype=torch.float32, device="cuda"))
        vv = 0.5 * vv - 0.1 * g
        yr = yr + 0.5 * vv - 0.1 * g
    assert np.abs(y.cpu().numpy() - yr).max() < 1e-5


def test_mbqml_trainer_graph_training_reduces_loss(rng):
    """A whole training iteration (adjoint step + device Adam) replayed as a
    HIP graph gives the same losses as the eager iteration and trains."""
    torch = _torch()
    import mentpy_amd as mp
    from mentpy_amd.optimizers import MBQMLTrainer
    g = mp.templates.grid_cluster(2, 6)
    sim = mp.PatternSimulator(g, backend="hip-sv")
    X, Y = _dataset(2, 32, rng)
    p0 = rng.uniform(0, 2 * np.pi, (16, len(g.trainable_nodes)))
    eager = MBQMLTrainer(sim, X, Y, p0, step_size=0.08)
    h_eager = eager.run(30, graph=False).cpu().numpy()
    graphed = MBQMLTrainer(sim, X, Y, p0, step_size=0.08)
    graphed.step_obj.angles.copy_(torch.tensor(p0, dtype=torch.float32, device="cuda"))
    first = graphed.run(1, graph=False)
    graphed.capture()
    h_graph = torch.cat([first, graphed.run(28)]).cpu().numpy()
    # the capture warm-up advanced one extra iteration: compare shifted histories
    assert np.abs(h_graph[0] - h_eager[0]).max() < 1e-6
    assert np.abs(h_graph[1:] - h_eager[2:]).max() < 1e-4
    assert (h_eager[-1] < h_eager[0]).mean() > 0.9


@pytest.mark.parametrize("rows,cols,S,B", [(1, 5, 4, 3), (2, 4, 9, 40), (3, 5, 40, 17), (4, 6, 300, 33),
                                          (4, 6, 64, 2100)])
def test_fused_gram_loss_matches_library_gemm(rows, cols, S, B, rng):
    """The fused Gram loss/seed kernel against the torch GEMM formulation of the
    same step (K = 4, 16, 64, 256; set counts not multiples of the tile, up to
    2,100 sets)."""
    torch = _torch()
    import mentpy_amd as mp
    from mentpy_amd.gradients import MBQMLStep
    g = mp.templates.grid_cluster(rows, cols) if rows > 1 else mp.templates.linear_cluster(cols)
    sim = mp.PatternSimulator(g, backend="hip-sv")
    X, Y = _dataset(rows, S, rng)
    th = rng.uniform(0, 2 * np.pi, (B, len(g.trainable_nodes)))
    a = MBQMLStep(sim, X, Y, n_sets=B, mode="gram", gram_kernel="fused")
    b = MBQMLStep(sim, X, Y, n_sets=B, mode="gram", gram_kernel="blas")
    la, ga = a.run(th)
    la, ga = la.clone(), ga.clone()
    lb, gb = b.run(th)
    print(f"{rows}x{cols} S={S} B={B}: loss diff {(la - lb).abs().max().item():.2e} "
          f"grad diff {(ga - gb).abs().max().item():.2e}", flush=True)
    assert (la - lb).abs().max().item() < 2e-6
    assert (ga - gb).abs().max().item() < 2e-5


@pytest.mark.parametrize("rows,cols,S,segs,mode", [
    (4, 16, 256, 8, "gram"), (4, 16, 40, 2, "transfer"), (2, 7, 9, 3, "gram"), (3, 6, 20, 4, "gram"),
    (5, 5, 40, 2, "transfer"), (1, 9, 5, 4, "gram"), (4, 6, 30, 5, "gram"), (3, 11, 24, 10, "gram"),
    (4, 20, 40, 15, "gram"), (4, 20, 40, 16, "transfer"), (4, 16, 64, 4, "gram"),
    (4, 20, 40, 4, "transfer")])
def test_segmented_adjoint_matches_unsegmented(rows, cols, S, segs, mode, rng):
    """The time-segmented transfer-matrix adjoint (segments' transfer matrices
    computed concurrently, chained per parameter set) gives the transfer
    matrix, loss and gradient of the single sweep, and the exact shift rule;
    segment counts up to one block per segment, set counts that leave
    partial waves; 16-amplitude windows chain on the matrix cores."""
    torch = _torch()
    import mentpy_amd as mp
    from mentpy_amd.gradients import MBQMLStep
    from oracle.gradients import exact_shift_gradient
    g = mp.templates.grid_cluster(rows, cols) if rows > 1 else mp.templates.linear_cluster(cols)
    sim = mp.PatternSimulator(g, backend="hip-sv")
    X, Y = _dataset(rows, S, rng)
    B = 7
    th = rng.uniform(0, 2 * np.pi, (B, len(g.trainable_nodes))).astype(np.float32)
    one = MBQMLStep(sim, X, Y, n_sets=B, mode=mode, segments=1)
    cut = MBQMLStep(sim, X, Y, n_sets=B, mode=mode, segments=segs)
    assert cut.segments == segs
    l1, g1 = one.run(th)
    l1, g1, t1 = l1.clone(), g1.clone(), one.out.clone()
    l2, g2 = cut.run(th)
    tscale = t1.abs().max().item()
    terr = (cut.out - t1).abs().max().item() / tscale
    lerr = (l1 - l2).abs().max().item()
    gerr = (g1 - g2).abs().max().item()
    print(f"{rows}x{cols} {mode} segments={segs}: T rel err {terr:.2e} loss diff {lerr:.2e} "
          f"grad diff {gerr:.2e}", flush=True)
    assert terr < 2e-6
    assert lerr < 2e-6
    assert gerr < 2e-5
    cost = _oracle_cost(g, X, Y)
    ref_g = exact_shift_gradient(cost, th[0].astype(np.float64))
    assert np.abs(g2[0].cpu().numpy() - ref_g).max() < 2e-4


@pytest.mark.parametrize("rows,cols,mode,S,B", [(4, 16, "gram", 64, 4096), (2, 7, "direct", 256, 256),
                                                (3, 6, "direct", 128, 512), (1, 9, "direct", 64, 1024),
                                                (5, 4, "direct", 128, 512)])
def test_adjoint_checkpoint_pairs(rows, cols, mode, S, B, rng):
    """From 65,536 columns the unsegmented adjoint keeps one checkpoint per
    two blocks and recomputes the other block in the backward sweep (odd and
    even block counts): it matches the exact shift rule of the oracle and,
    for the transfer matrix, the time-segmented path."""
    torch = _torch()
    import mentpy_amd as mp
    from mentpy_amd.gradients import MBQMLStep
    from oracle.gradients import exact_shift_gradient
    g = mp.templates.grid_cluster(rows, cols) if rows > 1 else mp.templates.linear_cluster(cols)
    sim = mp.PatternSimulator(g, backend="hip-sv")
    X, Y = _dataset(rows, S, rng)
    th = rng.uniform(0, 2 * np.pi, (B, len(g.trainable_nodes))).astype(np.float32)
    one = MBQMLStep(sim, X, Y, n_sets=B, mode=mode, segments=1)
    l1, g1 = one.run(th)
    l1, g1 = l1.clone(), g1.clone()
    if mode == "gram":
        cut = MBQMLStep(sim, X, Y, n_sets=B, mode=mode, segments=2)
        l2, g2 = cut.run(th)
        assert (l1 - l2).abs().max().item() < 2e-6
        assert (g1 - g2).abs().max().item() < 2e-5
    cost = _oracle_cost(g, X, Y)
    for b in (0, B - 1):
        ref = exact_shift_gradient(cost, th[b].astype(np.float64))
        err = np.abs(g1[b].cpu().numpy() - ref).max()
        print(f"{rows}x{cols} {mode} B={B} set {b}: grad err {err:.2e}", flush=True)
        assert err < 2e-4
        assert abs(l1[b].item() - cost(th[b].astype(np.float64))) < 2e-5


@pytest.mark.parametrize("rows,cols,w,S,mode,segs", [
    (3, 5, 5, 16, "transfer", 1), (2, 6, 4, 9, "gram", 1), (3, 5, 5, 4, "direct", 1),
    (3, 6, 5, 20, "gram", 3), (2, 7, 4, 9, "transfer", 2), (3, 8, 5, 20, "gram", 4),
    (3, 8, 5, 12, "transfer", 4)])
def test_adjoint_table_path_wider_window(rows, cols, w, S, mode, segs, rng):
    """Grids on a window above the minimum are not periodic: the adjoint
    kernels (single sweep and time-segmented) read the pair coefficients from
    the tables, and still match the exact shift rule."""
    torch = _torch()
    import mentpy_amd as mp
    from mentpy_amd.gradients import MBQMLStep
    from oracle.gradients import exact_shift_gradient
    g = mp.templates.grid_cluster(rows, cols)
    sim = mp.PatternSimulator(g, backend="hip-sv", window_size=w)
    assert sim.simulator.plan.info["periodic"] == 0
    X, Y = _dataset(rows, S, rng)
    B = 3
    th = rng.uniform(0, 2 * np.pi, (B, len(g.trainable_nodes))).astype(np.float32)
    step = MBQMLStep(sim, X, Y, n_sets=B, mode=mode, segments=segs)
    loss, grad = step.run(th)
    loss, grad = loss.cpu().numpy(), grad.cpu().numpy()
    cost = _oracle_cost(g, X, Y)
    for b in range(B):
        ref_g = exact_shift_gradient(cost, th[b].astype(np.float64))
        assert abs(loss[b] - cost(th[b].astype(np.float64))) < 2e-5
        assert np.abs(grad[b] - ref_g).max() < 2e-4
