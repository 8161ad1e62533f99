"""Batched persistent launch: uneven patterns in one launch, HIP-graph replay,
autotuned launch shape; each job checked against its own single-pattern
launch (bit-identical) and a sample against the complex128 oracle."""
import numpy as np
import pytest

from conftest import haar_state

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _jobs(rng, n_pat=10, B=40, S=3):
    import mentpy_amd as mp
    jobs = []
    for _ in range(n_pat):
        r = int(rng.integers(2, 6))
        c = int(rng.integers(4, 12))
        g = mp.templates.grid_cluster(r, c)
        sim = mp.PatternSimulator(g, backend="hip-sv")
        th = rng.uniform(0, 2 * np.pi, (B, len(g.trainable_nodes))).astype(np.float32)
        st = np.stack([haar_state(r, rng) for _ in range(S)])
        jobs.append((sim, th, st, g))
    return jobs


def test_batched_matches_single_launches(rng, tmp_path):
    torch = _torch()
    from mentpy_amd.batched import BatchedForward, autotune
    from oracle.vectorized import sv_batch
    jobs = _jobs(rng)
    bf = BatchedForward([(s, th, st) for s, th, st, _ in jobs], sims_per_item=64)
    cfg = autotune(bf, cache=tmp_path / "tune.json", reps=1,
                   candidates=[{"waves": 256, "wg_waves": 1, "sims_per_item": 64},
                               {"waves": 1024, "wg_waves": 4, "sims_per_item": 128}])
    assert (tmp_path / "tune.json").exists()
    outs = [o.clone() for o in bf.launch()]
    torch.cuda.synchronize()
    for (sim, th, st, g), o in zip(jobs, outs):
        single = sim.run_batch(th, st)
        assert (single - o).abs().max().item() < 1e-6, "batched result differs from the single launch"
        ref = sv_batch(g, th[:4].astype(np.float64), st)
        assert np.abs(o[:4].cpu().numpy() - ref).max() < 2e-5
    # graph capture and replay, with eager work in between
    graph = bf.capture_graph()
    for rep in range(3):
        for o in bf.outputs:
            o.zero_()
        graph.replay()
        torch.cuda.synchronize()
        for o, want in zip(bf.outputs, outs):
            assert torch.equal(o, want), f"graph replay {rep} differs"
    print(f"autotune chose {cfg}", flush=True)


@pytest.mark.parametrize("S_list", [(16, 16, 16, 8, 16, 16), (8, 3, 16, 8, 4, 8)])
def test_batched_shared_coefficient_groups(S_list, rng):
    """Window-size groups whose jobs all have S a multiple of the lane-group
    size (16 lanes for 2-3-slot windows, which evaluate all steps of an angle
    fetch at once; 4 for 4 slots, 8 for 5-6) evaluate the sines, cosines and
    reference phases once per group of lanes; a group with one job that cannot
    share (S = 3 or 8 with 2-3 slots, S = 4 with 6 slots) runs unshared.  Either
    way every job equals its single launch and the oracle."""
    torch = _torch()
    import mentpy_amd as mp
    from mentpy_amd.batched import BatchedForward
    from oracle.vectorized import sv_batch
    jobs = []
    for i, S in enumerate(S_list):
        r = 2 + i % 5
        g = mp.templates.grid_cluster(r, 5 + i)
        sim = mp.PatternSimulator(g, backend="hip-sv")
        th = rng.uniform(0, 2 * np.pi, (24, len(g.trainable_nodes))).astype(np.float32)
        st = np.stack([haar_state(r, rng) for _ in range(S)])
        jobs.append((sim, th, st, g))
    bf = BatchedForward([(s, th, st) for s, th, st, _ in jobs], sims_per_item=128)
    outs = [o.clone() for o in bf.launch(waves=512, wg_waves=2)]
    torch.cuda.synchronize()
    for (sim, th, st, g), o in zip(jobs, outs):
        single = sim.run_batch(th, st)
        assert (single - o).abs().max().item() < 1e-6
        ref = sv_batch(g, th[:3].astype(np.float64), st)
        assert np.abs(o[:3].cpu().numpy() - ref).max() < 2e-5


def test_batched_periodic_groups(rng):
    """Window-size groups of periodic plans (grids) take compile-time pair
    coefficients; a group holding one non-periodic plan (a 3-row grid on a
    5-qubit window joins the 4-row grids) reads the tables.  Both are
    bit-identical to the all-table launch and equal the single launches."""
    torch = _torch()
    import mentpy_amd as mp
    from mentpy_amd.batched import BatchedForward
    jobs = []
    for r, c, w in [(2, 6, None), (3, 5, None), (4, 7, None), (5, 4, None), (3, 4, 5), (4, 5, None)]:
        g = mp.templates.grid_cluster(r, c)
        sim = mp.PatternSimulator(g, backend="hip-sv", **({"window_size": w} if w else {}))
        assert sim.simulator.plan.info["periodic"] == (w is None)
        th = rng.uniform(-7, 7, (40, len(g.trainable_nodes))).astype(np.float32)
        st = np.stack([haar_state(r, rng) for _ in range(3)])
        jobs.append((sim, th, st))
    bf = BatchedForward(jobs, sims_per_item=64)
    for ref_phase in (False, True):
        a = [o.clone() for o in bf.launch(waves=256, wg_waves=2, ref_phase=ref_phase)]
        b = [o.clone() for o in bf.launch(waves=256, wg_waves=2, ref_phase=ref_phase,
                                          table_steps=True)]
        torch.cuda.synchronize()
        for x, y in zip(a, b):
            assert torch.equal(x, y)
    for (sim, th, st), o in zip(jobs, a):    # reference phase, as run_batch
        assert (sim.run_batch(th, st) - o).abs().max().item() < 1e-6


def test_batched_group_waves(rng):
    """Resident waves chosen per live-window group: every sizing gives the
    uniform launch's outputs bit for bit; a group given 0 waves is skipped
    (its outputs keep what they held) while the others are written; counts
    below the workgroup size and too many entries are rejected."""
    torch = _torch()
    from mentpy_amd.batched import MAX_GROUPS, BatchedForward
    jobs = _jobs(rng, n_pat=8)
    bf = BatchedForward([(s, th, st) for s, th, st, _ in jobs], sims_per_item=64)
    groups = bf.groups()
    assert sum(groups.values()) == bf.n_items and len(groups) > 1
    want = [o.clone() for o in bf.launch(waves=512, wg_waves=2)]
    torch.cuda.synchronize()
    for sizes in ([4, 8, 12, 16, 20, 24, 28], [64] * MAX_GROUPS, [2] * MAX_GROUPS):
        for o in bf.outputs:
            o.fill_(complex(float("nan"), 0.0))
        got = bf.launch(wg_waves=2, group_waves=sizes)
        torch.cuda.synchronize()
        for x, y in zip(got, want):
            assert torch.equal(x, y), f"group_waves {sizes} differs from the uniform launch"
    skip = min(groups)
    for o in bf.outputs:
        o.zero_()
    bf.launch(wg_waves=2, group_waves=[0 if ns == skip else 8 for ns in range(MAX_GROUPS)])
    torch.cuda.synchronize()
    for (sim, _, _, _), o, w in zip(jobs, bf.outputs, want):
        if sim.simulator.plan.info["n_slots"] == skip:
            assert o.abs().max().item() == 0.0
        else:
            assert torch.equal(o, w)
    with pytest.raises(ValueError):
        bf.launch(wg_waves=2, group_waves=[1] * MAX_GROUPS)
    with pytest.raises(ValueError):
        bf.launch(wg_waves=2, group_waves=[8] * (MAX_GROUPS + 1))


def test_batched_lds_groups(rng):
    """Groups of 7..14 live slots run persistent workgroups with the state in
    LDS (15..20: in a global workspace the batch allocates when it is created,
    so its launches stay capturable), concurrently with the
    register-kernel groups: every job equals its
    single launch bit for bit (the same LDS simulation code), a sample matches
    the oracle, every group sizing gives the same outputs, and a graph replay
    reproduces the launch."""
    torch = _torch()
    import mentpy_amd as mp
    from mentpy_amd.batched import MAX_GROUPS, BatchedForward
    from oracle.vectorized import sv_batch
    jobs = []
    for r, c, w, B, S in [(3, 5, None, 40, 3), (7, 3, None, 9, 2), (3, 5, 9, 30, 4), (8, 3, None, 5, 3),
                          (4, 6, None, 64, 2), (10, 3, None, 3, 2), (5, 6, 15, 4, 1),
                          (5, 6, 17, 3, 1)]:
        g = mp.templates.grid_cluster(r, c)
        sim = mp.PatternSimulator(g, backend="hip-sv", **({"window_size": w} if w else {}))
        th = rng.uniform(-7, 7, (B, len(g.trainable_nodes))).astype(np.float32)
        st = np.stack([haar_state(r, rng) for _ in range(S)])
        jobs.append((sim, th, st, g))
    bf = BatchedForward([(s, th, st) for s, th, st, _ in jobs], sims_per_item=64)
    groups = bf.groups()
    assert {7, 8, 10, 14, 16} <= set(groups) and {3, 4} <= set(groups)
    want = [o.clone() for o in bf.launch(waves=512, wg_waves=1)]
    torch.cuda.synchronize()
    for (sim, th, st, g), o in zip(jobs, want):
        single = sim.run_batch(th, st)
        ns = sim.simulator.plan.info["n_slots"]
        if ns > 6:
            assert torch.equal(single, o), f"{ns}-slot job differs from its single launch"
        else:
            assert (single - o).abs().max().item() < 1e-6
        ref = sv_batch(g, th[:2].astype(np.float64), st)
        err = np.abs(o[:2].cpu().numpy() - ref).max()
        print(f"{ns} slots: max amp err vs oracle {err:.2e}", flush=True)
        assert err < 2e-5
    for sizes in ([1] * MAX_GROUPS, [4, 8, 12, 16, 20, 24, 28, 3, 5, 1, 64, 2, 2, 9, 1000, 7, 3, 2, 4, 5, 6]):
        for o in bf.outputs:
            o.fill_(complex(float("nan"), 0.0))
        got = bf.launch(wg_waves=1, group_waves=sizes)
        torch.cuda.synchronize()
        for x, y in zip(got, want):
            assert torch.equal(x, y), f"group_waves {sizes} differs"
    graph = bf.capture_graph(waves=256, wg_waves=1)
    for rep in range(3):
        for o in bf.outputs:
            o.zero_()
        graph.replay()
        torch.cuda.synchronize()
        for o, w in zip(bf.outputs, want):
            assert torch.equal(o, w), f"graph replay {rep} differs"
