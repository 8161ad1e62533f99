"""Density-matrix noise channels on the GPU against the complex128 oracle.

The oracle applies each channel's Kraus operators (oracle/channels.py) on
every qubit after its CZs; the kernels fold the same channels into a
coherence factor of the measured qubit and a population/coherence map of the
output qubits.  Bar: trace distance < 1e-5, unit trace, Hermitian.
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


CASES = {
    "depolarizing": (dict(circuit_noise="depolarizing", p=0.05), lambda ch: ch.depolarizing(0.05)),
    "amplitude_damping": (dict(circuit_noise="amplitude_damping", gamma=0.15),
                          lambda ch: ch.amplitude_damping(0.15)),
    "phase_damping": (dict(circuit_noise="phase_damping", gamma=0.2),
                      lambda ch: ch.phase_damping(0.2)),
    "phase_flip": (dict(circuit_noise="phase_flip", p=0.08), lambda ch: ch.phase_flip(0.08)),
    "generalized_amplitude_damping": (
        dict(circuit_noise="generalized_amplitude_damping", gamma=0.2, p=0.3),
        lambda ch: ch.generalized_amplitude_damping(0.2, 0.3)),
}


@pytest.mark.parametrize("rows,cols", [(1, 6), (2, 5), (3, 6), (4, 5), (5, 4)])
@pytest.mark.parametrize("name", sorted(CASES))
def test_dm_channel_matches_kraus_oracle(rows, cols, name, rng):
    torch = _torch()
    import mentpy_amd as mp
    from oracle import channels as ch
    from oracle.vectorized import dm_batch
    kw, kraus = CASES[name]
    g = mp.templates.linear_cluster(cols) if rows == 1 else mp.templates.grid_cluster(rows, cols)
    sim = mp.PatternSimulator(g, backend="hip-dm", p_readout=0.01, **kw)
    B, S = 5, 3
    th = rng.uniform(0, 2 * np.pi, (B, len(g.trainable_nodes))).astype(np.float32)
    st = np.stack([haar_state(rows, rng) for _ in range(S)])
    out = sim.run_batch(th, st).cpu().numpy()
    ref = dm_batch(g, th.astype(np.float64), st, p_readout=0.01, kraus=kraus(ch))
    td = 0.5 * np.abs(np.linalg.eigvalsh(out - ref)).sum(-1).max()
    tr = np.trace(out, axis1=-2, axis2=-1)
    print(f"{name} {rows}x{cols}: trace distance {td:.2e}", flush=True)
    assert td < 1e-5
    assert np.abs(tr - 1).max() < 1e-5
    assert np.abs(out - np.conj(np.swapaxes(out, -1, -2))).max() < 1e-6


@pytest.mark.parametrize("rows,cols", [(2, 6), (4, 4)])
def test_dm_composed_channels(rows, cols, rng):
    """Several channels at once (composed in the documented order)."""
    torch = _torch()
    import mentpy_amd as mp
    from oracle import channels as ch
    from oracle.vectorized import dm_batch
    nm = mp.NoiseModel(depolarizing=0.03, readout_flip=0.02, amplitude_damping=0.1,
                       phase_damping=0.05, phase_flip=0.04, gad_gamma=0.12, gad_p=0.8)
    g = mp.templates.grid_cluster(rows, cols)
    sim = mp.PatternSimulator(g, backend="hip-dm", noise=nm)
    th = rng.uniform(0, 2 * np.pi, (4, len(g.trainable_nodes))).astype(np.float32)
    st = np.stack([haar_state(rows, rng) for _ in range(2)])
    out = sim.run_batch(th, st).cpu().numpy()
    ks = ch.noise_kraus(0.03, 0.1, 0.05, 0.04, 0.12, 0.8)
    ref = dm_batch(g, th.astype(np.float64), st, p_readout=0.02, kraus=ks)
    td = 0.5 * np.abs(np.linalg.eigvalsh(out - ref)).sum(-1).max()
    print(f"composed {rows}x{cols}: trace distance {td:.2e}", flush=True)
    assert td < 1e-5


@pytest.mark.parametrize("rows,cols,composed", [(2, 6, False), (3, 6, False), (3, 5, True),
                                                (1, 7, True)])
def test_dm_lane_and_workgroup_kernels_agree(rows, cols, composed, rng):
    """Batches of >= 32768 simulations run one simulation per lane (Hermitian
    upper-triangle registers); small ones hold one density-matrix entry per
    lane (1,024 and fewer at 3 slots, up to 16,384 at 2, 32,768 at 1); 3-slot
    batches in between keep the density matrix in LDS.  All paths agree on the
    same inputs, and a sample of the large batch matches the oracle."""
    torch = _torch()
    import mentpy_amd as mp
    from oracle import channels as ch
    from oracle.vectorized import dm_batch
    g = mp.templates.grid_cluster(rows, cols) if rows > 1 else mp.templates.linear_cluster(cols)
    if composed:
        nm = mp.NoiseModel(depolarizing=0.02, readout_flip=0.01, amplitude_damping=0.05,
                           phase_damping=0.03, phase_flip=0.02, gad_gamma=0.1, gad_p=0.3)
        sim = mp.PatternSimulator(g, backend="hip-dm", noise=nm)
        kraus = ch.noise_kraus(0.02, 0.05, 0.03, 0.02, 0.1, 0.3)
    else:
        sim = mp.PatternSimulator(g, backend="hip-dm", circuit_noise="amplitude_damping",
                                  gamma=0.1, p_readout=0.01)
        kraus = ch.amplitude_damping(0.1)
    B = 65536
    th = torch.rand((B, len(g.trainable_nodes)), generator=torch.Generator().manual_seed(4))
    th = (th * 2 * np.pi).cuda()
    psi = haar_state(rows, rng)
    big = sim.run_batch(th, psi)                 # one simulation per lane
    mid = sim.run_batch(th[:3000], psi)          # 3 slots: LDS kernel; fewer: one entry per lane
    small = sim.run_batch(th[:300], psi)         # one entry per lane
    diff = max((big[:300] - small).abs().max().item(), (big[:3000] - mid).abs().max().item())
    idx = rng.choice(B, 64, replace=False)
    ref = dm_batch(g, th[idx].double().cpu().numpy(), psi[None], p_readout=0.01, kraus=kraus)
    td = 0.5 * np.abs(np.linalg.eigvalsh(big[idx].cpu().numpy() - ref)).sum(-1).max()
    print(f"{rows}x{cols}: lane vs workgroup {diff:.2e}, trace distance {td:.2e}", flush=True)
    assert diff < 2e-6
    assert td < 1e-5
