"""Expressivity of a pattern from sampled fidelities, on the GPU.

Follows `mentpy/utils/expressivity.py:20-125`: sample random measurement
angles (uniform in [0, 2 pi)) and Haar-random input states, record the
fidelity of each output with its input, histogram the fidelities and compare
the histogram with the Haar density P(F) = (N-1)(1-F)^(N-2), N = 2^n
(Kullback-Leibler divergence, relative entropy or Jensen-Shannon distance).

mentpy's sampler cannot run as written (it calls simulator methods that do
not exist, `expressivity.py:98-106`, and compares against a kron of the input
vector); the fidelity used here is the one its docstrings describe,
F = |<psi|out(theta, psi)>|^2 for pure inputs.  All simulations of one call
run in one batched launch: every angle vector propagates the 2^|I| basis
inputs (the pattern's transfer matrix T, a unitary for flow patterns), and
each sample's output is T psi for its own Haar input, normalised.
"""
from __future__ import annotations

import numpy as np

__all__ = ["haar_probability_density_of_fidelities", "sample_probability_density_of_fidelities",
           "expressivity_with_histogram"]


def haar_probability_density_of_fidelities(F, n_qubits: int):
    """P_Haar(F) = (N-1)(1-F)^(N-2), N = 2^n (`expressivity.py:20-36`)."""
    N = int(2 ** n_qubits)
    return (N - 1) * ((1 - np.asarray(F)) ** (N - 2))


def sample_probability_density_of_fidelities(graph_state_circuit, n_samples: int = 1000,
                                             backend: str = "hip-sv", seed=None, device="cuda"):
    """``n_samples`` fidelities |<psi_k|out(theta_k, psi_k)>|^2 (numpy float64).

    ``theta_k`` uniform in [0, 2 pi) and ``psi_k`` Haar random, both drawn from
    ``numpy.random.default_rng(seed)``.
    """
    import torch
    from ..simulators import PatternSimulator
    from .generate_data import generate_haar_random_states
    if backend != "hip-sv":
        raise ValueError(f"Unsupported backend {backend}")
    circ = graph_state_circuit
    n_in, n_out = len(circ.input_nodes), len(circ.output_nodes)
    if n_in != n_out:
        raise ValueError("fidelity with the input needs as many outputs as inputs")
    rng = np.random.default_rng(seed)
    sim = PatternSimulator(circ, backend="hip-sv", device=device)
    n_ang = len(circ.trainable_nodes)
    theta = rng.uniform(0.0, 2.0 * np.pi, (n_samples, n_ang))
    psi = np.atleast_2d(generate_haar_random_states(n_in, n_samples, rng))
    d = 1 << n_in
    basis = np.eye(d, dtype=np.complex64)
    # unnormalised raw-frame columns: T e_c for every angle vector (exact for
    # any pattern, not only those whose T is a multiple of a unitary)
    cols = sim.run_batch(theta, basis, phase="raw", renormalize=False)   # [n, d_in, d_out]
    psi_t = torch.as_tensor(psi, dtype=torch.complex64, device=cols.device)
    out = (psi_t[:, :, None] * cols).sum(1)                              # T psi per sample
    ov = (psi_t.conj() * out).sum(-1)
    fid = (ov.abs() ** 2) / (out.abs() ** 2).sum(-1)
    return fid.double().cpu().numpy()


def expressivity_with_histogram(graph_state_circuit, n_samples: int = 10000, n_bins: int = 1000,
                                method: str = "KL", seed=None, device="cuda"):
    """Divergence between the sampled fidelity histogram and the Haar density
    (`expressivity.py:39-86`): ``method`` ``"KL"`` (sum of ``scipy.special.kl_div``),
    ``"RE"`` (``rel_entr``) or ``"JS"`` (``scipy.spatial.distance.jensenshannon``)."""
    from scipy.spatial import distance
    from scipy.special import kl_div, rel_entr
    if method not in ("KL", "RE", "JS"):
        raise UserWarning("Unsupported method for calculating expressivity")
    samples = sample_probability_density_of_fidelities(graph_state_circuit, n_samples=n_samples,
                                                       seed=seed, device=device)
    hist, edges = np.histogram(samples, bins=n_bins, density=True, range=(0, 1))
    mid = 0.5 * (edges[:-1] + edges[1:])
    haar = haar_probability_density_of_fidelities(mid, len(graph_state_circuit.output_nodes))
    p_data = hist * np.diff(edges)
    p_haar = haar * np.diff(edges)
    if method == "KL":
        return float(np.sum(kl_div(p_data, p_haar)))
    if method == "RE":
        return float(np.sum(rel_entr(p_data, p_haar)))
    return float(distance.jensenshannon(p_data, p_haar))
