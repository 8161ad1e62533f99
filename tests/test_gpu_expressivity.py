"""Expressivity sampler on the GPU: fidelities against the complex128 oracle
for the same angles and inputs, and the Haar limit of a long chain."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.mark.parametrize("rows,cols", [(1, 5), (2, 4), (3, 5)])
def test_sampled_fidelities_match_oracle(rows, cols):
    _torch()
    import mentpy_amd as mp
    from mentpy_amd.utils import generate_haar_random_states, sample_probability_density_of_fidelities
    from oracle.vectorized import sv_batch
    g = mp.templates.linear_cluster(cols) if rows == 1 else mp.templates.grid_cluster(rows, cols)
    n = 96
    fid = sample_probability_density_of_fidelities(g, n_samples=n, seed=11)
    # the same draws as the sampler: angles first, then the Haar states
    rng = np.random.default_rng(11)
    theta = rng.uniform(0.0, 2.0 * np.pi, (n, len(g.trainable_nodes)))
    psi = np.atleast_2d(generate_haar_random_states(rows, n, rng))
    ref = np.empty(n)
    for k in range(n):
        out = sv_batch(g, theta[k:k + 1].astype(np.float32).astype(np.float64), psi[k:k + 1])[0, 0]
        ref[k] = np.abs(np.vdot(psi[k], out)) ** 2 / np.vdot(out, out).real
    err = np.abs(fid - ref).max()
    print(f"{rows}x{cols} fidelity err {err:.2e}", flush=True)
    assert err < 2e-5


def test_long_chain_approaches_haar():
    """A 1-qubit chain with 24 random angles is close to Haar (mean F = 1/2,
    uniform density); a 2-angle chain is not (KL of a 50-bin histogram: about
    0.009 for the 2-angle chain, about bins / (2 n) = 2.5e-4 of sampling noise
    for a Haar ensemble)."""
    _torch()
    import mentpy_amd as mp
    from mentpy_amd.utils import expressivity_with_histogram, sample_probability_density_of_fidelities
    from scipy import stats
    long_chain = mp.templates.linear_cluster(25)
    f = sample_probability_density_of_fidelities(long_chain, n_samples=20000, seed=3)
    assert abs(f.mean() - 0.5) < 0.01
    assert stats.kstest(f, "uniform").pvalue > 1e-3
    kl_long = expressivity_with_histogram(long_chain, n_samples=100000, n_bins=50, seed=5)
    kl_short = expressivity_with_histogram(mp.templates.linear_cluster(3), n_samples=100000,
                                           n_bins=50, seed=5)
    print(f"KL long {kl_long:.3e} short {kl_short:.3e}", flush=True)
    assert kl_long < 0.002 and kl_short > 0.005
