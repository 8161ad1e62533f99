"""A numpy model of the per-step-probability density-matrix kernel (CPU).

``k_dm_general`` (``mentpy_amd/csrc/kernels.hip``) runs the reference's
density-matrix step in any plane (`mentpy/simulators/np_simulator_dm.py:307-346`)
on the compiled plan tables: the outcome-0 probability from the measured
slot's block traces, outcome 1 below 1e-4, the quartet update with population
weights W0, W1 and coherence coefficient C, and per-step normalisation.  This
model repeats that arithmetic (same tables, same formulas, noise-free) in
complex128 and must equal the vectorised oracle, so every formula the GPU
evaluates is checked without a GPU.  The chain cases have one live slot, where
the appended qubit's CZ with the measured one precedes the measurement and
removes the coherence term from the probability.
"""
import numpy as np
import pytest

import mentpy_amd as mp
from conftest import haar_state
from mentpy_amd.simulators.hip_simulator import compile_pattern
from oracle.vectorized import dm_batch


def kernel_model(plan, th, psi, keep_coh_on_cz=False):
    st, in_index, sign, out_pos = plan.tables()
    ns = plan.info['n_slots']; H = 1 << ns; HQ = H // 2
    v = psi[in_index] * np.where(sign, -1, 1)
    r = np.outer(v, v.conj())
    outs = []
    for k in range(plan.info['n_steps']):
        ai, flags, nmask, fixed = st[k]
        flags = int(flags); nmask = int(nmask); ai = int(ai)
        q = (flags >> 8) & 0xff; plane = (flags >> 2) & 3; app = flags & 1; sneg = flags & 2
        QB = 1 << q
        t = th[ai] if ai >= 0 else fixed
        c, s = np.cos(t), np.sin(t)
        nx, ny, nz = {0: (c, s, 0), 1: (c, 0, s), 2: (0, c, s), 3: (0, 0, 1)}[plane]
        t0 = t1 = x = y = 0
        def iz(a): return ((a >> q) << (q + 1)) | (a & ((1 << q) - 1))
        for a in range(HQ):
            j = iz(a); t0 += r[j, j].real; t1 += r[j | QB, j | QB].real
            b10 = r[j | QB, j]; x += b10.real; y += b10.imag
        tt = t0 + t1
        coh_term = 0 if (sneg and not keep_coh_on_cz) else 2 * (nx * x + ny * y)
        p0 = ((1 + nz) * t0 + (1 - nz) * t1 + coh_term) / (2 * tt)
        if plane == 3:
            w0 = w1 = 1; C = 0; den = tt; outc = 1 - p0
        else:
            one = p0 < 1e-4; g = -1 if one else 1
            w0 = 1 + g * nz; w1 = 1 - g * nz; C = g * (nx - 1j * ny)
            po = 1 - p0 if one else p0; den = 2 * po * tt; outc = int(one)
        sc = 1 / (2 * den if app else den)
        w0 *= sc; w1 *= sc; C *= sc
        outs.append(outc)
        new = np.zeros_like(r)
        for i in range(HQ * HQ):
            j = iz(i // HQ); kk = iz(i % HQ)
            B00 = r[j, kk]; B01 = r[j, kk | QB]; B10 = r[j | QB, kk]; B11 = r[j | QB, kk | QB]
            U = np.conj(C) * B01; V = C * B10
            P = w0 * B00 + w1 * B11
            r00 = P + U + V
            cz = lambda z: (-1) ** (bin(z & nmask).count('1') & 1)
            if app:
                sj, sk = cz(j), cz(kk)
                if sneg:
                    M = w0 * B00 - w1 * B11
                    new[j, kk] = r00; new[j, kk | QB] = sk * (M + V - U)
                    new[j | QB, kk] = sj * (M + U - V); new[j | QB, kk | QB] = sj * sk * (P - U - V)
                else:
                    new[j, kk] = r00; new[j, kk | QB] = sk * r00; new[j | QB, kk] = sj * r00
                    new[j | QB, kk | QB] = sj * sk * r00
            else:
                new[j, kk] = r00
        r = new
    d = 1 << plan.info['n_outputs']
    o = np.zeros((d, d), complex)
    for a in range(H):
        for b in range(H):
            if out_pos[a] >= 0 and out_pos[b] >= 0: o[out_pos[a], out_pos[b]] = r[a, b]
    return o / np.trace(o).real, outs



def _mixed(rows, cols, rng, z_nodes=0):
    g = mp.templates.grid_cluster(rows, cols) if rows > 1 else mp.templates.linear_cluster(cols)
    meas = [v for v in g.measurement_order if v not in g.output_nodes]
    for v in meas:
        g[v] = mp.Ment(plane=str(rng.choice(["XY", "XZ", "YZ"])))
    picks = rng.choice(meas, 2 + z_nodes, replace=False)
    g[int(picks[0])] = mp.Ment("X")
    g[int(picks[1])] = mp.Ment(float(np.float32(rng.uniform(0, 6))), "YZ")
    for v in picks[2:]:
        g[int(v)] = mp.Ment("Z")
    return g


@pytest.mark.parametrize("rows,cols,z", [(1, 5, 0), (1, 8, 0), (2, 3, 0), (2, 5, 1), (3, 4, 0),
                                         (3, 6, 1)])
def test_kernel_model_matches_oracle(rows, cols, z, rng):
    g = _mixed(rows, cols, rng, z_nodes=z)
    plan, *_ = compile_pattern(g, exclude="quantum_output", all_planes=True)
    th = rng.uniform(0, 2 * np.pi, (4, len(g.trainable_nodes))).astype(np.float32).astype(np.float64)
    st = np.stack([haar_state(len(g.input_nodes), rng) for _ in range(2)])
    mode = "expectation" if z else "sample"
    ref = dm_batch(g, th, st, mode=mode)
    for b in range(4):
        for s in range(2):
            got, _ = kernel_model(plan, th[b], st[s])
            assert np.abs(got - ref[b, s]).max() < 1e-6


def test_planted_coherence_term_is_caught():
    """Planted mistake: keeping the coherence term in the probability of the
    chain's CZ-first steps takes spurious outcomes 1 and breaks agreement."""
    rng = np.random.default_rng(1234)
    g = _mixed(1, 6, rng)
    plan, *_ = compile_pattern(g, exclude="quantum_output", all_planes=True)
    th = rng.uniform(0, 2 * np.pi, (24, len(g.trainable_nodes))).astype(np.float32).astype(np.float64)
    st = np.stack([haar_state(1, rng) for _ in range(2)])
    ref = dm_batch(g, th, st)
    good = max(np.abs(kernel_model(plan, th[b], st[s])[0] - ref[b, s]).max()
               for b in range(24) for s in range(2))
    bad = max(np.abs(kernel_model(plan, th[b], st[s], keep_coh_on_cz=True)[0] - ref[b, s]).max()
              for b in range(24) for s in range(2))
    assert good < 1e-6 and bad > 0.5


def test_kernel_model_outcome_one(rng):
    """Input |1> with the first chain node measured at XZ pi/2 takes outcome 1."""
    g = mp.templates.linear_cluster(6)
    first = [v for v in g.measurement_order if v not in g.output_nodes][0]
    g[first] = mp.Ment(float(np.float32(np.pi / 2)), "XZ")
    plan, *_ = compile_pattern(g, exclude="quantum_output", all_planes=True)
    th = rng.uniform(0, 2 * np.pi, len(g.trainable_nodes)).astype(np.float32).astype(np.float64)
    one = np.array([0, 1], dtype=complex)
    got, outs = kernel_model(plan, th, one)
    assert outs[0] == 1
    assert np.abs(got - dm_batch(g, th[None], one[None])[0, 0]).max() < 1e-6
