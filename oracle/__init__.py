"""CPU complex128 oracle for the measurement-pattern simulate path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` may import this package, and only to
check or time against it; the product (``mentpy_amd``) never imports it.

Contents (each module cites the mentpy 0.1.0a9 lines it restates):

* ``reference_sv`` / ``reference_dm`` -- step-by-step restatements of
  ``NumpySimulatorSV`` / ``NumpySimulatorDM`` (dense kron-built projectors and
  CZ matrices, amplitude-sum partial trace, per-step renormalisation).  The DM
  one adds the noise model of ``mentpy_amd.noise`` at the places the PennyLane
  backend puts ``circuit_noise``.  These are the timed CPU baseline.
* ``vectorized`` -- the same per-step operations vectorised over a batch of
  simulations (used for large-batch parity checks).
* ``brute`` -- full-register simulation of tiny patterns in gate order (all
  CZs, then noise, then projections), sharing no code with the window logic.
* ``pattern`` -- restatement of mentpy's causal-flow layering and measurement
  order, used to check the product's pattern compiler.
* ``closed_forms`` -- gate-model forms that pin the phase convention.
* ``gradients`` -- mentpy's finite-difference and parameter-shift gradients and
  the exact shift rule.

Parity status: every function here is pinned by a test in ``tests/`` against
an independent check (closed form, brute force, known-answer test of the
reference's own test suite, or a mathematical invariant); see the module
docstrings.
"""
