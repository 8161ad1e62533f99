"""Many patterns, angle vectors and inputs in one persistent launch.

mentpy simulates one (pattern, angles, input) per Python call; a training or
expressivity sweep loops over them (`docs/tutorials/intro-to-mbqml.rst:43-54`,
`docs/tutorials/intro-to-mbqml-parallel.rst:43-47`).  ``BatchedForward`` packs
jobs (pattern, angles ``[B, n]``, inputs ``[S, 2^|I|]``) of uneven sizes into
one work list, sorted longest first, and runs them with one persistent kernel
whose wavefronts pull work items from a device counter.  The launch (counter
reset kernel + persistent kernel) can be captured in a HIP graph and replayed.

``autotune`` picks the resident wavefront count, wavefronts per workgroup and
simulations per work item by timing graph replays of candidate launches with
HIP events, then
(with several live-window groups) wavefront counts per group in proportion to
each group's solo time, and caches the choice per (device, job signature) in a
JSON file.
"""
from __future__ import annotations

import ctypes as C
import hashlib
import json
import os
from pathlib import Path
from typing import List, Optional, Sequence, Tuple

from . import _native as nat
from .simulators.hip_simulator import _angle_strides, _as_angles, _as_states, _require_gpu

__all__ = ["BatchedForward", "autotune", "default_cache_path"]

# live-window sizes 0..20 of the batched statevector kernels (one simulation per
# lane up to 6 slots, one per workgroup with the state in LDS up to 14, in a
# global workspace above)
MAX_GROUPS = 21
# tag of the cached tuning results (entries timed another way are not reused)
_TUNE_VERSION = "graph-replay"


def _torch():
    import torch
    return torch


def default_cache_path() -> Path:
    base = os.environ.get("MENTPY_AMD_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "mentpy_amd"))
    return Path(base) / "autotune.json"


class BatchedForward:
    """One persistent statevector launch over a list of jobs.

    ``jobs``: sequence of ``(sim, angles, inputs)`` where ``sim`` is a hip-sv
    ``PatternSimulator`` (or its backend), ``angles`` ``[B, n]`` and
    ``inputs`` ``[S, 2^|I|]``.  Outputs are ``[B, S, 2^|O|]`` device tensors in
    ``self.outputs`` (same order as the jobs).
    """

    def __init__(self, jobs: Sequence[Tuple], sims_per_item: int = 64, device="cuda"):
        torch = _torch()
        self.device = torch.device(device)
        _require_gpu(self.device)
        self.sims_per_item = int(sims_per_item)
        self._jobs = []
        self.outputs = []
        for sim, angles, inputs in jobs:
            core = getattr(sim, "simulator", sim)
            plan = core.plan
            a = _as_angles(angles, plan.info["n_angles"], self.device)
            s = _as_states(inputs, plan.info["n_inputs"], self.device)
            out = torch.empty((a.shape[0], s.shape[0], 1 << plan.info["n_outputs"]),
                              dtype=torch.complex64, device=self.device)
            self._jobs.append((plan, a, s, out))
            self.outputs.append(out)
        self._handle = None
        self._build(self.sims_per_item)
        self.config = None

    def _build(self, sims_per_item: int):
        nat.release_pending()
        if self._handle is not None:
            nat.lib().mbqc_batch_destroy(self._handle)
            self._handle = None
        arr = (nat.BatchJob * len(self._jobs))()
        for k, (plan, a, s, out) in enumerate(self._jobs):
            sb, sa = _angle_strides(a)
            arr[k] = nat.BatchJob(plan=plan.handle.value, angles=a.data_ptr(),
                                  angle_stride=sb, angle_stride_a=sa, psi_in=s.data_ptr(),
                                  in_stride=s.shape[1], out=out.data_ptr(), B=a.shape[0], S=s.shape[0])
        h = C.c_void_p()
        with _torch().cuda.device(self.device):
            nat.check(nat.lib().mbqc_batch_create(arr, len(self._jobs), int(sims_per_item), C.byref(h)))
        self._handle = h
        self.sims_per_item = int(sims_per_item)

    def __del__(self):
        # queued, not destroyed here: a finalizer may run inside a graph capture
        h = getattr(self, "_handle", None)
        if h is not None and h.value and nat is not None:
            nat.defer_destroy("mbqc_batch_destroy", h.value)
            self._handle = None

    @property
    def n_items(self) -> int:
        return int(nat.lib().mbqc_batch_items(self._handle))

    @property
    def total_sims(self) -> int:
        return sum(a.shape[0] * s.shape[0] for _, a, s, _ in self._jobs)

    def signature(self) -> str:
        parts = sorted((p.info["n_slots"], p.info["n_steps"], a.shape[0] * s.shape[0])
                       for p, a, s, _ in self._jobs)
        return hashlib.sha1(repr(parts).encode()).hexdigest()[:16]

    def groups(self) -> dict:
        """Work items per live-window size (the groups that run as separate,
        concurrent persistent kernels)."""
        L = nat.lib()
        out = {}
        for ns in range(MAX_GROUPS):
            n = int(L.mbqc_batch_group_items(self._handle, ns))
            if n > 0:
                out[ns] = n
        return out

    def default_config(self):
        torch = _torch()
        cus = torch.cuda.get_device_properties(self.device).multi_processor_count
        return {"waves": cus * 4 * 4, "wg_waves": 4, "sims_per_item": self.sims_per_item}

    def launch(self, waves: Optional[int] = None, wg_waves: Optional[int] = None,
               group_waves: Optional[Sequence[int]] = None, ref_phase: bool = True,
               renormalize: bool = True, table_steps: bool = False):
        """One persistent launch over every job.  ``waves`` resident
        wavefronts per live-window group, or ``group_waves[ns]`` for the group
        of size ns (0 skips that group); without either, the tuned (or
        default) configuration.  ``table_steps`` makes groups of periodic plans
        read the pair coefficients from the tables (same results)."""
        torch = _torch()
        cfg = self.config or self.default_config()
        wg = int(wg_waves or cfg["wg_waves"])
        if group_waves is None and waves is None:
            group_waves = cfg.get("group_waves")
        flags = ((nat.FLAG_REF_PHASE if ref_phase else 0) | (nat.FLAG_RENORMALIZE if renormalize else 0)
                 | (nat.FLAG_TABLE_STEPS if table_steps else 0))
        stream = C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)
        with torch.cuda.device(self.device):
            if group_waves is not None:
                gw = [int(x) for x in group_waves]
                if len(gw) > MAX_GROUPS:
                    raise ValueError(f"group_waves has at most {MAX_GROUPS} entries")
                arr = (C.c_int32 * len(gw))(*gw)
                nat.check(nat.lib().mbqc_batch_launch_groups(self._handle, arr, len(gw), wg, flags,
                                                             stream))
            else:
                waves = int(waves or cfg["waves"])
                nat.check(nat.lib().mbqc_batch_launch(self._handle, waves, wg, flags, stream))
        return self.outputs

    def capture_graph(self, **launch_kwargs):
        """Capture one launch (counter reset + persistent kernel) in a HIP graph."""
        torch = _torch()
        self.launch(**launch_kwargs)            # warm-up outside capture
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.launch(**launch_kwargs)
        torch.cuda.synchronize(self.device)
        return g


def _time(fn, reps=5):
    """Mean ms of ``fn``'s launch captured in a HIP graph and replayed
    ``reps`` times (how the batch runs in production: the replays carry no
    host-side launch cost, which would otherwise blur candidates apart)."""
    torch = _torch()
    fn()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        with torch.cuda.graph(g):
            fn()
    torch.cuda.synchronize()
    g.replay()
    torch.cuda.synchronize()
    s = torch.cuda.Event(enable_timing=True)
    e = torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(reps):
        g.replay()
    e.record()
    torch.cuda.synchronize()
    del g
    return s.elapsed_time(e) / reps


def autotune(batch: BatchedForward, cache: Optional[Path] = None, reps: int = 10,
             candidates: Optional[List[dict]] = None, verbose: bool = False) -> dict:
    """Time candidate launch shapes, keep the fastest, cache it per signature."""
    torch = _torch()
    cache = Path(cache) if cache is not None else default_cache_path()
    props = torch.cuda.get_device_properties(batch.device)
    # architecture and CU count (the marketing name can be empty, e.g. under a profiler)
    key = f"{props.gcnArchName}|{props.multi_processor_count}|{batch.signature()}|{_TUNE_VERSION}"
    table = {}
    if cache.exists():
        try:
            table = json.loads(cache.read_text())
        except (OSError, ValueError):
            table = {}
    if key in table:
        cfg = table[key]
        if cfg["sims_per_item"] != batch.sims_per_item:
            batch._build(cfg["sims_per_item"])
        batch.config = cfg
        return cfg
    cus = torch.cuda.get_device_properties(batch.device).multi_processor_count
    if candidates is None:
        candidates = [{"waves": cus * 4 * occ, "wg_waves": wg, "sims_per_item": spi}
                      for spi in (64, 128, 256) for occ in (2, 4, 8) for wg in (1, 4)]
    best, best_t = None, float("inf")
    for cfg in candidates:
        if cfg["sims_per_item"] != batch.sims_per_item:
            batch._build(cfg["sims_per_item"])
        t = _time(lambda: batch.launch(waves=cfg["waves"], wg_waves=cfg["wg_waves"]), reps)
        if verbose:
            print(f"autotune {cfg}: {t:.3f} ms", flush=True)
        if t < best_t:
            best, best_t = dict(cfg), t
    if best["sims_per_item"] != batch.sims_per_item:
        batch._build(best["sims_per_item"])
    # second pass: with several window-size groups, size each group's
    # resident waves by its solo time, so that the concurrent kernels finish
    # together (equal counts leave the cheapest-per-wave group running alone)
    groups = batch.groups()
    if len(groups) > 1:
        wg, W = best["wg_waves"], best["waves"]
        total = W * len(groups)
        solo = {}
        for ns in groups:
            gw = [0] * MAX_GROUPS
            gw[ns] = total
            solo[ns] = _time(lambda: batch.launch(wg_waves=wg, group_waves=gw), reps)
        norm = sum(solo.values())
        for f in (0.5, 0.75, 1.0, 1.5):
            gw = [0] * MAX_GROUPS
            for ns, t in solo.items():
                gw[ns] = max(wg, int(round(f * total * t / norm / wg)) * wg)
            t = _time(lambda: batch.launch(wg_waves=wg, group_waves=gw), reps)
            if verbose:
                print(f"autotune group_waves {gw}: {t:.3f} ms", flush=True)
            if t < best_t:
                best = {**best, "group_waves": gw}
                best_t = t
    best["ms"] = best_t
    batch.config = best
    table[key] = best
    try:
        cache.parent.mkdir(parents=True, exist_ok=True)
        cache.write_text(json.dumps(table, indent=1))
    except OSError:
        pass
    return best
