"""ctypes binding of ``libmbqc_hip.so`` (C ABI declared in ``include/mbqc_hip.h``).

The library is loaded from ``mentpy_amd/lib`` (built in-tree by
``mentpy_amd/csrc/build.py``).  There is no fallback: if the library is
missing, every entry point raises ``RuntimeError``.  Status codes map to the
exception classes the reference raises for the same conditions.
"""
from __future__ import annotations

import ctypes as C
import os
import threading
from pathlib import Path
from typing import Optional

import numpy as np

# MENTPY_AMD_LIB selects another build of the same library (the host-sanitized
# one of ``mentpy_amd/csrc/build.py --sanitize``).
LIB_PATH = Path(os.environ.get("MENTPY_AMD_LIB")
                or Path(__file__).resolve().parent / "lib" / "libmbqc_hip.so")

MBQC_OK = 0
MBQC_ERR_VALUE = -1
MBQC_ERR_TYPE = -2
MBQC_ERR_NOT_IMPLEMENTED = -3
MBQC_ERR_RUNTIME = -4

FLAG_REF_PHASE = 1
FLAG_RENORMALIZE = 2
FLAG_WRITE_PHASE = 4
FLAG_TABLE_STEPS = 8
FLAG_STEP_PROBS = 16

PLANE_XY = 0
PLANE_XZ = 1
PLANE_YZ = 2
PLANE_Z = 3

EXPORTED_SYMBOLS = (
    "mbqc_causal_flow", "mbqc_plan_create", "mbqc_plan_upload", "mbqc_plan_destroy", "mbqc_plan_get_info", "mbqc_plan_get_tables",
    "mbqc_plan_get_rescale", "mbqc_sv_forward_c64", "mbqc_sv_sample_c64", "mbqc_dm_forward_c64",
    "mbqc_dm_forward_outcomes_c64",
    "mbqc_plan_window_qubits", "mbqc_sv_window_c64", "mbqc_dm_window_c64",
    "mbqc_gram_workspace_bytes", "mbqc_gram_loss_seed_c64", "mbqc_gram_loss_seed_c64_noalloc",
    "mbqc_gram_finish_c64",
    "mbqc_overlap_loss_seed_c64",
    "mbqc_sv_adjoint_workspace_bytes", "mbqc_sv_adjoint_forward_c64",
    "mbqc_sv_adjoint_partials", "mbqc_sv_adjoint_backward_c64", "mbqc_sv_segments_auto",
    "mbqc_sv_segmented_workspace_bytes", "mbqc_sv_segmented_forward_c64",
    "mbqc_sv_segmented_backward_c64", "mbqc_batch_create",
    "mbqc_batch_destroy", "mbqc_batch_items", "mbqc_batch_launch", "mbqc_batch_group_items",
    "mbqc_batch_launch_groups", "mbqc_adam_step_f32",
    "mbqc_sgd_step_f32", "mbqc_last_error", "mbqc_abi_version",
)


class PatternDesc(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_int32), ("n_edges", C.c_int32), ("edges", C.POINTER(C.c_int32)),
        ("schedule", C.POINTER(C.c_int32)), ("n_measure", C.c_int32), ("window", C.c_int32),
        ("n_inputs", C.c_int32), ("input_nodes", C.POINTER(C.c_int32)),
        ("n_outputs", C.c_int32), ("output_nodes", C.POINTER(C.c_int32)),
        ("angle_index", C.POINTER(C.c_int32)), ("fixed_angle", C.POINTER(C.c_float)),
        ("n_angles", C.c_int32), ("flow", C.POINTER(C.c_int32)),
        ("plane", C.POINTER(C.c_int32)), ("measure_order", C.POINTER(C.c_int32)),
    ]


class PlanInfo(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "n_slots", "lag", "n_steps", "n_inputs", "n_outputs", "n_angles", "n_nodes", "window",
        "n_appends", "n_dropped_edges", "min_window", "can_sample", "periodic", "general_planes",
        "shared_angles", "rotating")]


class Step(C.Structure):
    _fields_ = [("angle_index", C.c_int32), ("flags", C.c_uint32), ("nmask", C.c_uint32),
                ("fixed_angle", C.c_float)]


class Noise(C.Structure):
    _fields_ = [("p_depol", C.c_float), ("p_readout", C.c_float), ("amp_damping", C.c_float),
                ("phase_damping", C.c_float), ("phase_flip", C.c_float), ("gad_gamma", C.c_float),
                ("gad_p", C.c_float)]


class BatchJob(C.Structure):
    _fields_ = [("plan", C.c_void_p), ("angles", C.c_void_p), ("angle_stride", C.c_int64),
                ("angle_stride_a", C.c_int64), ("psi_in", C.c_void_p), ("in_stride", C.c_int64), ("out", C.c_void_p),
                ("B", C.c_int64), ("S", C.c_int64)]


_lib: Optional[C.CDLL] = None


def library_available() -> bool:
    return LIB_PATH.exists()


def lib() -> C.CDLL:
    """Load the native library (once) and declare the signatures."""
    global _lib
    if _lib is not None:
        return _lib
    # torch ships its own HIP runtime with the same SONAME; loading torch first
    # makes this library bind to that runtime, so both share one HIP context.
    import torch  # noqa: F401
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"native library {LIB_PATH} is not built; run `python mentpy_amd/csrc/build.py` first")
    L = C.CDLL(str(LIB_PATH))
    vp, i32, i64, u32, sz = C.c_void_p, C.c_int32, C.c_int64, C.c_uint32, C.c_size_t
    L.mbqc_plan_create.argtypes = [C.POINTER(PatternDesc), C.POINTER(vp)]
    L.mbqc_plan_create.restype = C.c_int
    L.mbqc_plan_upload.argtypes = [vp]
    L.mbqc_plan_upload.restype = C.c_int
    L.mbqc_plan_destroy.argtypes = [vp]
    L.mbqc_plan_destroy.restype = None
    L.mbqc_plan_get_info.argtypes = [vp, C.POINTER(PlanInfo)]
    L.mbqc_plan_get_info.restype = C.c_int
    L.mbqc_plan_get_tables.argtypes = [vp, C.POINTER(Step), C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)]
    L.mbqc_plan_get_tables.restype = C.c_int
    L.mbqc_plan_get_rescale.argtypes = [vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)]
    L.mbqc_plan_get_rescale.restype = C.c_int
    L.mbqc_sv_forward_c64.argtypes = [vp, vp, i64, i64, vp, i64, vp, vp, i64, i64, u32, vp]
    L.mbqc_sv_forward_c64.restype = C.c_int
    L.mbqc_gram_workspace_bytes.argtypes = [i64, i32]
    L.mbqc_gram_workspace_bytes.restype = sz
    L.mbqc_gram_loss_seed_c64.argtypes = [vp, vp, i64, i32, i32, vp, vp, vp, vp]
    L.mbqc_gram_loss_seed_c64.restype = C.c_int
    L.mbqc_gram_loss_seed_c64_noalloc.argtypes = [vp, vp, i64, i32, i32, vp, vp, vp]
    L.mbqc_gram_loss_seed_c64_noalloc.restype = C.c_int
    L.mbqc_plan_window_qubits.argtypes = [vp, i32]
    L.mbqc_plan_window_qubits.restype = i32
    L.mbqc_sv_window_c64.argtypes = [vp, vp, i64, vp, i64, i32, vp, i64, i64, u32, vp]
    L.mbqc_sv_window_c64.restype = C.c_int
    L.mbqc_dm_window_c64.argtypes = [vp, vp, i64, vp, i64, i32, vp, i64, i64, C.POINTER(Noise), vp,
                                     vp]
    L.mbqc_dm_window_c64.restype = C.c_int
    L.mbqc_causal_flow.argtypes = [i32, i32, vp, i32, vp, i32, vp, vp, vp, vp, vp]
    L.mbqc_causal_flow.restype = C.c_int
    L.mbqc_sv_sample_c64.argtypes = [vp, vp, i64, i64, vp, i64, vp, vp, i64, i64, C.c_uint64, vp]
    L.mbqc_sv_sample_c64.restype = C.c_int
    L.mbqc_dm_forward_c64.argtypes = [vp, vp, i64, i64, vp, i64, vp, i64, i64, C.POINTER(Noise), u32, vp]
    L.mbqc_dm_forward_c64.restype = C.c_int
    L.mbqc_dm_forward_outcomes_c64.argtypes = [vp, vp, i64, i64, vp, i64, vp, vp, i64, i64,
                                               C.POINTER(Noise), u32, vp]
    L.mbqc_dm_forward_outcomes_c64.restype = C.c_int
    L.mbqc_sv_adjoint_workspace_bytes.argtypes = [vp, i64]
    L.mbqc_sv_adjoint_workspace_bytes.restype = sz
    L.mbqc_sv_adjoint_forward_c64.argtypes = [vp, vp, i64, i64, vp, i64, vp, i64, i64, vp, sz, vp]
    L.mbqc_sv_adjoint_forward_c64.restype = C.c_int
    L.mbqc_sv_adjoint_partials.argtypes = [i64]
    L.mbqc_sv_adjoint_partials.restype = i32
    L.mbqc_sv_adjoint_backward_c64.argtypes = [vp, vp, i64, i64, vp, i64, i64, vp, sz, vp, vp]
    L.mbqc_sv_adjoint_backward_c64.restype = C.c_int
    L.mbqc_gram_finish_c64.argtypes = [vp, vp, i64, i32, i32, vp, vp, vp]
    L.mbqc_gram_finish_c64.restype = C.c_int
    L.mbqc_overlap_loss_seed_c64.argtypes = [vp, vp, i64, i32, i32, vp, vp, vp]
    L.mbqc_overlap_loss_seed_c64.restype = C.c_int
    L.mbqc_sv_segments_auto.argtypes = [vp, i64]
    L.mbqc_sv_segments_auto.restype = i32
    L.mbqc_sv_segmented_workspace_bytes.argtypes = [vp, i64, i32]
    L.mbqc_sv_segmented_workspace_bytes.restype = sz
    L.mbqc_sv_segmented_forward_c64.argtypes = [vp, vp, i64, i64, vp, i64, i32, vp, sz, vp]
    L.mbqc_sv_segmented_forward_c64.restype = C.c_int
    L.mbqc_sv_segmented_backward_c64.argtypes = [vp, vp, i64, i64, vp, i64, i32, vp, sz, vp, vp]
    L.mbqc_sv_segmented_backward_c64.restype = C.c_int
    L.mbqc_batch_create.argtypes = [C.POINTER(BatchJob), i32, i32, C.POINTER(vp)]
    L.mbqc_batch_create.restype = C.c_int
    L.mbqc_batch_destroy.argtypes = [vp]
    L.mbqc_batch_destroy.restype = None
    L.mbqc_batch_items.argtypes = [vp]
    L.mbqc_batch_items.restype = i64
    L.mbqc_batch_launch.argtypes = [vp, i32, i32, u32, vp]
    L.mbqc_batch_launch.restype = C.c_int
    L.mbqc_batch_group_items.argtypes = [vp, i32]
    L.mbqc_batch_group_items.restype = i64
    L.mbqc_batch_launch_groups.argtypes = [vp, C.POINTER(i32), i32, i32, u32, vp]
    L.mbqc_batch_launch_groups.restype = C.c_int
    f32 = C.c_float
    L.mbqc_adam_step_f32.argtypes = [vp, vp, vp, vp, i64, f32, f32, f32, f32, vp, vp]
    L.mbqc_adam_step_f32.restype = C.c_int
    L.mbqc_sgd_step_f32.argtypes = [vp, vp, vp, i64, f32, f32, i32, vp]
    L.mbqc_sgd_step_f32.restype = C.c_int
    L.mbqc_last_error.argtypes = []
    L.mbqc_last_error.restype = C.c_char_p
    L.mbqc_abi_version.argtypes = []
    L.mbqc_abi_version.restype = i32
    _lib = L
    return L


def check(rc: int) -> None:
    """Raise the Python exception matching a native status code."""
    if rc == MBQC_OK:
        return
    msg = lib().mbqc_last_error().decode(errors="replace")
    if rc == MBQC_ERR_VALUE:
        raise ValueError(msg)
    if rc == MBQC_ERR_TYPE:
        raise TypeError(msg)
    if rc == MBQC_ERR_NOT_IMPLEMENTED:
        raise NotImplementedError(msg)
    raise RuntimeError(msg)


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


# Native objects whose Python owners were garbage-collected.  Finalizers only
# queue them: the cyclic garbage collector can run a finalizer at any
# allocation, including while a stream is being captured into a HIP graph,
# where freeing device memory or destroying streams is not allowed (it breaks
# the capture).  The queue is emptied by release_pending() at points that
# allocate device memory themselves (plan and batch construction), which are
# never inside a capture.
_pending_lock = threading.Lock()
_pending: list = []


def defer_destroy(fn_name: str, handle_value: int) -> None:
    with _pending_lock:
        _pending.append((fn_name, handle_value))


def release_pending() -> int:
    """Destroy the queued native objects; returns how many were destroyed."""
    with _pending_lock:
        items = list(_pending)
        _pending.clear()
    L = _lib
    if L is None:
        return 0
    for fn_name, h in items:
        getattr(L, fn_name)(C.c_void_p(h))
    return len(items)


class NativePlan:
    """Owner of an ``mbqc_plan*`` compiled from flat pattern arrays."""

    def __init__(self, n_nodes: int, edges, schedule, n_measure: int, window: int,
                 input_nodes, output_nodes, angle_index, fixed_angle, n_angles: int,
                 flow=None, plane=None, measure_order=None):
        L = lib()
        release_pending()
        self._keep = dict(
            edges=_i32(np.asarray(edges, dtype=np.int32).reshape(-1, 2) if len(edges) else np.zeros((0, 2))),
            schedule=None if schedule is None else _i32(schedule), inputs=_i32(input_nodes),
            outputs=_i32(output_nodes),
            angle_index=_i32(angle_index),
            fixed=np.ascontiguousarray(np.asarray(fixed_angle, dtype=np.float32)),
            flow=None if flow is None else _i32(flow),
            plane=None if plane is None else _i32(plane),
            order=None if measure_order is None else _i32(measure_order),
        )
        k = self._keep
        p32 = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
        desc = PatternDesc(
            n_nodes=n_nodes, n_edges=len(k["edges"]), edges=p32(k["edges"]),
            schedule=None if k["schedule"] is None else p32(k["schedule"]), n_measure=n_measure,
            window=window,
            n_inputs=len(k["inputs"]), input_nodes=p32(k["inputs"]),
            n_outputs=len(k["outputs"]), output_nodes=p32(k["outputs"]),
            angle_index=p32(k["angle_index"]),
            fixed_angle=k["fixed"].ctypes.data_as(C.POINTER(C.c_float)), n_angles=n_angles,
            flow=None if k["flow"] is None else p32(k["flow"]),
            plane=None if k["plane"] is None else p32(k["plane"]),
            measure_order=None if k["order"] is None else p32(k["order"]))
        handle = C.c_void_p()
        check(L.mbqc_plan_create(C.byref(desc), C.byref(handle)))
        self.handle = handle
        info = PlanInfo()
        check(L.mbqc_plan_get_info(self.handle, C.byref(info)))
        self.info = {f: getattr(info, f) for f, _ in PlanInfo._fields_}

    def upload(self) -> None:
        """Copy the plan's tables to the current device (done at creation for
        the device current then); launches on that device then allocate
        nothing, so they can be captured in a HIP graph."""
        check(lib().mbqc_plan_upload(self.handle))

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and defer_destroy is not None:
            defer_destroy("mbqc_plan_destroy", h.value)
            self.handle = None

    def tables(self):
        n, H = self.info["n_steps"], 1 << self.info["n_slots"]
        steps = (Step * max(n, 1))()
        in_index = np.zeros(H, np.int32)
        sign = np.zeros(H, np.int32)
        out_pos = np.zeros(H, np.int32)
        p32 = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
        check(lib().mbqc_plan_get_tables(self.handle, steps, p32(in_index), p32(sign), p32(out_pos)))
        st = np.array([(steps[i].angle_index, steps[i].flags, steps[i].nmask, steps[i].fixed_angle)
                       for i in range(n)], dtype=np.float64).reshape(n, 4)
        return st, in_index, sign, out_pos

    def rescale(self):
        nb = -(-self.info["n_steps"] // self.info["n_slots"])
        sv = np.zeros(max(nb, 1), np.int32)
        dm = np.zeros(max(nb, 1), np.int32)
        e = np.zeros(2, np.int32)
        p32 = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
        check(lib().mbqc_plan_get_rescale(self.handle, p32(sv), p32(dm), p32(e)))
        return sv[:nb], dm[:nb], e
