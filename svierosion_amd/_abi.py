"""ctypes binding of libsvimx (include/svimx.h).

The library is built in-tree by ``svierosion_amd.build``; importing this
module loads it and fails loudly if it is missing -- there is no fallback.
Device buffers are torch tensors on the ROCm device; their ``data_ptr()`` and
the current torch stream are passed through the C ABI.
"""

from __future__ import annotations

import ctypes as C
import os
import threading
from pathlib import Path

import torch

LIB_PATH = Path(os.environ.get("SVIMX_LIB", Path(__file__).resolve().parent / "libsvimx.so"))

OK, EBADARG, EMARGIN, ECAPACITY, EDEVICE = 0, -1, -2, -3, -4

# result slots (SVIMX_R_*)
R_VOLUME_SUM, R_HAS_NEG, R_HAS_POS, R_VMAX = 0, 1, 2, 3
R_QMIN, R_QMAX, R_DVMAX, R_CAPPED, R_VSUM, R_NNEAR, R_DMAX = 4, 5, 6, 7, 8, 9, 10
R_QMIN_ALL, R_QMAX_ALL, R_DVMAX_ALL, R_CAPPED_ALL, R_VSUM_ALL, R_NALL, R_DMAX_ALL = 11, 12, 13, 14, 15, 16, 17
R_TILES = 18
R_DT = 31
DT_DEVICE = -1.0
R_COUNT = 32


class Law(C.Structure):
    """svimx_law_t."""
    _fields_ = [("kind", C.c_int32), ("k", C.c_int32), ("s", C.c_int32), ("reserved", C.c_int32),
                ("t", C.c_double), ("eps", C.c_double), ("n_cap", C.c_double),
                ("alpha", C.c_double), ("delta", C.c_double)]


_P, _I, _I64, _D, _U8P = C.c_void_p, C.c_int, C.c_int64, C.c_double, C.c_void_p


class JobDesc(C.Structure):
    """svimx_job_desc_t."""
    _fields_ = [("phi", C.c_void_p * 3), ("occupancy", C.c_void_p), ("band", C.c_void_p), ("band_cap", C.c_int64),
                ("ws", C.c_void_p), ("ws_bytes", C.c_int64), ("nx", C.c_int32), ("ny", C.c_int32),
                ("nz", C.c_int32), ("cur", C.c_int32), ("nb", C.c_int64), ("inside_outside", C.c_int64),
                ("h", C.c_double), ("h3", C.c_double), ("width", C.c_double), ("margin", C.c_int32),
                ("frac_valid", C.c_int32), ("law", Law), ("nr_plan", C.c_void_p), ("field", C.c_void_p),
                ("k12", C.c_void_p), ("reserved_ptr", C.c_void_p), ("radius", C.c_double), ("cfl", C.c_double),
                ("feedback_cfl", C.c_double), ("parabolic_cfl", C.c_double), ("reinit_distance", C.c_double),
                ("cutoff_lo", C.c_double), ("cutoff_hi", C.c_double), ("reinit_iterations", C.c_int32),
                ("reserved", C.c_int32), ("time", C.c_double), ("initial_volume", C.c_double),
                ("displacement", C.c_double), ("steps", C.c_int64)]


class ReinitGate(C.Structure):
    """svimx_reinit_gate_t."""
    _fields_ = [("displacement", C.c_double), ("threshold", C.c_double), ("inside_outside", C.c_double),
                ("h3", C.c_double), ("target", C.c_double), ("target_stop", C.c_double),
                ("has_target", C.c_int32), ("reserved", C.c_int32)]


class StepRec(C.Structure):
    """svimx_step_record_t."""
    _fields_ = [("step", C.c_int64), ("time", C.c_double), ("dt", C.c_double), ("volume", C.c_double),
                ("fraction", C.c_double), ("vmax", C.c_double), ("vmean", C.c_double), ("qmin", C.c_double),
                ("qmax", C.c_double), ("capped", C.c_int64), ("reinit", C.c_int32), ("reserved", C.c_int32),
                ("wall", C.c_double)]


class JobStatus(C.Structure):
    """svimx_job_status_t."""
    _fields_ = [("running", C.c_int32), ("stop_reason", C.c_int32), ("error", C.c_int32), ("cur", C.c_int32),
                ("frac_valid", C.c_int32), ("reserved", C.c_int32), ("nb", C.c_int64),
                ("inside_outside", C.c_int64), ("steps", C.c_int64), ("nrecords", C.c_int64),
                ("time", C.c_double), ("volume", C.c_double), ("displacement", C.c_double),
                ("message", C.c_char * 256)]


_SIGNATURES = {
    "svimx_last_error": (C.c_char_p, []),
    "svimx_version": (_I, []),
    "svimx_workspace_bytes": (_I64, [_I, _I, _I]),
    "svimx_occupancy_elems": (_I64, [_I, _I, _I]),
    "svimx_band_rebuild": (_I, [_P, _P, _P, _P, _P, _I, _I, _I, _D, _D, _I, _P, _I64, _P, _I64, _P]),
    "svimx_fractions": (_I, [_P, _P, _P, _I64, _I, _I, _I, _D, _P, _P, _D, _P, _I64]),
    "svimx_curvatures": (_I, [_P, _P, _P, _I64, _I, _I, _I, _D, _P]),
    "svimx_velocity": (_I, [_P, _P, _P, _I64, _I, _I, _I, _D, C.POINTER(Law), _P, _P, _D, _D, _I, _P,
                            _P, _I64]),
    "svimx_velocity_rk": (_I, [_P, _I, _P, _P, _P, _P, _I64, _I, _I, _I, _D, C.POINTER(Law), _P, _P, _D, _D, _D,
                               _P, _P, _I64]),
    "svimx_velocity_rk_points": (_I, [_P, _I, _P, _P, _P, _P, _I64, _I, _I, _I, _D, C.POINTER(Law), _P, _P, _D, _D,
                                      _D, _P, _P, _P, _I64]),
    "svimx_velocity_static": (_I, [_P, _P, _P, _I64, _I, _I, _I, _D, C.POINTER(Law), _P, _P, _D, _D, _I, _P,
                                   _P]),
    "svimx_rk_stage": (_I, [_P, _I, _P, _P, _P, _P, _I64, _I, _I, _I, _D, _D, _P, _P]),
    "svimx_rk_stage_dt": (_I, [_P, _I, _P, _P, _P, _P, _I64, _I, _I, _I, _D, _P, _P, _P]),
    "svimx_stable_dt": (_I, [_P, C.POINTER(Law), _D, _D, _D, _D, _D, _I, _D, _P, _I64]),
    "svimx_nr_plan_create": (_I, [_P, _I, _D, _D, _I, _I, _I, C.POINTER(C.c_void_p)]),
    "svimx_nr_plan_destroy": (_I, [_P]),
    "svimx_nr_reset": (_I, [_P, _P]),
    "svimx_nr_mark_nodes": (_I, [_P, _P, _P, _I64]),
    "svimx_nr_mark_points": (_I, [_P, _P, _P, _I64]),
    "svimx_mesh_signed_volume": (_I, [_P, _I64, _P, _I64, C.POINTER(_D)]),
    "svimx_mesh_edge_topology": (_I, [_P, _I64, _I64, C.POINTER(_I64)]),
    "svimx_sdf_prepare": (_I, [_P, _I64, _P, _I64, _P, _D, _P, _D, _D, _P, _P, _P, _P, _P, _P, _P, _P,
                               C.POINTER(_I64), C.POINTER(_I64)]),
    "svimx_nr_exec": (_I, [_P, _P, _P, _P, _P, _I64]),
    "svimx_nr_select": (_I, [_P, _P]),
    "svimx_nr_sum": (_I, [_P, _P, _P, _P, _P, _I64]),
    "svimx_nr_plan_tune": (_I, [_P, _P, _P, _P, _I64, _I, _I, _I, _D, _D, _P, _P, _P, _I64, C.POINTER(_D)]),
    "svimx_nr_plan_run_length": (_I, [_P]),
    "svimx_nr_plan_set_run_length": (_I, [_P, _I]),
    "svimx_nr_plan_set_lines": (_I, [_P, _I]),
    "svimx_engine_capture_fallbacks": (_I64, [_P]),
    "svimx_nr_plan_lines": (_I, [_P]),
    "svimx_nr_tune_shape": (_I, [_I, C.POINTER(_I), C.POINTER(_I), C.POINTER(_I)]),
    "svimx_nr_plan_set_staging": (_I, [_P, _I]),
    "svimx_nr_plan_staging": (_I, [_P]),
    "svimx_nr_plan_set_method": (_I, [_P, _I]),
    "svimx_nr_plan_method": (_I, [_P]),
    "svimx_reinit": (_I, [_P, _P, _P, _P, _P, _P, _I64, _I64, _I, _I, _I, _D, _D, _I, _D, _I, _I, _P, _P, _P, _P,
                          _P, _I64, _P, _I64, _P]),
    "svimx_reinit_gated": (_I, [_P, _P, _P, _P, _P, _P, _I64, _I64, _I, _I, _I, _D, _D, _I, _D, _I, _I, _P, _P, _P,
                                _P, _P, _I64, _P, _I64, C.POINTER(ReinitGate), _P]),
    "svimx_sdf_build": (_I, [_P, _P, _P, _P, _I64, _P, _P, _P, _P, _I64, _P, _I64, _I, _I, _I, _I, _I, _D, _D, _D,
                             _P, _P, _P, _P, _I64]),
    "svimx_extract_surface": (_I, [_P, _P, _P, _I64, _I, _I, _I, _D, C.POINTER(C.c_void_p),
                                   C.POINTER(C.c_int64), C.POINTER(C.c_void_p), C.POINTER(C.c_int64)]),
    "svimx_point_mesh_distance": (_I, [_P, _P, _P, _I64, _I, _I, _I, _D, _P, _I64, _P]),
    "svimx_free": (_I, [_P, _P]),
    "svimx_copy_to_host": (_I, [_P, _P, _P, _I64]),
    "svimx_timing_event_create": (_I, [C.POINTER(C.c_void_p)]),
    "svimx_timing_event_record": (_I, [_P, _P]),
    "svimx_timing_event_elapsed": (_I, [_P, _P, C.POINTER(C.c_float)]),
    "svimx_timing_event_destroy": (_I, [_P]),
    "svimx_engine_create": (_I, [_P, _I, C.POINTER(_P)]),
    "svimx_engine_destroy": (_I, [_P]),
    "svimx_engine_add": (_I, [_P, C.POINTER(JobDesc), C.POINTER(C.c_int)]),
    "svimx_engine_run": (_I, [_P, _I, _D, _D, _I64, _D]),
    "svimx_engine_tick": (_I, [_P, C.POINTER(C.c_int)]),
    "svimx_engine_ticks": (_I, [_P, _I, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "svimx_engine_status": (_I, [_P, _I, C.POINTER(JobStatus)]),
    "svimx_engine_records": (_I, [_P, _I, _I64, _I64, C.POINTER(StepRec)]),
    "svimx_engine_remove": (_I, [_P, _I]),
    "svimx_engine_use_graphs": (_I, [_P, _I]),
    "svimx_engine_group_nodes": (_I, [_P, _I64]),
    "svimx_engine_short_work": (_I, [_P, _D]),
    "svimx_engine_schedule": (_I, [_P, _I, _I64, _D]),
    "svimx_engine_stopped": (_I, [_P, C.POINTER(C.c_int), _I, C.POINTER(C.c_int)]),
    "svimx_engine_timers": (_I, [_P, C.POINTER(_D)]),
    "svimx_op_godunov": (_I, [_P, _P, _P, _I64, _I, _I, _I, _D, _P]),
    "svimx_op_cell_fractions": (_I, [_P, _P, _P, _I64, _I, _I, _I, _P]),
    "svimx_op_closest_points": (_I, [_P, _P, _P, _I64, _I, _I, _I, _D, _P]),
    "svimx_closest_points_newton": (_I, [_P, _P, _P, _I64, _I, _I, _I, _D, _I, _P]),
    "svimx_op_tricubic": (_I, [_P, _P, _I, _I, _I, _P, _I64, _P]),
    "svimx_op_tricubic_gradient": (_I, [_P, _P, _I, _I, _I, _P, _I64, _P, _P]),
    "svimx_op_trilinear": (_I, [_P, _P, _I, _I, _I, _P, _I64, _P]),
    "svimx_op_principal_curvatures": (_I, [_P, _P, _P, _I64, _I, _I, _I, _D, _P, _P]),
    "svimx_op_interface_mask": (_I, [_P, _P, _P, _I64, _I, _I, _I, _P]),
    "svimx_op_reinit_geometry": (_I, [_P, _P, _P, _I64, _I, _I, _I, _D, _D, _P]),
    "svimx_op_point_triangle_distance": (_I, [_P, _P, _P, _I64, _P]),
    "svimx_op_law": (_I, [_P, C.POINTER(Law), _P, _P, _I64, _P, _P]),
    "svimx_annotate_points": (_I, [_P, C.POINTER(Law), _P, _P, _I, _I, _I, _P, _I64, _P, _P, _P]),
    "svimx_smooth_size": (_I, [_I64, C.POINTER(_I64)]),
    "svimx_grid_around": (_I, [C.POINTER(_D), C.POINTER(_D), _D, _D, _I, C.POINTER(_I64), C.POINTER(_D)]),
    "svimx_law_init": (_I, [_I, _D, _D, _D, _D, _D, _D, _D, _D, C.POINTER(Law)]),
    "svimx_law_raw": (_I, [C.POINTER(Law), _D, C.POINTER(_D)]),
}


def _load():
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is not built; run `python -m svierosion_amd.build` "
                          "(svierosion_amd has no CPU fallback)")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()

# Held around the calls that allocate or free device memory synchronously
# (N_r plan creation and destruction, which run on set-up threads or in the
# garbage collector) and around the batch engine's ticks, which may capture a
# HIP graph: a synchronous free in another thread would otherwise invalidate
# a capture in flight.
device_lock = threading.RLock()


class SvimxError(RuntimeError):
    pass


def check(rc: int, what: str = "") -> None:
    """Map a libsvimx return code onto the reference's exception types."""
    if rc == OK:
        return
    msg = lib.svimx_last_error().decode(errors="replace")
    if what:
        msg = f"{what}: {msg}"
    if rc == EBADARG:
        raise ValueError(msg)
    if rc == EMARGIN:
        raise RuntimeError(msg)
    raise SvimxError(f"libsvimx error {rc}: {msg}")


def ptr(t):
    """Raw pointer of a tensor (None -> NULL)."""
    return None if t is None else C.c_void_p(t.data_ptr())


class TimingEvent:
    """A HIP timing event recorded on torch's current stream, without the
    system-scope fence of a default event (svimx_timing_event_*); the
    record / elapsed_time interface of torch.cuda.Event."""

    def __init__(self):
        h = C.c_void_p()
        check(lib.svimx_timing_event_create(C.byref(h)), "timing event")
        self._h = h

    def record(self) -> None:
        check(lib.svimx_timing_event_record(self._h, stream()), "timing event record")

    def elapsed_time(self, end: "TimingEvent") -> float:
        ms = C.c_float()
        check(lib.svimx_timing_event_elapsed(self._h, end._h, C.byref(ms)), "timing event elapsed")
        return float(ms.value)

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h and lib is not None:
            lib.svimx_timing_event_destroy(h)


def ptr_at(t, offset_bytes: int):
    """Raw pointer ``offset_bytes`` into a tensor's storage."""
    return C.c_void_p(t.data_ptr() + int(offset_bytes))


def stream():
    """hipStream_t of torch's current stream on the current device."""
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def require_device(device) -> torch.device:
    dev = torch.device(device)
    if dev.type != "cuda" or not torch.cuda.is_available():
        raise RuntimeError("svierosion_amd computes on a ROCm GPU; no GPU device is available "
                           f"(requested {dev}) and there is no CPU fallback")
    return dev


def workspace(shape, device) -> torch.Tensor:
    nbytes = int(lib.svimx_workspace_bytes(*map(int, shape)))
    return torch.zeros(nbytes, dtype=torch.uint8, device=device)


OCC_PAD = 16  # zero cells in front of every axis of the occupancy layout


def occupancy_zeros(shape, device) -> torch.Tensor:
    """Zeroed occupancy array in the padded layout of svimx_occupancy_elems."""
    return torch.zeros(int(lib.svimx_occupancy_elems(*map(int, shape))), dtype=torch.float32, device=device)


def occupancy_view(occ: torch.Tensor, shape) -> torch.Tensor:
    """Dense (nx, ny, nz) view of a padded occupancy array."""
    nx, ny, nz = map(int, shape)
    pz = occ.numel() // ((nx + 2 * OCC_PAD) * (ny + 2 * OCC_PAD))
    v = occ.view(nx + 2 * OCC_PAD, ny + 2 * OCC_PAD, pz)
    return v[OCC_PAD:OCC_PAD + nx, OCC_PAD:OCC_PAD + ny, OCC_PAD:OCC_PAD + nz]


def results_view(ws: torch.Tensor) -> torch.Tensor:
    return ws[: 8 * R_COUNT].view(torch.float64)


def exported_symbols():
    return sorted(_SIGNATURES)
