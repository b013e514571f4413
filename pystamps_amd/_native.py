"""ctypes binding of libpystamps_hip.so (C ABI in include/pystamps_hip.h).

The library is the product: there is no CPU fallback.  Loading fails loudly
when the library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os
import re
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("PS2_HIP_LIB") or PKG / "libpystamps_hip.so")
HEADER = PKG.parent / "include" / "pystamps_hip.h"

ABI_VERSION = 2
FORM_FP64 = 0
FORM_FP32 = 1
WEIGHT_PSQUARE = 0
WEIGHT_SNR = 1
COH_BINS = 100
KERNEL_SLOTS = ("fill_grid", "clap_strip", "clap_combine", "scan", "rescore",
                "refine", "post", "weights", "outputs")

_f64p = C.POINTER(C.c_double)
_f32p = C.POINTER(C.c_float)
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)


class PatchIn(C.Structure):
    _fields_ = [
        ("n_ps", C.c_int64), ("n_ifg", C.c_int64), ("n_i", C.c_int64), ("n_j", C.c_int64),
        ("ph_nm", _f32p), ("amp", _f32p), ("row_invariant", C.c_int32),
        ("bperp_row", _f64p), ("bperp_mat", _f64p), ("grid_lin", _i64p), ("d_a", _f64p),
        ("low_pass", _f64p), ("n_win_ex", C.c_int64), ("clap_window", C.c_int64),
        ("clap_alpha", C.c_double), ("clap_beta", C.c_double),
        ("trial_values", _f64p), ("n_trial", C.c_int64),
        ("nr_base", _f64p), ("nr_max_nz_ix", C.c_double), ("coh_bins", _f64p),
        ("low_coh_thresh", C.c_int64), ("weighting", C.c_int32),
        ("gamma_change_convergence", C.c_double), ("gamma_max_iterations", C.c_int64),
    ]


class PatchOut(C.Structure):
    _fields_ = [
        ("k_ps", _f64p), ("c_ps", _f64p), ("coh_ps", _f64p), ("n_opt", _f64p),
        ("coh_ps_save", _f64p), ("ph_res", _f32p), ("ph_patch", _f32p), ("ph_grid", _f32p),
        ("ph_weight", _f32p), ("nr", _f64p), ("gamma_change_save", C.c_double),
        ("i_loop", C.c_int64), ("trial_ix", _i32p), ("n_rescored", C.c_int64),
        ("n_rechecked", C.c_int64), ("loop_seconds", C.c_double),
    ]


class DeviceOutputs(C.Structure):
    _fields_ = [
        ("device", C.c_int32), ("n_ps", C.c_int64), ("n_ifg", C.c_int64),
        ("k_ps", C.c_void_p), ("c_ps", C.c_void_p), ("coh_ps", C.c_void_p), ("n_opt", C.c_void_p),
        ("ph_patch", C.c_void_p), ("ph_res", C.c_void_p), ("nr", C.c_double * COH_BINS),
        ("gamma_change_save", C.c_double), ("i_loop", C.c_int64),
    ]


class RunOpts(C.Structure):
    _fields_ = [
        ("form", C.c_int32), ("device", C.c_int32), ("rescore_bound", C.c_double),
        ("verbose", C.c_int32), ("reserved", C.c_int32), ("autotune", C.c_int32),
        ("autotune_cache", C.c_char_p),
    ]


class Tiles(C.Structure):
    _fields_ = [
        ("fill_cells", C.c_int32), ("scan_step_table", C.c_int32), ("clap_segments", C.c_int32),
        ("clap_windows_per_segment", C.c_int32), ("source", C.c_int32),
    ]


TILE_SOURCES = {0: "model", 1: "cache", 2: "measured"}


_lib = None


def declared_functions() -> list[str]:
    """Names of the functions include/pystamps_hip.h declares."""
    text = HEADER.read_text()
    return sorted(set(re.findall(r"\b(ps2_hip_\w+)\s*\(", text)))


# ps2_iteration_fn: int (*)(void *user, int64_t iteration, ps2_patch_out *state)
ITERATION_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int64, C.POINTER(PatchOut))


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is not built; run `python -m pystamps_amd.build` (hipcc, gfx950)")
    L = C.CDLL(str(LIB_PATH))
    L.ps2_hip_abi_version.restype = C.c_int
    L.ps2_hip_last_error.restype = C.c_char_p
    L.ps2_hip_device_count.restype = C.c_int
    L.ps2_hip_run.argtypes = [C.POINTER(PatchIn), C.POINTER(PatchOut), C.POINTER(RunOpts)]
    L.ps2_hip_run_batch.argtypes = [C.c_size_t, C.POINTER(PatchIn), C.POINTER(PatchOut),
                                    C.POINTER(RunOpts)]
    L.ps2_hip_random_coherence_histogram.argtypes = [
        _f64p, C.c_int64, _f64p, C.c_int64, _f64p, C.c_int64, C.c_uint32, C.c_int32, _f64p, _f64p]
    _i32p = C.POINTER(C.c_int32)
    L.ps2_hip_random_coherence_histogram_sb.argtypes = [
        _f64p, C.c_int64, _i32p, _i32p, C.c_int64, _f64p, C.c_int64, _f64p, C.c_int64, C.c_uint32,
        C.c_int32, _f64p, _f64p]
    L.ps2_hip_topofit_rows.argtypes = [_f64p, _f64p, C.c_int32, C.c_int64, C.c_int64, C.c_double,
                                       C.c_int32, _f64p, _f64p, _f64p, C.POINTER(C.c_float)]
    L.ps2_hip_topofit_coh_row_invariant.argtypes = [_f64p, _f64p, C.c_int64, C.c_int64, C.c_double,
                                                    C.c_int32, _f64p]
    L.ps2_hip_grid_accumulate.argtypes = [C.POINTER(C.c_float), _i64p, C.c_int64, C.c_int64,
                                          C.c_int64, C.c_int64, C.c_int32, C.POINTER(C.c_float)]
    L.ps2_hip_histogram_with_centers.argtypes = [_f64p, C.c_int64, _f64p, C.c_int64, _f64p]
    L.ps2_hip_session_export_device.argtypes = [C.c_void_p, C.POINTER(DeviceOutputs)]
    L.ps2_hip_normalize_phase.argtypes = [C.POINTER(C.c_float), C.c_int64, C.c_int64, _i64p, C.c_int64,
                                          C.c_int32, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.ps2_hip_select_columns_f64.argtypes = [_f64p, C.c_int64, C.c_int64, _i64p, C.c_int64, C.c_int32,
                                             _f64p]
    L.ps2_hip_session_create.argtypes = [C.POINTER(PatchIn), C.POINTER(RunOpts), C.POINTER(C.c_void_p)]
    L.ps2_hip_session_run.argtypes = [C.c_void_p, _i64p, _f64p]
    L.ps2_hip_session_download.argtypes = [C.c_void_p, C.POINTER(PatchOut)]
    L.ps2_hip_session_run_iterations.argtypes = [C.c_void_p, C.POINTER(PatchOut), ITERATION_FN, C.c_void_p,
                                                 _i64p]
    L.ps2_hip_session_destroy.argtypes = [C.c_void_p]
    L.ps2_hip_session_destroy.restype = None
    L.ps2_hip_step_fill_grid.argtypes = [C.c_void_p, _f64p, _f64p, C.c_void_p]
    L.ps2_hip_step_clap.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    L.ps2_hip_step_fit.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(PatchOut)]
    L.ps2_hip_sessions_run.argtypes = [C.c_size_t, C.POINTER(C.c_void_p), _i64p, _f64p]
    L.ps2_hip_session_set_timing.argtypes = [C.c_void_p, C.c_int32]
    L.ps2_hip_session_kernel_times.argtypes = [C.c_void_p, _f64p, _i64p]
    if L.ps2_hip_abi_version() != ABI_VERSION:
        raise ImportError("libpystamps_hip.so ABI version mismatch")
    _lib = L
    return L


def last_error() -> str:
    return lib().ps2_hip_last_error().decode("utf-8", "replace")


def device_count() -> int:
    n = lib().ps2_hip_device_count()
    return max(n, 0)
