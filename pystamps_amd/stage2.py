"""Stage-2 gamma estimation on the GPU: the host-side mirror of the reference's
stage-2 entry points.

``run_stage2_hip(prepared)`` replaces the loop of ``run_stage2_native_inner``
(crates/pystamps-core/src/native_stage2.rs:242-406) and returns the
``write_pm1`` content (:1899-1965) as a dict keyed like pm1.mat.
``run_stage2_patch(raw)`` is prepare + loop, returning the reference's summary
string (:408-412) with the pm1 dict.  Failures raise :class:`Stage2Error`,
whose text matches CoreError::NativeStage{stage: 2} (lib.rs:47-48).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _native
from .prepare import RawPatch, Stage2Error, Stage2Prepared, prepare_stage2_inputs

FORMS = {"fp64": _native.FORM_FP64, "fp32": _native.FORM_FP32}

_f64p = C.POINTER(C.c_double)
_f32p = C.POINTER(C.c_float)
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)


def _ptr(a, t):
    return None if a is None else a.ctypes.data_as(t)


class _InHolder:
    """Keeps the contiguous arrays behind one PatchIn alive."""

    def __init__(self, prep: Stage2Prepared):
        self.ph_nm = np.ascontiguousarray(prep.ph_nm, dtype=np.complex64)
        self.amp = np.ascontiguousarray(prep.amp, dtype=np.float32)
        self.bperp_row = np.ascontiguousarray(prep.row_bperp_nm, dtype=np.float64)
        self.bperp_mat = (None if prep.bperp_mat is None
                          else np.ascontiguousarray(prep.bperp_mat, dtype=np.float64))
        self.grid_lin = np.ascontiguousarray(prep.grid_lin, dtype=np.int64)
        self.d_a = np.ascontiguousarray(prep.d_a, dtype=np.float64)
        self.low_pass = np.ascontiguousarray(prep.low_pass, dtype=np.float64)
        self.trial_values = np.ascontiguousarray(prep.trial_values, dtype=np.float64)
        self.nr_base = np.ascontiguousarray(prep.nr_base, dtype=np.float64)
        self.coh_bins = np.ascontiguousarray(prep.coh_bins, dtype=np.float64)
        s = _native.PatchIn()
        s.n_ps, s.n_ifg, s.n_i, s.n_j = prep.n_ps, prep.n_ifg, prep.n_i, prep.n_j
        s.ph_nm = self.ph_nm.ctypes.data_as(_f32p)
        s.amp = _ptr(self.amp, _f32p)
        s.row_invariant = int(bool(prep.row_invariant_bperp))
        s.bperp_row = _ptr(self.bperp_row, _f64p)
        s.bperp_mat = _ptr(self.bperp_mat, _f64p)
        s.grid_lin = _ptr(self.grid_lin, _i64p)
        s.d_a = _ptr(self.d_a, _f64p)
        s.low_pass = _ptr(self.low_pass, _f64p)
        s.n_win_ex = int(round(float(np.sqrt(self.low_pass.size))))
        s.clap_window = int(prep.clap_window)
        s.clap_alpha, s.clap_beta = float(prep.clap_alpha), float(prep.clap_beta)
        s.trial_values = _ptr(self.trial_values, _f64p)
        s.n_trial = self.trial_values.size
        s.nr_base = _ptr(self.nr_base, _f64p)
        s.nr_max_nz_ix = float(prep.nr_max_nz_ix)
        s.coh_bins = _ptr(self.coh_bins, _f64p)
        s.low_coh_thresh = int(prep.low_coh_thresh)
        s.weighting = (_native.WEIGHT_PSQUARE if prep.filter_weighting.lower() == "p-square"
                       else _native.WEIGHT_SNR)
        s.gamma_change_convergence = float(prep.gamma_change_convergence)
        s.gamma_max_iterations = int(prep.gamma_max_iterations)
        self.s = s


def _opts(form: str, device: int, rescore_bound: float, verbose: bool,
          autotune_cache: Optional[str] = None, autotune: bool = False) -> _native.RunOpts:
    if form not in FORMS:
        raise Stage2Error(f"unknown stage-2 form '{form}' (use fp32 or fp64)")
    o = _native.RunOpts()
    o.form = FORMS[form]
    o.device = int(device)
    o.rescore_bound = float(rescore_bound)
    o.verbose = int(bool(verbose))
    o.autotune = int(bool(autotune or autotune_cache))
    o.autotune_cache = None if autotune_cache is None else str(autotune_cache).encode()
    return o


def _check(rc: int):
    if rc != 0:
        raise Stage2Error(_native.last_error())


class _OutHolder:
    def __init__(self, prep: Stage2Prepared, loop_outputs: bool = True):
        P, I, Cc = prep.n_ps, prep.n_ifg, prep.n_i * prep.n_j
        self.d = {k: np.zeros(P) for k in ("K_ps", "C_ps", "coh_ps", "N_opt")}
        self.d["ph_res"] = np.zeros((P, I), np.float32)
        self.d["ph_patch"] = np.zeros((P, I), np.complex64)
        self.d["trial_ix"] = np.zeros(P, np.int32)
        s = _native.PatchOut()
        s.k_ps = _ptr(self.d["K_ps"], _f64p)
        s.c_ps = _ptr(self.d["C_ps"], _f64p)
        s.coh_ps = _ptr(self.d["coh_ps"], _f64p)
        s.n_opt = _ptr(self.d["N_opt"], _f64p)
        s.ph_res = _ptr(self.d["ph_res"], _f32p)
        s.ph_patch = self.d["ph_patch"].ctypes.data_as(_f32p)
        s.trial_ix = _ptr(self.d["trial_ix"], _i32p)
        if loop_outputs:
            self.d["coh_ps_save"] = np.zeros(P)
            self.d["ph_grid"] = np.zeros((Cc, I), np.complex64)
            self.d["ph_weight"] = np.zeros((P, I), np.complex64)
            self.d["Nr"] = np.zeros(_native.COH_BINS)
            s.coh_ps_save = _ptr(self.d["coh_ps_save"], _f64p)
            s.ph_grid = self.d["ph_grid"].ctypes.data_as(_f32p)
            s.ph_weight = self.d["ph_weight"].ctypes.data_as(_f32p)
            s.nr = _ptr(self.d["Nr"], _f64p)
        self.s = s

    def finish(self, prep: Stage2Prepared) -> dict:
        d = self.d
        d["gamma_change_save"] = self.s.gamma_change_save
        d["i_loop"] = int(self.s.i_loop)
        d["n_rescored"] = int(self.s.n_rescored)
        d["n_rechecked"] = int(self.s.n_rechecked)
        d["loop_seconds"] = float(self.s.loop_seconds)
        d["Nr_max_nz_ix"] = prep.nr_max_nz_ix
        d["step_number"] = 1.0
        d["n_trial_wraps"] = np.float32(prep.n_trial_wraps)
        d["grid_ij"] = prep.grid_ij
        d["grid_size"] = prep.grid_size
        d["low_pass"] = prep.low_pass
        d["coh_bins"] = prep.coh_bins
        return d


def run_stage2_hip(prep: Stage2Prepared, form: str = "fp32", device: int = 0,
                   rescore_bound: float = 0.0, verbose: bool = False,
                   autotune_cache: Optional[str] = None) -> dict:
    """Whole gamma loop on the GPU; returns the pm1.mat content."""
    L = _native.lib()
    inp = _InHolder(prep)
    out = _OutHolder(prep)
    _check(L.ps2_hip_run(C.byref(inp.s), C.byref(out.s),
                         C.byref(_opts(form, device, rescore_bound, verbose, autotune_cache))))
    return out.finish(prep)


def run_stage2_patch(raw: RawPatch, form: str = "fp32", device: int = 0,
                     nr_rows: Optional[int] = None) -> tuple[str, dict]:
    """prepare + loop for one in-memory patch; returns (summary, pm1 dict)."""
    kw = {} if nr_rows is None else {"nr_rows": nr_rows}
    prep = prepare_stage2_inputs(raw, device=device, **kw)
    pm1 = run_stage2_hip(prep, form=form, device=device)
    summary = (f"Stage 2 computed coherence for {prep.n_ps} candidates in {pm1['i_loop']} "
               f"iterations using the hip backend")
    return summary, pm1


def run_sessions(sessions) -> tuple[list[int], list[float]]:
    """Runs the loops of several device-resident sessions concurrently."""
    L = _native.lib()
    n = len(sessions)
    hs = (C.c_void_p * n)(*[x.h.value for x in sessions])
    il = np.zeros(n, np.int64)
    sec = np.zeros(n)
    _check(L.ps2_hip_sessions_run(n, hs, _ptr(il, _i64p), _ptr(sec, _f64p)))
    return [int(v) for v in il], [float(v) for v in sec]


def run_stage2_hip_batch(preps, form: str = "fp32", device: int = 0,
                         rescore_bound: float = 0.0, autotune_cache: Optional[str] = None) -> list[dict]:
    """Several patches in one batched call (ps2_hip_run_batch)."""
    L = _native.lib()
    ins = [_InHolder(p) for p in preps]
    outs = [_OutHolder(p) for p in preps]
    n = len(preps)
    arr_in = (_native.PatchIn * n)(*[x.s for x in ins])
    arr_out = (_native.PatchOut * n)(*[x.s for x in outs])
    _check(L.ps2_hip_run_batch(n, arr_in, arr_out,
                               C.byref(_opts(form, device, rescore_bound, False, autotune_cache))))
    res = []
    for k, p in enumerate(preps):
        outs[k].s = arr_out[k]
        res.append(outs[k].finish(p))
    return res


class Stage2Session:
    """Device-resident patch: repeated loop runs and single steps."""

    def __init__(self, prep: Stage2Prepared, form: str = "fp32", device: int = 0,
                 rescore_bound: float = 0.0, verbose: bool = False,
                 autotune_cache: Optional[str] = None, autotune: bool = False):
        """autotune / autotune_cache: FP32 form tiles measured for the patch's
        shape (kept in the cache file when one is named); else the model's."""
        self.L = _native.lib()
        self.prep = prep
        self.form = form
        self._in = _InHolder(prep)
        self._opts = _opts(form, device, rescore_bound, verbose, autotune_cache, autotune)
        self.h = C.c_void_p()
        _check(self.L.ps2_hip_session_create(C.byref(self._in.s), C.byref(self._opts), C.byref(self.h)))

    @property
    def real_dtype(self):
        return np.float64 if self.form == "fp64" else np.float32

    @property
    def complex_dtype(self):
        return np.complex128 if self.form == "fp64" else np.complex64

    def run(self) -> tuple[int, float]:
        i_loop = C.c_int64()
        secs = C.c_double()
        _check(self.L.ps2_hip_session_run(self.h, C.byref(i_loop), C.byref(secs)))
        return int(i_loop.value), float(secs.value)

    def tiles(self) -> dict:
        t = _native.Tiles()
        _check(self.L.ps2_hip_session_tiles(self.h, C.byref(t)))
        return {"fill_cells": t.fill_cells, "scan_step_table": bool(t.scan_step_table),
                "clap_segments": t.clap_segments, "clap_windows_per_segment": t.clap_windows_per_segment,
                "source": _native.TILE_SOURCES.get(t.source, str(t.source))}

    def set_timing(self, enable: bool = True):
        _check(self.L.ps2_hip_session_set_timing(self.h, int(bool(enable))))

    def kernel_times(self) -> dict:
        n = len(_native.KERNEL_SLOTS)
        ms = np.zeros(n)
        cnt = np.zeros(n, np.int64)
        _check(self.L.ps2_hip_session_kernel_times(self.h, _ptr(ms, _f64p), _ptr(cnt, _i64p)))
        return {name: (float(ms[k]), int(cnt[k])) for k, name in enumerate(_native.KERNEL_SLOTS)}

    def run_iterations(self, on_iteration) -> int:
        """The loop with the state after every iteration:
        on_iteration(iteration, pm1_dict) is called after each gamma iteration
        with that state (the pm1.mat content at that point; the arrays are
        reused by the next call) -- the reference's
        PYSTAMPS_STAGE2_NATIVE_DEBUG_PM snapshots.  Returns the iteration count."""
        out = _OutHolder(self.prep)
        err = []

        def cb(user, iteration, state):
            try:
                out.s = state.contents
                on_iteration(int(iteration), out.finish(self.prep))
                return 0
            except Exception as e:   # reported after the call returns
                err.append(e)
                return 1

        fn = _native.ITERATION_FN(cb)
        il = C.c_int64(0)
        rc = self.L.ps2_hip_session_run_iterations(self.h, C.byref(out.s), fn, None, C.byref(il))
        if err:
            raise err[0]
        _check(rc)
        return int(il.value)

    def download(self) -> dict:
        out = _OutHolder(self.prep)
        _check(self.L.ps2_hip_session_download(self.h, C.byref(out.s)))
        return out.finish(self.prep)

    def export_device(self) -> dict:
        """Stage-3 inputs left on the device in row order (device pointers as
        ints, valid until the session's next run, export or close)."""
        d = _native.DeviceOutputs()
        _check(self.L.ps2_hip_session_export_device(self.h, C.byref(d)))
        return {"device": int(d.device), "n_ps": int(d.n_ps), "n_ifg": int(d.n_ifg),
                "K_ps": d.k_ps, "C_ps": d.c_ps, "coh_ps": d.coh_ps, "N_opt": d.n_opt,
                "ph_patch": d.ph_patch, "ph_res": d.ph_res, "Nr": np.array(d.nr[:]),
                "gamma_change_save": float(d.gamma_change_save), "i_loop": int(d.i_loop)}

    def fill_grid(self, k_ps, weighting) -> np.ndarray:
        k = np.ascontiguousarray(k_ps, dtype=np.float64)
        w = np.ascontiguousarray(weighting, dtype=np.float64)
        g = np.zeros((self.prep.n_i * self.prep.n_j, self.prep.n_ifg), self.complex_dtype)
        _check(self.L.ps2_hip_step_fill_grid(self.h, _ptr(k, _f64p), _ptr(w, _f64p), g.ctypes.data))
        return g

    def clap(self, ph_grid) -> np.ndarray:
        g = np.ascontiguousarray(ph_grid, dtype=self.complex_dtype)
        f = np.zeros_like(g)
        _check(self.L.ps2_hip_step_clap(self.h, g.ctypes.data, f.ctypes.data))
        return f

    def fit(self, ph_filt) -> dict:
        f = np.ascontiguousarray(ph_filt, dtype=self.complex_dtype)
        out = _OutHolder(self.prep, loop_outputs=False)
        _check(self.L.ps2_hip_step_fit(self.h, f.ctypes.data, C.byref(out.s)))
        d = out.d
        d["n_rescored"] = int(out.s.n_rescored)
        d["n_rechecked"] = int(out.s.n_rechecked)
        return d

    def close(self):
        if self.h:
            self.L.ps2_hip_session_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
