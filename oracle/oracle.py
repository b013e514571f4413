"""ctypes wrapper of the stage-2 oracle (oracle/liboracle_stage2.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py``, as the checker or the timed CPU
baseline.  The product path never imports it.

The oracle restates crates/pystamps-core/src/native_stage2.rs in C; see
oracle/stage2_oracle.c for the function-by-function citations and what pins
each function.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle_stage2.so"

REF32 = 32
REF64 = 64
COH_BINS = 100

_f64p = C.POINTER(C.c_double)
_f32p = C.POINTER(C.c_float)
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)


class _In(C.Structure):
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


class _Out(C.Structure):
    _fields_ = [
        ("k_ps", _f64p), ("c_ps", _f64p), ("coh_ps", _f64p), ("n_opt", _f64p),
        ("coh_ps_save", _f64p),
        ("ph_res", C.c_void_p), ("ph_patch", C.c_void_p), ("ph_grid", C.c_void_p),
        ("ph_weight", C.c_void_p), ("nr", _f64p), ("gamma_change_save", C.c_double),
        ("i_loop", C.c_int64), ("trial_ix", _i32p), ("trial_margin", _f32p),
        ("weighting_last", _f64p), ("k_prev", _f64p), ("trial_margin_min", _f32p),
        ("decision_margin_min", _f64p), ("refine_margin_min", _f64p),
    ]


def build() -> Path:
    """Compile the oracle with its Makefile (gcc only)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = _bind(C.CDLL(str(LIB_PATH)))
    return _lib


def build_variant(out_path, defines=()) -> Path:
    """Compile the oracle with extra -D defines (the planted-mistake builds of
    tests/test_oracle_pins.py) into out_path; the default library is untouched."""
    out_path = Path(out_path)
    subprocess.run(["make", "-s", "-B", "-C", str(HERE), f"LIB={out_path}",
                    "EXTRA=" + " ".join(f"-D{d}" for d in defines)], check=True)
    return out_path


class use_library:
    """Context manager: the module's functions call the library at `path`
    (a build_variant output) instead of the default oracle."""

    def __init__(self, path):
        self.path = str(path)

    def __enter__(self):
        global _lib
        self.saved = _lib
        _lib = _bind(C.CDLL(self.path))
        return _lib

    def __exit__(self, *exc):
        global _lib
        _lib = self.saved
        return False


def _bind(L):
    L.ora_stage2_run.argtypes = [C.POINTER(_In), C.POINTER(_Out), C.c_int, C.c_int]
    L.ora_fill_phase_weight.argtypes = [C.POINTER(_In), _f64p, _f64p, C.c_void_p, C.c_int, C.c_int]
    L.ora_accumulate_grid.argtypes = [C.POINTER(_In), C.c_void_p, C.c_void_p, C.c_int]
    L.ora_clap_filter_grid_stack.argtypes = [C.POINTER(_In), C.c_void_p, C.c_void_p, C.c_int, C.c_int]
    L.ora_extract_patch_phase.argtypes = [C.POINTER(_In), C.c_void_p, C.c_void_p, C.c_int, C.c_int]
    L.ora_fit_rows.argtypes = [C.POINTER(_In), C.c_void_p, _f64p, _f64p, _f64p, _f64p,
                               C.c_void_p, _i32p, _f32p, C.c_int, C.c_int]
    L.ora_rms_difference.argtypes = [_f64p, _f64p, C.c_int64]
    L.ora_rms_difference.restype = C.c_double
    L.ora_hist_with_centers.argtypes = [_f64p, C.c_int64, _f64p, C.c_int64, _f64p]
    L.ora_psquare_weighting.argtypes = [_f64p, _f64p, C.c_int64, C.c_int64, C.c_double,
                                        _f64p, C.c_int64, _f64p, _f64p]
    L.ora_rng_uniform.argtypes = [C.c_uint32, C.c_int64, _f64p]
    L.ora_trial_values.argtypes = [C.c_double, _f64p, C.c_int64]
    L.ora_trial_values.restype = C.c_int64
    L.ora_low_pass.argtypes = [C.c_double, C.c_double, C.c_double, _f64p, _i64p]
    L.ora_grid_indices.argtypes = [_f32p, C.c_int64, C.c_double, _f32p, _i64p, _i64p]
    L.ora_clap_window_weight.argtypes = [C.c_int64, _f64p]
    L.ora_random_coherence_histogram.argtypes = [_f64p, C.c_int64, _f64p, C.c_int64, _f64p,
                                                 C.c_int64, C.c_uint32, _f64p, _f64p, C.c_int]
    L.ora_random_coherence_histogram_sb.argtypes = [_f64p, C.c_int64, _i32p, _i32p, C.c_int64,
                                                    _f64p, C.c_int64, _f64p, C.c_int64,
                                                    C.c_uint32, _f64p, _f64p, C.c_int]
    L.ora_py_near_max_candidates.argtypes = [_f64p, C.c_int64, _i64p]
    L.ora_py_near_max_candidates.restype = C.c_int64
    L.ora_py_select_candidate.argtypes = [_i64p, _f64p, _f64p, C.c_int64, C.c_int64]
    L.ora_py_select_candidate.restype = C.c_int64
    L.ora_py_topofit_rows.argtypes = [_f64p, _f64p, C.c_int, C.c_int64, C.c_int64, C.c_double,
                                      _f64p, _f64p, _f64p, _f64p]
    L.ora_py_topofit_coh_row_invariant.argtypes = [_f64p, _f64p, C.c_int64, C.c_int64, C.c_double,
                                                   _f64p, C.POINTER(C.c_int8)]
    L.ora_py_topofit_coh_row_invariant.restype = None
    L.ora_py_grid_accumulate.argtypes = [_f32p, _i64p, C.c_int64, C.c_int64, C.c_int64, _f32p]
    L.ora_py_grid_accumulate.restype = C.c_int
    L.ora_topofit_phase_row_coh.argtypes = [_f64p, _f64p, C.c_int64, _f64p, C.c_int64]
    L.ora_topofit_phase_row_coh.restype = C.c_double
    L.ora_fft32_2d.argtypes = [_f64p, C.c_int]
    L.ora_snr_weighting.argtypes = [_f32p, C.c_void_p, C.c_int64, C.c_int64, _f64p, C.c_int]
    return L


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


class Inputs:
    """Holds the contiguous arrays behind one ``_In`` struct (keeps them alive)."""

    def __init__(self, prep):
        self.prep = prep
        self.ph_nm = _c(prep.ph_nm, np.complex64)
        self.amp = _c(prep.amp, np.float32)
        self.bperp_row = _c(prep.row_bperp_nm, np.float64)
        self.bperp_mat = None if prep.bperp_mat is None else _c(prep.bperp_mat, np.float64)
        self.grid_lin = _c(prep.grid_lin, np.int64)
        self.d_a = _c(prep.d_a, np.float64)
        self.low_pass = _c(prep.low_pass, np.float64)
        self.trial_values = _c(prep.trial_values, np.float64)
        self.nr_base = _c(prep.nr_base, np.float64)
        self.coh_bins = _c(prep.coh_bins, np.float64)
        s = _In()
        s.n_ps, s.n_ifg = prep.n_ps, prep.n_ifg
        s.n_i, s.n_j = prep.n_i, prep.n_j
        s.ph_nm = self.ph_nm.ctypes.data_as(_f32p)
        s.amp = _p(self.amp, _f32p)
        s.row_invariant = int(bool(prep.row_invariant_bperp))
        s.bperp_row = _p(self.bperp_row, _f64p)
        s.bperp_mat = _p(self.bperp_mat, _f64p)
        s.grid_lin = _p(self.grid_lin, _i64p)
        s.d_a = _p(self.d_a, _f64p)
        s.low_pass = _p(self.low_pass, _f64p)
        s.n_win_ex = int(round(np.sqrt(self.low_pass.size)))
        s.clap_window = prep.clap_window
        s.clap_alpha, s.clap_beta = prep.clap_alpha, prep.clap_beta
        s.trial_values = _p(self.trial_values, _f64p)
        s.n_trial = self.trial_values.size
        s.nr_base = _p(self.nr_base, _f64p)
        s.nr_max_nz_ix = prep.nr_max_nz_ix
        s.coh_bins = _p(self.coh_bins, _f64p)
        s.low_coh_thresh = prep.low_coh_thresh
        s.weighting = 0 if prep.filter_weighting.lower() == "p-square" else 1
        s.gamma_change_convergence = prep.gamma_change_convergence
        s.gamma_max_iterations = prep.gamma_max_iterations
        self.s = s


def _rdt(mode):
    return (np.float64, np.complex128) if mode == REF64 else (np.float32, np.complex64)


def run(prep, mode=REF32, threads=0, diagnostics=True):
    """Whole gamma loop.  Returns a dict keyed like pm1.mat (+ diagnostics)."""
    L = lib()
    inp = Inputs(prep)
    P, I, Cc = prep.n_ps, prep.n_ifg, prep.n_i * prep.n_j
    rt, ct = _rdt(mode)
    o = {k: np.zeros(P) for k in ("K_ps", "C_ps", "coh_ps", "N_opt", "coh_ps_save")}
    o["ph_res"] = np.zeros((P, I), rt)
    o["ph_patch"] = np.zeros((P, I), ct)
    o["ph_grid"] = np.zeros((Cc, I), ct)
    o["ph_weight"] = np.zeros((P, I), ct)
    o["Nr"] = np.zeros(COH_BINS)
    if diagnostics:
        o["trial_ix"] = np.zeros(P, np.int32)
        o["trial_margin"] = np.zeros(P, np.float32)
        o["weighting_last"] = np.zeros(P)
        o["k_prev"] = np.zeros(P)
        o["trial_margin_min"] = np.zeros(P, np.float32)
        o["decision_margin_min"] = np.zeros(P)
        o["refine_margin_min"] = np.zeros(P)
    s = _Out()
    s.k_ps, s.c_ps = _p(o["K_ps"], _f64p), _p(o["C_ps"], _f64p)
    s.coh_ps, s.n_opt = _p(o["coh_ps"], _f64p), _p(o["N_opt"], _f64p)
    s.coh_ps_save = _p(o["coh_ps_save"], _f64p)
    s.ph_res = o["ph_res"].ctypes.data
    s.ph_patch = o["ph_patch"].ctypes.data
    s.ph_grid = o["ph_grid"].ctypes.data
    s.ph_weight = o["ph_weight"].ctypes.data
    s.nr = _p(o["Nr"], _f64p)
    if diagnostics:
        s.trial_ix = _p(o["trial_ix"], _i32p)
        s.trial_margin = _p(o["trial_margin"], _f32p)
        s.weighting_last = _p(o["weighting_last"], _f64p)
        s.k_prev = _p(o["k_prev"], _f64p)
        s.trial_margin_min = _p(o["trial_margin_min"], _f32p)
        s.decision_margin_min = _p(o["decision_margin_min"], _f64p)
        s.refine_margin_min = _p(o["refine_margin_min"], _f64p)
    rc = L.ora_stage2_run(C.byref(inp.s), C.byref(s), mode, threads)
    if rc != 0:
        raise RuntimeError(f"oracle stage 2 failed with code {rc}")
    o["gamma_change_save"] = s.gamma_change_save
    o["i_loop"] = s.i_loop
    o["Nr_max_nz_ix"] = prep.nr_max_nz_ix
    return o


# --- single steps (state injection) ----------------------------------------

def fill_phase_weight(prep, k_ps, weighting, mode=REF32, threads=0):
    inp = Inputs(prep)
    _, ct = _rdt(mode)
    out = np.zeros((prep.n_ps, prep.n_ifg), ct)
    k = _c(k_ps, np.float64)
    w = _c(weighting, np.float64)
    rc = lib().ora_fill_phase_weight(C.byref(inp.s), _p(k, _f64p), _p(w, _f64p), out.ctypes.data,
                                     mode, threads)
    if rc:
        raise RuntimeError(rc)
    return out


def accumulate_grid(prep, ph_weight, mode=REF32):
    inp = Inputs(prep)
    _, ct = _rdt(mode)
    pw = _c(ph_weight, ct)
    out = np.zeros((prep.n_i * prep.n_j, prep.n_ifg), ct)
    lib().ora_accumulate_grid(C.byref(inp.s), pw.ctypes.data, out.ctypes.data, mode)
    return out


def clap_filter_grid_stack(prep, ph_grid, mode=REF32, threads=0):
    inp = Inputs(prep)
    _, ct = _rdt(mode)
    g = _c(ph_grid, ct)
    out = np.zeros((prep.n_i * prep.n_j, prep.n_ifg), ct)
    lib().ora_clap_filter_grid_stack(C.byref(inp.s), g.ctypes.data, out.ctypes.data, mode, threads)
    return out


def extract_patch_phase(prep, ph_filt, mode=REF32, threads=0):
    inp = Inputs(prep)
    _, ct = _rdt(mode)
    f = _c(ph_filt, ct)
    out = np.zeros((prep.n_ps, prep.n_ifg), ct)
    lib().ora_extract_patch_phase(C.byref(inp.s), f.ctypes.data, out.ctypes.data, mode, threads)
    return out


def fit_rows(prep, ph_patch, mode=REF32, threads=0):
    inp = Inputs(prep)
    rt, ct = _rdt(mode)
    pp = _c(ph_patch, ct)
    P, I = prep.n_ps, prep.n_ifg
    o = {k: np.zeros(P) for k in ("K_ps", "C_ps", "coh_ps", "N_opt")}
    o["ph_res"] = np.zeros((P, I), rt)
    o["trial_ix"] = np.zeros(P, np.int32)
    o["trial_margin"] = np.zeros(P, np.float32)
    rc = lib().ora_fit_rows(C.byref(inp.s), pp.ctypes.data, _p(o["K_ps"], _f64p),
                            _p(o["C_ps"], _f64p), _p(o["coh_ps"], _f64p), _p(o["N_opt"], _f64p),
                            o["ph_res"].ctypes.data, _p(o["trial_ix"], _i32p),
                            _p(o["trial_margin"], _f32p), mode, threads)
    if rc:
        raise RuntimeError(rc)
    return o


# --- host-sized helpers ----------------------------------------------------

def rng_uniform(n, seed=2005):
    out = np.zeros(n)
    lib().ora_rng_uniform(seed, n, _p(out, _f64p))
    return out


def hist_with_centers(values, centers):
    v = _c(values, np.float64)
    c = _c(centers, np.float64)
    out = np.zeros(c.size)
    lib().ora_hist_with_centers(_p(v, _f64p), v.size, _p(c, _f64p), c.size, _p(out, _f64p))
    return out


def psquare_weighting(nr, na, low_coh_thresh, nr_max_nz_ix, coh_ps):
    nr = _c(nr, np.float64)
    na = _c(na, np.float64)
    coh = _c(coh_ps, np.float64)
    w = np.zeros(coh.size)
    hi = np.zeros(max(nr.size + 1, 1) * 10)
    lib().ora_psquare_weighting(_p(nr, _f64p), _p(na, _f64p), nr.size, low_coh_thresh,
                                float(nr_max_nz_ix), _p(coh, _f64p), coh.size, _p(w, _f64p),
                                _p(hi, _f64p))
    keep = max((nr.size + 1) * 10 - 9, 0)
    return w, hi[:keep]


def trial_values(n_trial_wraps):
    out = np.zeros(4096)
    n = lib().ora_trial_values(float(n_trial_wraps), _p(out, _f64p), out.size)
    return out[:n].copy()


def low_pass(clap_win=32.0, low_pass_wavelength=800.0, grid_size=50.0):
    n = int(max(round(clap_win), 1))
    out = np.zeros(n * n)
    dim = C.c_int64()
    lib().ora_low_pass(clap_win, low_pass_wavelength, grid_size, _p(out, _f64p), C.byref(dim))
    return out.reshape(n, n)


def grid_indices(xy, grid_size):
    xy = _c(xy, np.float32)
    n = xy.shape[0]
    out = np.zeros((n, 2), np.float32)
    ni, nj = C.c_int64(), C.c_int64()
    lib().ora_grid_indices(_p(xy, _f32p), n, grid_size, _p(out, _f32p), C.byref(ni), C.byref(nj))
    return out, ni.value, nj.value


def clap_window_weight(n_win=24):
    out = np.zeros(n_win * n_win)
    lib().ora_clap_window_weight(n_win, _p(out, _f64p))
    return out.reshape(n_win, n_win)


def random_coherence_histogram(bperp_nm, trial_vals, coh_bins, n_rows=300_000, seed=2005, threads=0):
    bp = _c(bperp_nm, np.float64)
    tv = _c(trial_vals, np.float64)
    cb = _c(coh_bins, np.float64)
    nr = np.zeros(COH_BINS)
    mx = C.c_double()
    lib().ora_random_coherence_histogram(_p(bp, _f64p), bp.size, _p(tv, _f64p), tv.size,
                                         _p(cb, _f64p), n_rows, seed, _p(nr, _f64p), C.byref(mx),
                                         threads)
    return nr, mx.value


def random_coherence_histogram_sb(bperp_nm, image_a, image_b, n_image, trial_vals, coh_bins,
                                  n_rows=300_000, seed=2005, threads=0):
    """Small-baseline Nr: image_a / image_b are 0-based image indices per ifg."""
    bp = _c(bperp_nm, np.float64)
    ia = _c(image_a, np.int32)
    ib = _c(image_b, np.int32)
    tv = _c(trial_vals, np.float64)
    cb = _c(coh_bins, np.float64)
    nr = np.zeros(COH_BINS)
    mx = C.c_double()
    lib().ora_random_coherence_histogram_sb(_p(bp, _f64p), bp.size, _p(ia, _i32p), _p(ib, _i32p),
                                            int(n_image), _p(tv, _f64p), tv.size, _p(cb, _f64p),
                                            n_rows, seed, _p(nr, _f64p), C.byref(mx), threads)
    return nr, mx.value


def topofit_phase_row_coh(phase, bperp, trial_vals):
    ph = _c(phase, np.float64)
    bp = _c(bperp, np.float64)
    tv = _c(trial_vals, np.float64)
    return lib().ora_topofit_phase_row_coh(_p(ph, _f64p), _p(bp, _f64p), bp.size, _p(tv, _f64p), tv.size)


def fft32_2d(x, inverse=False):
    a = np.ascontiguousarray(np.asarray(x, np.complex128).reshape(32, 32))
    buf = a.view(np.float64).copy()
    lib().ora_fft32_2d(_p(buf, _f64p), int(inverse))
    return buf.view(np.complex128).reshape(32, 32)


def snr_weighting(amp, ph_res, mode=REF64):
    """snr_weighting (native_stage2.rs:1875-1897) of amp [P][I] and ph_res [P][I]."""
    a = _c(amp, np.float32)
    rt, _ = _rdt(mode)
    pr = _c(ph_res, rt)
    P, I = a.shape
    out = np.zeros(P)
    lib().ora_snr_weighting(_p(a, _f32p), pr.ctypes.data, P, I, _p(out, _f64p), mode)
    return out


def rms_difference(a, b):
    x = _c(a, np.float64)
    y = _c(b, np.float64)
    return lib().ora_rms_difference(_p(x, _f64p), _p(y, _f64p), x.size)


def cpu_threads_default():
    return os.cpu_count() or 1


# --- Python-pipeline registry kernels (pytopofit_oracle.c) ---------------------

def py_near_max_candidates(coh_trial):
    c = _c(coh_trial, np.float64)
    out = np.zeros(max(c.size, 1), np.int64)
    n = lib().ora_py_near_max_candidates(_p(c, _f64p), c.size, _p(out, _i64p))
    return out[:n]


def py_select_candidate(cand, coarse, refined, trial_count):
    ca = _c(cand, np.int64)
    co = _c(coarse, np.float64)
    rf = _c(refined, np.float64)
    return int(lib().ora_py_select_candidate(_p(ca, _i64p), _p(co, _f64p), _p(rf, _f64p), ca.size,
                                             int(trial_count)))


def py_topofit_rows(cpxphase, bperp, n_trial_wraps):
    """stage2_topofit / _row_invariant semantics in FP64: (K, C, coh, residual)."""
    cpx = np.ascontiguousarray(cpxphase, dtype=np.complex128)
    P, I = cpx.shape
    bp = np.asarray(bperp, dtype=np.float64)
    rowinv = bp.ndim == 1
    bp = np.ascontiguousarray(bp)
    K, Cc, coh = np.zeros(P), np.zeros(P), np.zeros(P)
    res = np.zeros((P, I), np.complex128)
    lib().ora_py_topofit_rows(cpx.ctypes.data_as(_f64p), _p(bp, _f64p), int(rowinv), P, I,
                              float(n_trial_wraps), _p(K, _f64p), _p(Cc, _f64p), _p(coh, _f64p),
                              res.ctypes.data_as(_f64p))
    return K, Cc, coh, res


def py_topofit_coh_row_invariant(cpxphase, bperp_vec, n_trial_wraps):
    """stage2_topofit_coh_row_invariant in FP64: (coh, single) where single
    marks the rows refined from the whole-vector trial grid."""
    cpx = np.ascontiguousarray(cpxphase, dtype=np.complex128)
    P, I = cpx.shape
    bp = _c(np.asarray(bperp_vec, dtype=np.float64).reshape(-1), np.float64)
    coh = np.zeros(P)
    single = np.zeros(P, np.int8)
    lib().ora_py_topofit_coh_row_invariant(cpx.ctypes.data_as(_f64p), _p(bp, _f64p), P, I,
                                           float(n_trial_wraps), _p(coh, _f64p),
                                           single.ctypes.data_as(C.POINTER(C.c_int8)))
    return coh, single.astype(bool)


def py_grid_accumulate(ph_weight, grid_lin, n_i, n_j):
    ph = np.ascontiguousarray(ph_weight, dtype=np.complex64)
    gl = _c(grid_lin, np.int64)
    P, I = ph.shape
    out = np.zeros((n_i * n_j, I), np.complex64)
    rc = lib().ora_py_grid_accumulate(ph.ctypes.data_as(_f32p), _p(gl, _i64p), P, I, n_i * n_j,
                                      out.ctypes.data_as(_f32p))
    if rc != 0:
        raise IndexError("grid index out of range")
    return out.reshape(n_i, n_j, I)
