"""Stage-2 prep: the host side that feeds the HIP gamma loop.

Mirrors ``prepare_stage2_inputs`` (crates/pystamps-core/src/native_stage2.rs:
415-561) and the parameter loaders (:2105-2128) on in-memory arrays instead of
MAT files, with the reference's f32/f64 arithmetic (numpy float32 ops round
like Rust f32; float32 ``np.hypot`` is libm ``hypotf``).  The random-coherence
histogram Nr is computed on the GPU through the C ABI
(``ps2_hip_random_coherence_histogram``).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from typing import Callable, Optional

import numpy as np

from . import _native

# native_stage2.rs:12-29
DEFAULT_GRID_SIZE = 50.0
DEFAULT_CLAP_WIN = 32.0
DEFAULT_CLAP_LOW_PASS_WAVELENGTH = 800.0
DEFAULT_CLAP_ALPHA = 1.0
DEFAULT_CLAP_BETA = 0.3
DEFAULT_MAX_TOPO_ERR = 15.0
DEFAULT_LAMBDA_M = 0.0555
DEFAULT_MEAN_INCIDENCE = math.radians(23.0)
COH_BIN_COUNT = 100
COH_BIN_START = 0.005
COH_BIN_STEP = 0.01
STAGE2_RANDOM_SEED = 2005
STAGE2_RANDOM_COUNT = 300_000
TRIAL_RHO_M = 830_000.0


class Stage2Error(RuntimeError):
    """Stage-2 failure; displays like CoreError::NativeStage{stage: 2} (lib.rs:47-48)."""

    def __init__(self, message: str):
        self.message = message
        super().__init__(f"stage 2 native implementation error: {message}")


@dataclass
class Stage2Options:
    grid_size: float = DEFAULT_GRID_SIZE
    clap_win: float = DEFAULT_CLAP_WIN
    clap_low_pass_wavelength: float = DEFAULT_CLAP_LOW_PASS_WAVELENGTH
    clap_alpha: float = DEFAULT_CLAP_ALPHA
    clap_beta: float = DEFAULT_CLAP_BETA
    max_topo_err: float = DEFAULT_MAX_TOPO_ERR
    lambda_m: float = DEFAULT_LAMBDA_M


@dataclass
class Stage2Parms:
    small_baseline_flag: str = "n"
    filter_weighting: str = "P-square"
    gamma_change_convergence: float = 1.0e-4
    gamma_max_iterations: int = 25


def _round_half_away(v: float) -> float:
    return math.floor(v + 0.5) if v >= 0.0 else math.ceil(v - 0.5)


def load_stage2_options(parms: dict) -> Stage2Options:
    """load_stage2_options, native_stage2.rs:2105-2116."""
    o = Stage2Options()
    o.grid_size = float(parms.get("filter_grid_size", o.grid_size))
    o.clap_win = float(parms.get("clap_win", o.clap_win))
    o.clap_low_pass_wavelength = float(parms.get("clap_low_pass_wavelength", o.clap_low_pass_wavelength))
    o.clap_alpha = float(parms.get("clap_alpha", o.clap_alpha))
    o.clap_beta = float(parms.get("clap_beta", o.clap_beta))
    o.max_topo_err = float(parms.get("max_topo_err", o.max_topo_err))
    o.lambda_m = float(parms.get("lambda", o.lambda_m))
    return o


def load_stage2_parms(parms: dict) -> Stage2Parms:
    """load_stage2_parms, native_stage2.rs:2118-2128."""
    p = Stage2Parms()
    sb = str(parms.get("small_baseline_flag", "") or "").strip()
    p.small_baseline_flag = sb or p.small_baseline_flag
    fw = str(parms.get("filter_weighting", "") or "").strip()
    p.filter_weighting = fw or p.filter_weighting
    p.gamma_change_convergence = float(parms.get("gamma_change_convergence", p.gamma_change_convergence))
    p.gamma_max_iterations = int(_round_half_away(float(parms.get("gamma_max_iterations",
                                                                   p.gamma_max_iterations))))
    return p


@dataclass
class RawPatch:
    """In-memory content of a PATCH_* directory's stage-2 inputs."""

    ph: np.ndarray                      # ph1.ph, complex64 [n_ps][n_ifg_full]
    bperp: np.ndarray                   # ps1.bperp [n_ifg_full]
    xy: np.ndarray                      # ps1.xy float32 [n_ps][3] (id, x, y)
    master_ix: int = 1                  # ps1.master_ix (1-based)
    bperp_mat: Optional[np.ndarray] = None   # bp1.bperp_mat
    d_a: Optional[np.ndarray] = None         # da1.D_A
    mean_incidence: Optional[float] = None   # ps1.mean_incidence (rad)
    inc: Optional[np.ndarray] = None         # inc1.inc
    la: Optional[np.ndarray] = None          # la1.la
    parms: dict = field(default_factory=dict)
    ifgday_ix: Optional[np.ndarray] = None   # ps1.ifgday_ix (small-baseline only)


@dataclass
class Stage2Prepared:
    """Mirror of Stage2Prepared (native_stage2.rs:159-182) plus the options and
    parms the loop reads."""

    n_ps: int
    n_ifg: int
    ph_nm: np.ndarray          # complex64 [n_ps][n_ifg]
    amp: np.ndarray            # float32 [n_ps][n_ifg]
    bperp_mat: Optional[np.ndarray]
    row_invariant_bperp: bool
    row_bperp_nm: np.ndarray
    grid_ij: np.ndarray        # float32 [n_ps][2]
    grid_lin: np.ndarray       # int64 [n_ps]
    n_i: int
    n_j: int
    d_a: np.ndarray
    low_pass: np.ndarray       # [clap_win][clap_win]
    coh_bins: np.ndarray
    nr_base: np.ndarray
    nr_max_nz_ix: float
    n_trial_wraps: float
    trial_values: np.ndarray
    grid_size: float
    clap_window: int
    low_coh_thresh: int
    clap_alpha: float
    clap_beta: float
    filter_weighting: str
    gamma_change_convergence: float
    gamma_max_iterations: int
    small_baseline: bool = False
    bperp_nm: Optional[np.ndarray] = None


def stage2_grid_indices(xy: np.ndarray, grid_size: float) -> np.ndarray:
    """stage2_grid_indices, native_stage2.rs:1971-2011 (f32 arithmetic)."""
    xy = np.asarray(xy, dtype=np.float32)
    x = xy[:, 1]
    y = xy[:, 2]
    min_x = np.fmin.reduce(x) if x.size else np.float32(np.inf)
    min_y = np.fmin.reduce(y) if y.size else np.float32(np.inf)
    scale = np.float32(grid_size)
    eps = np.float32(1.0e-6)
    with np.errstate(invalid="ignore"):
        fi = np.ceil((y - min_y + eps) / scale)
        fj = np.ceil((x - min_x + eps) / scale)
    fi = np.where(fi >= np.float32(1.0), fi, np.float32(1.0))
    fj = np.where(fj >= np.float32(1.0), fj, np.float32(1.0))
    i = np.clip(fi, -2**31, 2**31 - 1).astype(np.int64)
    j = np.clip(fj, -2**31, 2**31 - 1).astype(np.int64)
    max_i = max(1, int(i.max()) if i.size else 1)
    max_j = max(1, int(j.max()) if j.size else 1)
    if max_i > 1 or max_j > 1:
        if max_i > 1:
            i[i == max_i] = max_i - 1
        if max_j > 1:
            j[j == max_j] = max_j - 1
    return np.stack([i, j], axis=1).astype(np.float32)


def build_low_pass(options: Stage2Options) -> np.ndarray:
    """build_low_pass, native_stage2.rs:2013-2043."""
    n = int(max(_round_half_away(options.clap_win), 1.0))
    freq0 = 1.0 / options.clap_low_pass_wavelength
    butter = np.empty(n)
    for ix in range(n):
        fi = (ix - n / 2.0) / (options.grid_size * n)
        q = fi / freq0
        q2 = q * q
        q8 = (q2 * q2) * (q2 * q2)
        butter[ix] = 1.0 / (1.0 + q8 * q2)
    raw = butter[:, None] * butter[None, :]
    sh = n // 2
    return np.roll(np.roll(raw, -sh, axis=0), -sh, axis=1)


def trial_values(n_trial_wraps: float) -> np.ndarray:
    """trial_values, native_stage2.rs:1305-1308."""
    tn = int(math.ceil(8.0 * n_trial_wraps))
    return np.arange(-tn, tn + 1, dtype=np.float64)


def coh_bins() -> np.ndarray:
    return np.array([COH_BIN_START + COH_BIN_STEP * float(k) for k in range(COH_BIN_COUNT)])


def trial_wrap_mean_incidence(raw: RawPatch) -> float:
    """stage2_trial_wrap_mean_incidence, native_stage2.rs:2130-2153."""
    for values, is_la, offset in ((raw.inc, False, 0.0), (raw.la, True, 0.052)):
        if values is None:
            continue
        v = np.asarray(values, dtype=np.float64).ravel()
        valid = [float(a) for a in v if math.isfinite(a) and (a != 0.0 or is_la)]
        if valid:
            s = 0.0
            for a in valid:
                s += a
            return s / len(valid) + offset
    if raw.mean_incidence is not None:
        return float(raw.mean_incidence) + 0.052
    return DEFAULT_MEAN_INCIDENCE


def _bperp_rows_invariant(mat: Optional[np.ndarray]) -> bool:
    if mat is None or mat.shape[0] <= 1:
        return True
    # per-pixel baselines usually differ within the first rows: decide there
    # before comparing the whole matrix
    if not np.all(mat[1:17] == mat[0:1]):
        return False
    return bool(np.all(mat[17:] == mat[0:1]))


def nr_hip(bperp_nm, trials, bins, n_rows=STAGE2_RANDOM_COUNT, seed=STAGE2_RANDOM_SEED, device=0):
    """Nr on the GPU through the C ABI."""
    L = _native.lib()
    bp = np.ascontiguousarray(bperp_nm, dtype=np.float64)
    tv = np.ascontiguousarray(trials, dtype=np.float64)
    cb = np.ascontiguousarray(bins, dtype=np.float64)
    nr = np.zeros(COH_BIN_COUNT)
    mx = C.c_double()
    rc = L.ps2_hip_random_coherence_histogram(
        bp.ctypes.data_as(C.POINTER(C.c_double)), bp.size,
        tv.ctypes.data_as(C.POINTER(C.c_double)), tv.size,
        cb.ctypes.data_as(C.POINTER(C.c_double)), int(n_rows), int(seed), int(device),
        nr.ctypes.data_as(C.POINTER(C.c_double)), C.byref(mx))
    if rc != 0:
        raise Stage2Error(_native.last_error())
    return nr, mx.value


def nr_hip_sb(bperp_nm, image_a, image_b, n_image, trials, bins, n_rows=STAGE2_RANDOM_COUNT,
              seed=STAGE2_RANDOM_SEED, device=0):
    """Small-baseline Nr on the GPU (0-based image indices per interferogram)."""
    L = _native.lib()
    bp = np.ascontiguousarray(bperp_nm, dtype=np.float64)
    ia = np.ascontiguousarray(image_a, dtype=np.int32)
    ib = np.ascontiguousarray(image_b, dtype=np.int32)
    tv = np.ascontiguousarray(trials, dtype=np.float64)
    cb = np.ascontiguousarray(bins, dtype=np.float64)
    nr = np.zeros(COH_BIN_COUNT)
    mx = C.c_double()
    i32 = C.POINTER(C.c_int32)
    f64 = C.POINTER(C.c_double)
    rc = L.ps2_hip_random_coherence_histogram_sb(
        bp.ctypes.data_as(f64), bp.size, ia.ctypes.data_as(i32), ib.ctypes.data_as(i32),
        int(n_image), tv.ctypes.data_as(f64), tv.size, cb.ctypes.data_as(f64), int(n_rows),
        int(seed), int(device), nr.ctypes.data_as(f64), C.byref(mx))
    if rc != 0:
        raise Stage2Error(_native.last_error())
    return nr, mx.value


def small_baseline_random_indices(ifgday_ix, n_ifg: int):
    """small_baseline_random_indices, native_stage2.rs:1358-1403: 0-based image
    indices (a, b) per interferogram and the image count from ps1.ifgday_ix."""
    if ifgday_ix is None:
        raise Stage2Error("ps1.ifgday_ix is missing or invalid: variable not found")
    m = np.asarray(ifgday_ix, dtype=np.float64)
    if m.ndim == 1:
        m = m.reshape(1, -1)
    if m.shape == (n_ifg, 2):
        pass
    elif m.shape == (2, n_ifg):
        m = m.T
    else:
        raise Stage2Error(f"ps1.ifgday_ix has incompatible shape {m.shape[0]}x{m.shape[1]} "
                          f"for n_ifg={n_ifg}")
    a = np.array([_round_half_away(v) for v in m[:, 0]], dtype=np.int64)
    b = np.array([_round_half_away(v) for v in m[:, 1]], dtype=np.int64)
    if np.any(a <= 0) or np.any(b <= 0):
        raise Stage2Error("ps1.ifgday_ix must contain positive one-based image ids")
    n_image = int(max(a.max(initial=0), b.max(initial=0)))
    return (a - 1).astype(np.int32), (b - 1).astype(np.int32), n_image


NrFn = Callable[..., tuple]


def _threads() -> int:
    import os
    return max(1, min(os.cpu_count() or 1, 16))


def normalize_phase(ph_full: np.ndarray, cols) -> tuple[np.ndarray, np.ndarray]:
    """ph_nm and amp of the kept columns (native_stage2.rs:449-467) with the
    library's threaded host routine; f32 results equal numpy's hypot and
    divisions."""
    ph = np.ascontiguousarray(ph_full, dtype=np.complex64)
    P, n_full = ph.shape
    c = np.ascontiguousarray(cols, dtype=np.int64)
    ph_nm = np.empty((P, c.size), np.complex64)
    amp = np.empty((P, c.size), np.float32)
    f32p = C.POINTER(C.c_float)
    rc = _native.lib().ps2_hip_normalize_phase(
        ph.ctypes.data_as(f32p), P, n_full, c.ctypes.data_as(C.POINTER(C.c_int64)), c.size, _threads(),
        ph_nm.ctypes.data_as(f32p), amp.ctypes.data_as(f32p))
    if rc != 0:
        raise Stage2Error(_native.last_error())
    return ph_nm, amp


def select_columns(src: np.ndarray, cols) -> np.ndarray:
    """Row-major f64 column selection with the library's threaded routine."""
    a = np.ascontiguousarray(src, dtype=np.float64)
    c = np.ascontiguousarray(cols, dtype=np.int64)
    out = np.empty((a.shape[0], c.size), np.float64)
    f64p = C.POINTER(C.c_double)
    rc = _native.lib().ps2_hip_select_columns_f64(
        a.ctypes.data_as(f64p), a.shape[0], a.shape[1], c.ctypes.data_as(C.POINTER(C.c_int64)), c.size,
        _threads(), out.ctypes.data_as(f64p))
    if rc != 0:
        raise Stage2Error(_native.last_error())
    return out


def prepare_stage2_inputs(raw: RawPatch, nr_fn: Optional[NrFn] = None,
                          nr_rows: int = STAGE2_RANDOM_COUNT, device: int = 0,
                          nr_sb_fn: Optional[NrFn] = None) -> Stage2Prepared:
    """prepare_stage2_inputs (native_stage2.rs:415-561) on in-memory arrays.

    ``nr_fn(bperp_nm, trial_values, coh_bins, n_rows)`` computes the random
    coherence histogram and ``nr_sb_fn(bperp_nm, image_a, image_b, n_image,
    trial_values, coh_bins, n_rows)`` its small-baseline branch; the defaults
    run them on the GPU.
    """
    parms = load_stage2_parms(raw.parms)
    options = load_stage2_options(raw.parms)
    ph_full = np.asarray(raw.ph)
    if ph_full.ndim != 2 or ph_full.shape[0] == 0:
        raise Stage2Error("ps1.mat missing valid n_ps")
    n_ps, n_ifg_full = ph_full.shape
    master_ix = int(_round_half_away(float(raw.master_ix)))
    if master_ix == 0 or master_ix > n_ifg_full:
        raise Stage2Error(
            f"ps1.master_ix must be 1-based within ph1 width {n_ifg_full}; got {master_ix}")
    bperp_full = np.asarray(raw.bperp, dtype=np.float64).ravel().astype(np.float32).astype(np.float64)
    if bperp_full.size != n_ifg_full:
        raise Stage2Error(
            f"ps1.bperp has length {bperp_full.size} but ph1.ph has {n_ifg_full} interferograms")
    small_baseline = parms.small_baseline_flag.lower() == "y"
    no_master = [ix for ix in range(n_ifg_full) if small_baseline or ix != master_ix - 1]
    n_ifg = len(no_master)
    bperp_nm = bperp_full[no_master]

    # the random-coherence histogram needs only bperp and the trial grid: it
    # runs on a worker thread (its C call releases the GIL) while the phases
    # are normalised
    bins = coh_bins()
    mean_incidence = trial_wrap_mean_incidence(raw)
    max_k = options.max_topo_err / (options.lambda_m * TRIAL_RHO_M * math.sin(mean_incidence)
                                    / (4.0 * math.pi))
    min_bp = min(bperp_nm) if bperp_nm.size else math.inf
    max_bp = max(bperp_nm) if bperp_nm.size else -math.inf
    n_trial_wraps = float((max_bp - min_bp) * max_k / (2.0 * math.pi))
    trials = trial_values(n_trial_wraps)
    low_coh_thresh = 15 if small_baseline else 31
    if small_baseline:
        ia, ib, n_image = small_baseline_random_indices(raw.ifgday_ix, n_ifg)
        sb_fn = nr_sb_fn if nr_sb_fn is not None else (
            lambda bp, a, b, ni, tv, cb, n: nr_hip_sb(bp, a, b, ni, tv, cb, n, device=device))
        nr_call = lambda: sb_fn(bperp_nm, ia, ib, n_image, trials, bins, nr_rows)  # noqa: E731
    else:
        sm_fn = nr_fn if nr_fn is not None else (lambda bp, tv, cb, n: nr_hip(bp, tv, cb, n, device=device))
        nr_call = lambda: sm_fn(bperp_nm, trials, bins, nr_rows)  # noqa: E731
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=1) as pool:
        nr_future = pool.submit(nr_call)
        ph_nm, amp = normalize_phase(ph_full, no_master)
        nr_base, nr_max_nz_ix = nr_future.result()

    # load_bperp_mat, native_stage2.rs:563-612
    bperp_mat = None
    if raw.bperp_mat is None:
        if small_baseline:
            bperp_mat = np.tile(bperp_nm, (n_ps, 1))
    else:
        src = np.asarray(raw.bperp_mat, dtype=np.float64)
        if src.ndim == 1:
            src = src.reshape(-1, 1)
        if src.shape[0] != n_ps and src.shape[1] == n_ps:
            src = src.T
        if src.shape[0] != n_ps:
            raise Stage2Error(f"bp1.bperp_mat has incompatible shape {src.shape[0]}x{src.shape[1]} "
                              f"for n_ps={n_ps}")
        if src.shape[1] == n_ifg:
            bperp_mat = np.ascontiguousarray(src)
        elif not small_baseline and src.shape[1] == n_ifg_full:
            bperp_mat = select_columns(src, no_master)
        else:
            raise Stage2Error(
                f"bp1.bperp_mat has incompatible shape {src.shape[0]}x{src.shape[1]} for stage-2 "
                f"ph shape {n_ps}x{n_ifg}")
    row_invariant = _bperp_rows_invariant(bperp_mat)
    if row_invariant:
        row_bperp = bperp_mat[0].copy() if (bperp_mat is not None and bperp_mat.shape[0] > 0) else bperp_nm.copy()
    else:
        row_bperp = bperp_nm.copy()

    d_a = np.ones(n_ps)
    if raw.d_a is not None:
        v = np.asarray(raw.d_a, dtype=np.float64).ravel()
        if v.size == n_ps:
            d_a = v
    xy = np.asarray(raw.xy, dtype=np.float32)
    if xy.shape != (n_ps, 3):
        if xy.shape == (3, n_ps):
            xy = np.ascontiguousarray(xy.T)
        else:
            raise Stage2Error(f"ps1.xy has incompatible shape {xy.shape[0]}x{xy.shape[1]}; "
                              f"expected {n_ps}x3")
    grid_ij = stage2_grid_indices(xy, options.grid_size)
    n_i = max(1, int(grid_ij[:, 0].max()))
    n_j = max(1, int(grid_ij[:, 1].max()))
    grid_lin = (grid_ij[:, 0].astype(np.int64) - 1) * n_j + (grid_ij[:, 1].astype(np.int64) - 1)

    low_pass = build_low_pass(options)

    return Stage2Prepared(
        n_ps=n_ps, n_ifg=n_ifg, ph_nm=ph_nm, amp=amp, bperp_mat=bperp_mat,
        row_invariant_bperp=row_invariant, row_bperp_nm=row_bperp, grid_ij=grid_ij,
        grid_lin=grid_lin, n_i=n_i, n_j=n_j, d_a=d_a, low_pass=low_pass, coh_bins=bins,
        nr_base=np.asarray(nr_base, dtype=np.float64), nr_max_nz_ix=float(nr_max_nz_ix),
        n_trial_wraps=n_trial_wraps, trial_values=trials, grid_size=options.grid_size,
        clap_window=int(max(_round_half_away(options.clap_win * 0.75), 1.0)),
        low_coh_thresh=low_coh_thresh, clap_alpha=options.clap_alpha,
        clap_beta=options.clap_beta, filter_weighting=parms.filter_weighting,
        gamma_change_convergence=parms.gamma_change_convergence,
        gamma_max_iterations=parms.gamma_max_iterations, small_baseline=small_baseline,
        bperp_nm=bperp_nm,
    )
