"""`hip` provider for the pySTAMPS kernel registry.

The Python pipeline's stage 2 resolves five kernels through a registry
(pystamps/kernels/accelerated.py:392-426): stage2_grid_accumulate,
stage2_topofit, stage2_topofit_row_invariant,
stage2_topofit_coh_row_invariant and stage2_histogram.  The functions below
have the same signatures and return types as the reference's `python` and
`native` implementations and run on the GPU through the C ABI:

* the topofit kernels fit each row like ``_ps_topofit_single``
  (pystamps/pipeline/ported.py:2355-2430) in FP64, with its near-max candidate
  refinement and selection;
* the grid accumulation sums rows in index order in complex64, like
  ``np.add.at`` in ``_stage2_grid_accumulate_cpu`` (accelerated.py:183-200);
* the histogram is ``_stage2_histogram_with_centers_cpu``'s rule
  (accelerated.py:240-264), host code in the same library.

``register(registry)`` adds the provider to a registry with the reference's
``register_provider`` / ``register`` interface (registry.py:112-152).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native

_f64p = C.POINTER(C.c_double)
_f32p = C.POINTER(C.c_float)
_i64p = C.POINTER(C.c_int64)

KERNELS = ("stage2_grid_accumulate", "stage2_topofit", "stage2_topofit_row_invariant",
           "stage2_topofit_coh_row_invariant", "stage2_histogram")


class BackendError(RuntimeError):
    """A failure of the HIP library, with its message."""


def _check(rc: int):
    if rc != 0:
        raise BackendError(_native.last_error())


def available() -> bool:
    """True when the HIP library loads and sees a device."""
    try:
        return _native.device_count() > 0
    except Exception:  # noqa: BLE001 - availability probe
        return False


def _shared_bperp(bperp, P, I):
    """The one bperp vector of a row-invariant call (1-D, or the first row of
    a [1|n_row][n_ifg] matrix)."""
    bp = np.asarray(bperp, dtype=np.float64)
    if bp.ndim == 2:
        if bp.shape[0] not in (1, P):
            raise BackendError(f"Row-invariant stage-2 bperp has incompatible shape {bp.shape} "
                               f"for n_row={P}")
        bp = bp.reshape(-1) if bp.shape[0] == 1 else bp[0, :]
    bp = np.ascontiguousarray(bp.reshape(-1))
    if bp.size != I:
        raise BackendError(f"bperp has {bp.size} values for {I} interferograms")
    return bp


def _topofit(cpxphase, bperp, n_trial_wraps, row_invariant: bool, device: int = 0):
    cpx = np.ascontiguousarray(cpxphase, dtype=np.complex128)
    if cpx.ndim != 2:
        raise BackendError("stage-2 topofit expects a 2-D cpxphase")
    P, I = cpx.shape
    if row_invariant:
        bp = _shared_bperp(bperp, P, I)
    else:
        bp = np.ascontiguousarray(bperp, dtype=np.float64)
        if bp.shape != (P, I):
            raise BackendError("ps_topofit batch expects cpxphase and bperp with matching 2-D shapes")
    K, Cc, coh = np.zeros(P), np.zeros(P), np.zeros(P)
    res = np.zeros((P, I), np.complex64)
    if P:
        _check(_native.lib().ps2_hip_topofit_rows(
            cpx.ctypes.data_as(_f64p), bp.ctypes.data_as(_f64p), int(row_invariant), P, I,
            float(n_trial_wraps), int(device), K.ctypes.data_as(_f64p), Cc.ctypes.data_as(_f64p),
            coh.ctypes.data_as(_f64p), res.ctypes.data_as(_f32p)))
    return K, Cc, coh, res


def stage2_topofit(cpxphase, bperp, n_trial_wraps, threads: int = 0):
    """(K, C, coh, phase_residual) per row; bperp [n_row][n_ifg]."""
    return _topofit(cpxphase, bperp, n_trial_wraps, row_invariant=False)


def stage2_topofit_row_invariant(cpxphase, bperp, n_trial_wraps, threads: int = 0):
    """As stage2_topofit with one bperp vector for every row (1-D, or a 2-D
    matrix whose rows are equal)."""
    return _topofit(cpxphase, bperp, n_trial_wraps, row_invariant=True)


def stage2_topofit_coh_row_invariant(cpxphase, bperp, n_trial_wraps, threads: int = 0, device: int = 0):
    """Coherence per row against one bperp vector, by the rule of
    _ps_topofit_batch_row_invariant_coh (pystamps/pipeline/ported.py:2555-2616):
    rows with one near-maximum peak on the trial grid of the whole vector's
    range are refined from it; the others take stage2_topofit_row_invariant's
    coherence."""
    cpx = np.ascontiguousarray(cpxphase, dtype=np.complex128)
    if cpx.ndim != 2:
        raise BackendError("stage-2 topofit expects a 2-D cpxphase")
    P, I = cpx.shape
    bp = _shared_bperp(bperp, P, I)
    coh = np.zeros(P)
    if P:
        _check(_native.lib().ps2_hip_topofit_coh_row_invariant(
            cpx.ctypes.data_as(_f64p), bp.ctypes.data_as(_f64p), P, I, float(n_trial_wraps),
            int(device), coh.ctypes.data_as(_f64p)))
    return coh


def stage2_grid_accumulate(ph_weight, grid_lin, n_i, n_j, threads: int = 0):
    """Per-cell sums of the rows' weighted phases: complex64 [n_i][n_j][n_ifg]."""
    ph = np.ascontiguousarray(ph_weight, dtype=np.complex64)
    gl = np.ascontiguousarray(np.asarray(grid_lin, dtype=np.int64).reshape(-1))
    if ph.ndim != 2 or ph.shape[0] != gl.size:
        raise BackendError("stage2_grid_accumulate expects ph_weight [n_row][n_ifg] and n_row indices")
    P, I = ph.shape
    out = np.zeros((int(n_i) * int(n_j), I), np.complex64)
    _check(_native.lib().ps2_hip_grid_accumulate(
        ph.ctypes.data_as(_f32p), gl.ctypes.data_as(_i64p), P, I, int(n_i), int(n_j), 0,
        out.ctypes.data_as(_f32p)))
    return out.reshape(int(n_i), int(n_j), I)


def stage2_histogram(values, centers):
    """Counts of the finite values per center (float64)."""
    v = np.ascontiguousarray(np.asarray(values, dtype=np.float64).reshape(-1))
    c = np.ascontiguousarray(np.asarray(centers, dtype=np.float64).reshape(-1))
    out = np.zeros(c.size)
    _check(_native.lib().ps2_hip_histogram_with_centers(
        v.ctypes.data_as(_f64p), v.size, c.ctypes.data_as(_f64p), c.size, out.ctypes.data_as(_f64p)))
    return out


IMPLEMENTATIONS = {
    "stage2_grid_accumulate": stage2_grid_accumulate,
    "stage2_topofit": stage2_topofit,
    "stage2_topofit_row_invariant": stage2_topofit_row_invariant,
    "stage2_topofit_coh_row_invariant": stage2_topofit_coh_row_invariant,
    "stage2_histogram": stage2_histogram,
}


def register(registry, name: str = "hip"):
    """Adds the provider and its five kernels to a registry that has the
    reference's ``register_provider(name, description=..., aliases=...,
    availability_probe=..., unavailable_reason=...)`` and ``register(kernel,
    **{provider: fn})`` methods."""
    registry.register_provider(name, description="HIP backend (MI355X)", aliases=(),
                               availability_probe=available,
                               unavailable_reason="HIP backend requested but no GPU or "
                                                  "libpystamps_hip.so is available")
    for kernel, fn in IMPLEMENTATIONS.items():
        registry.register(kernel, **{name: fn})
    return registry
