"""The allocation-free topofit entry point (ps2_hip_topofit_rows_noalloc)
against ps2_hip_topofit_rows, through the C ABI.  GPU.

The device buffers are allocated here, by the caller, with the HIP runtime the
library itself links; the entry point brings no allocation of its own.  Calls
follow one another with different sizes, trial counts and layouts (the second
smaller than the first), so the shared scratch array is reused, and a trial
count past the scratch's limit is rejected on the host.
"""
import ctypes as C

import numpy as np
import pytest

from pystamps_amd import _native

pytestmark = pytest.mark.gpu

_f64p = C.POINTER(C.c_double)
_f32p = C.POINTER(C.c_float)


def _runtime():
    """The HIP runtime libpystamps_hip.so is linked against (same handle)."""
    _native.lib()
    with open("/proc/self/maps") as fh:
        path = next(line.split()[-1] for line in fh if "libamdhip64" in line)
    rt = C.CDLL(path)
    rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    rt.hipFree.argtypes = [C.c_void_p]
    rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return rt


class _Dev:
    """Device copy of a host array (caller-owned memory)."""

    def __init__(self, rt, host=None, nbytes=None):
        self.rt = rt
        self.nbytes = host.nbytes if host is not None else nbytes
        self.p = C.c_void_p()
        assert rt.hipMalloc(C.byref(self.p), max(self.nbytes, 8)) == 0
        if host is not None:
            assert rt.hipMemcpy(self.p, np.ascontiguousarray(host).ctypes.data, self.nbytes, 1) == 0

    def get(self, shape, dtype):
        out = np.empty(shape, dtype)
        assert self.rt.hipMemcpy(out.ctypes.data, self.p, out.nbytes, 2) == 0
        return out

    def free(self):
        self.rt.hipFree(self.p)


def _rows(n_rows, n_ifg, row_invariant, seed):
    rng = np.random.default_rng(seed)
    k = rng.uniform(-0.01, 0.01, n_rows)
    bp = rng.uniform(-300.0, 300.0, n_ifg if row_invariant else (n_rows, n_ifg))
    bpm = np.broadcast_to(bp, (n_rows, n_ifg))
    cpx = np.exp(1j * (k[:, None] * bpm + rng.normal(0.0, 0.6, (n_rows, n_ifg))))
    cpx *= rng.uniform(0.5, 2.0, (n_rows, n_ifg))
    cpx[rng.random((n_rows, n_ifg)) < 0.03] = 0.0     # elements left out of the fit
    cpx[3] = 0.0                                       # a row with nothing to fit
    return np.ascontiguousarray(cpx, np.complex128), np.ascontiguousarray(bp, np.float64)


def _host_call(cpx, bp, row_invariant, wraps, with_res):
    n_rows, n_ifg = cpx.shape
    k, c, coh = np.zeros(n_rows), np.zeros(n_rows), np.zeros(n_rows)
    res = np.zeros((n_rows, n_ifg), np.complex64) if with_res else None
    rc = _native.lib().ps2_hip_topofit_rows(
        cpx.ctypes.data_as(_f64p), bp.ctypes.data_as(_f64p), C.c_int32(int(row_invariant)),
        C.c_int64(n_rows), C.c_int64(n_ifg), C.c_double(wraps), C.c_int32(0),
        k.ctypes.data_as(_f64p), c.ctypes.data_as(_f64p), coh.ctypes.data_as(_f64p),
        res.ctypes.data_as(_f32p) if with_res else None)
    assert rc == 0, _native.last_error()
    return k, c, coh, res


def _noalloc_call(rt, cpx, bp, row_invariant, wraps, with_res):
    n_rows, n_ifg = cpx.shape
    d_cpx, d_bp = _Dev(rt, cpx), _Dev(rt, bp)
    d_k, d_c, d_coh = (_Dev(rt, nbytes=8 * n_rows) for _ in range(3))
    d_res = _Dev(rt, nbytes=8 * n_rows * n_ifg) if with_res else None
    try:
        rc = _native.lib().ps2_hip_topofit_rows_noalloc(
            d_cpx.p, d_bp.p, C.c_int32(int(row_invariant)), C.c_int64(n_rows), C.c_int64(n_ifg),
            C.c_double(wraps), C.c_int32(0), None, d_k.p, d_c.p, d_coh.p,
            d_res.p if with_res else None)
        assert rc == 0, _native.last_error()
        out = (d_k.get(n_rows, np.float64), d_c.get(n_rows, np.float64),
               d_coh.get(n_rows, np.float64),
               d_res.get((n_rows, n_ifg), np.complex64) if with_res else None)
    finally:
        for d in (d_cpx, d_bp, d_k, d_c, d_coh, d_res):
            if d is not None:
                d.free()
    return out


def test_topofit_rows_noalloc_equals_topofit_rows():
    rt = _runtime()
    L = _native.lib()
    L.ps2_hip_topofit_rows_noalloc.argtypes = [
        C.c_void_p, C.c_void_p, C.c_int32, C.c_int64, C.c_int64, C.c_double, C.c_int32, C.c_void_p,
        C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    # (rows, ifgs, row-invariant bperp, n_trial_wraps, residual output): the
    # second call is smaller than the first; the last takes 16,381 trials, so
    # the scratch holds 256 rows and the call runs three launches
    cases = [(5000, 24, False, 1.1, True), (1200, 24, True, 0.7, False),
             (700, 12, False, 1023.75, True)]
    for n_rows, n_ifg, rowinv, wraps, with_res in cases:
        cpx, bp = _rows(n_rows, n_ifg, rowinv, seed=n_rows)
        ref = _host_call(cpx, bp, rowinv, wraps, with_res)
        got = _noalloc_call(rt, cpx, bp, rowinv, wraps, with_res)
        for name, a, b in zip(("K", "C", "coh", "res"), got, ref):
            if b is None:
                continue
            assert np.array_equal(a, b, equal_nan=True), (n_rows, n_ifg, wraps, name)
        assert np.all(np.isnan(got[2][3])) and np.isfinite(got[2][4])
    # more trials than the scratch holds (16,385): rejected on the host
    rc = L.ps2_hip_topofit_rows_noalloc(None, None, 0, 0, 8, 1024.0, 0, None, None, None, None, None)
    assert rc == -1 and "16384 trials" in _native.last_error(), _native.last_error()
