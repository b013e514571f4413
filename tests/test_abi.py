"""The C-ABI library loads, exports every symbol include/pystamps_hip.h
declares, and rejects bad input on the host before touching a device.  CPU."""
import ctypes as C
import subprocess

import numpy as np
import pytest

from pystamps_amd import _native


def test_library_exports_every_declared_function():
    L = _native.lib()
    names = _native.declared_functions()
    assert len(names) == 27
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", str(_native.LIB_PATH)],
                         capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    for n in names:
        assert n in exported
    # nothing else with the ABI prefix is exported undeclared
    extra = sorted(e for e in exported if e.startswith("ps2_hip_") and e not in names)
    assert not extra, extra


def test_abi_version_and_error_string():
    L = _native.lib()
    assert L.ps2_hip_abi_version() == _native.ABI_VERSION
    assert isinstance(_native.last_error(), str)


def test_device_count_is_nonnegative():
    assert _native.device_count() >= 0


def test_invalid_patch_rejected_on_host():
    L = _native.lib()
    inp = _native.PatchIn()
    inp.n_ps = 0
    out = _native.PatchOut()
    opts = _native.RunOpts()
    rc = L.ps2_hip_run(C.byref(inp), C.byref(out), C.byref(opts))
    assert rc == -1
    assert "n_ps" in _native.last_error()


def test_null_arguments_rejected():
    L = _native.lib()
    assert L.ps2_hip_run(None, None, None) == -1
    assert L.ps2_hip_session_run(None, None, None) == -1
    nr = np.zeros(100)
    mx = C.c_double()
    assert L.ps2_hip_random_coherence_histogram(None, 0, None, 0, None, 0, 0, 0,
                                                nr.ctypes.data_as(C.POINTER(C.c_double)),
                                                C.byref(mx)) == -1


def test_header_structs_match_ctypes_layout():
    # the ctypes mirror names every header member, in the header's order
    names = [f[0] for f in _native.PatchIn._fields_]
    text = _native.HEADER.read_text()
    start = text.index("typedef struct {", text.index("Prepared per-patch inputs"))
    body = text[start:text.index("} ps2_patch_in;")]
    pos = [body.find(n) for n in names]
    assert all(p >= 0 for p in pos), [n for n, p in zip(names, pos) if p < 0]
    assert pos == sorted(pos)


@pytest.mark.parametrize("mutate, message", [
    (lambda h: setattr(h.s, "n_ifg", 0), "n_ifg must be positive"),
    (lambda h: h.grid_lin.__setitem__(7, h.s.n_i * h.s.n_j), "grid_lin[7] outside"),
    (lambda h: h.grid_lin.__setitem__(0, -1), "grid_lin[0] outside"),
    (lambda h: h.trial_values.__setitem__(1, 0.5), "consecutive integers"),
    (lambda h: setattr(h.s, "n_trial", 4), "2*tn+1"),
    # CLAP windows: 1 <= clap_window <= n_win_ex <= 64 (any clap_win up to 64)
    (lambda h: setattr(h.s, "clap_window", 33), "CLAP window outside"),
    (lambda h: setattr(h.s, "clap_window", 0), "CLAP window outside"),
    (lambda h: setattr(h.s, "n_win_ex", 65), "CLAP window outside"),
    (lambda h: (setattr(h.s, "row_invariant", 0), setattr(h.s, "bperp_mat", None)), "bperp_mat is required"),
    (lambda h: (setattr(h.s, "weighting", _native.WEIGHT_SNR), setattr(h.s, "amp", None)), "amp is required"),
    (lambda h: setattr(h.s, "d_a", None), "required input array is NULL"),
    # size limits are checked before any per-pixel array is read, so these
    # oversized extents never touch the 500-row arrays they point at
    (lambda h: setattr(h.s, "n_ps", (1 << 31) // 6 + 1), "patch too large"),
    # 2^31 pixels in one interferogram, and exactly 2^31 elements: the int32
    # counting sort of the cells would overflow
    (lambda h: (setattr(h.s, "n_ps", 1 << 31), setattr(h.s, "n_ifg", 1)), "patch too large"),
    (lambda h: (setattr(h.s, "n_ps", 1 << 30), setattr(h.s, "n_ifg", 2)), "patch too large"),
    (lambda h: (setattr(h.s, "n_i", 1 << 15), setattr(h.s, "n_j", 1 << 15), setattr(h.s, "n_ifg", 2)),
     "patch too large"),
    (lambda h: (setattr(h.s, "n_i", 1 << 40), setattr(h.s, "n_j", 1 << 40)), "patch too large"),
    (lambda h: (setattr(h.s, "n_i", 1 << 16), setattr(h.s, "n_j", 1 << 14)), "patch too large"),
])
def test_patch_validation_messages_on_host(mutate, message):
    """Inputs the library rejects before any device call, with the message."""
    import dataclasses
    from conftest import small_prep
    from pystamps_amd.stage2 import _InHolder
    prep = small_prep(n_ps=500, n_ifg=6, seed=0x5EED1900)
    # private copies of the arrays a case edits (the prepared patch is cached)
    h = _InHolder(dataclasses.replace(prep, grid_lin=prep.grid_lin.copy(),
                                      trial_values=prep.trial_values.copy()))
    mutate(h)
    out = _native.PatchOut()
    rc = _native.lib().ps2_hip_run(C.byref(h.s), C.byref(out), C.byref(_native.RunOpts()))
    assert rc == -1
    assert message in _native.last_error(), _native.last_error()
