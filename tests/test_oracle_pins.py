"""Independent pins of the oracle's CLAP gain path, low-pass multiply, fill,
extract, gamma-change RMS and SNR weighting.  CPU only.

Each oracle function is compared with a composition of library routines
(numpy.fft, scipy.signal.convolve2d, scipy.signal.windows.gaussian, np.median)
that follows the reference's definition step by step, or with a closed form.
The last test compiles the oracle with planted mistakes (-DORA_MUTANT=n) and
checks that the pin meant for each mistake fails on it.

Reference: clap_filter_grid_stack / clap_filter_one_ifg and helpers
(crates/pystamps-core/src/native_stage2.rs:742-1184), fill_phase_weight
(:629-657), extract_patch_phase (:1186-1205), snr_weighting (:1875-1897),
rms_difference (:2045-2057).
"""
import dataclasses
import math

import numpy as np
import pytest
import scipy.signal

from conftest import small_prep

pytestmark = pytest.mark.filterwarnings("ignore::RuntimeWarning")

N_WIN, N_EX, N_INC = 24, 32, 6


# --- library restatement of one interferogram's CLAP -------------------------

def _window_weight(n_win=N_WIN):
    # clap_window_weight (native_stage2.rs:1014-1037): quadrant value
    # row + col + 1e-6 on a (n_win // 2)-square quadrant, mirrored into the
    # other three; an odd window's last row and column stay 0
    h2 = 2 * (n_win // 2)
    r = np.arange(h2)
    m = np.minimum(r, h2 - 1 - r).astype(np.float64)
    w = np.zeros((n_win, n_win))
    w[:h2, :h2] = m[:, None] + m[None, :] + 1e-6
    return w


def _windows(n_i, n_j, n_win=N_WIN):
    # clap_windows (native_stage2.rs:957-995): stride max(n_win // 4, 1), a
    # window past the grid edge is moved back inside and remembers the shift
    n_inc = max(n_win // 4, 1)
    nwi = math.ceil(n_i / n_inc) - 3
    nwj = math.ceil(n_j / n_inc) - 3
    out = []
    for a in range(max(nwi, 0)):
        i1, rs = a * n_inc, 0
        if i1 + n_win > n_i:
            rs, i1 = i1 + n_win - n_i, n_i - n_win
        for b in range(max(nwj, 0)):
            j1, cs = b * n_inc, 0
            if j1 + n_win > n_j:
                cs, j1 = j1 + n_win - n_j, n_j - n_win
            out.append((i1, j1, rs, cs))
    return out if nwi > 0 and nwj > 0 else []


def clap_plane_by_library(plane, low_pass, alpha, beta, n_win=N_WIN):
    """One interferogram plane [n_i, n_j] complex128 through CLAP, built from
    numpy / scipy routines; window n_win, transform size = low_pass's."""
    n_i, n_j = plane.shape
    n_ex = low_pass.shape[0]
    g7 = scipy.signal.windows.gaussian(7, std=(7 - 1) / (2 * 2.5))   # gausswin(7, 2.5)
    k2 = np.outer(g7, g7)
    base_w = _window_weight(n_win)
    acc = np.zeros((n_i, n_j), np.complex128)
    for i1, j1, rs, cs in _windows(n_i, n_j, n_win):
        blk = plane[i1:i1 + n_win, j1:j1 + n_win]
        blk = np.where(np.isnan(blk.real) | np.isnan(blk.imag), 0.0, blk)
        x = np.zeros((n_ex, n_ex), np.complex128)
        x[:n_win, :n_win] = blk
        F = np.fft.fft2(x)
        # the reference's shifts (shift_real_into, native_stage2.rs:1097-1120)
        # read at offset floor(n / 2) (fftshift) and ceil(n / 2) (ifftshift):
        # numpy's fftshift / ifftshift for an even n, swapped for an odd one
        sm = scipy.signal.convolve2d(np.roll(np.abs(F), -(n_ex // 2), axis=(0, 1)), k2, mode="same",
                                     boundary="fill", fillvalue=0.0)
        H = np.roll(sm, -((n_ex + 1) // 2), axis=(0, 1))
        med = np.median(H)
        Hn = H / med if med != 0.0 else H
        gain = np.maximum(Hn ** alpha - 1.0, 0.0) * beta + low_pass
        y = np.fft.ifft2(F * gain)[:n_win, :n_win]      # ifft2 carries the 1/n_ex^2
        w = np.zeros((n_win, n_win))
        w[rs:, cs:] = base_w[:n_win - rs, :n_win - cs]
        acc[i1:i1 + n_win, j1:j1 + n_win] += y * w
    return acc


def _prep(n_i, n_j, n_ifg=2, low_pass=None, alpha=1.0, beta=0.3):
    base = small_prep(n_ps=400, n_ifg=n_ifg)
    kw = dict(n_i=n_i, n_j=n_j, clap_alpha=alpha, clap_beta=beta,
              grid_lin=np.zeros(base.n_ps, np.int64))
    if low_pass is not None:
        kw["low_pass"] = np.ascontiguousarray(low_pass, np.float64)
    return dataclasses.replace(base, **kw)


def _random_grid(n_i, n_j, n_ifg, seed, nan_cells=0):
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((n_i * n_j, n_ifg)) + 1j * rng.standard_normal((n_i * n_j, n_ifg))
    if nan_cells:
        idx = rng.choice(n_i * n_j, nan_cells, replace=False)
        g[idx[: nan_cells // 2], 0] = np.nan + 0j
        g[idx[nan_cells // 2:], 1 % n_ifg] = complex(0.0, np.nan)
    return g


def _asym_low_pass(seed=11):
    lp = np.random.default_rng(seed).uniform(0.0, 1.0, (N_EX, N_EX))
    assert np.abs(lp - lp.T).max() > 0.1
    return lp


def _compare_clap(O, n_i, n_j, alpha, beta, low_pass, seed, nan_cells=0, n_ifg=2):
    prep = _prep(n_i, n_j, n_ifg=n_ifg, low_pass=low_pass, alpha=alpha, beta=beta)
    g = _random_grid(n_i, n_j, n_ifg, seed, nan_cells)
    ref = np.stack([clap_plane_by_library(g[:, f].reshape(n_i, n_j), prep.low_pass, alpha, beta)
                    .reshape(-1) for f in range(n_ifg)], axis=1)
    scale = np.abs(ref).max()
    assert scale > 0
    out64 = O.clap_filter_grid_stack(prep, g, mode=O.REF64)
    err64 = np.abs(out64 - ref).max() / scale
    assert err64 < 1e-12, (n_i, n_j, alpha, beta, err64)
    # reference precision map: complex64 in and out, f64 inside
    g32 = g.astype(np.complex64)
    ref32 = np.stack([clap_plane_by_library(g32[:, f].astype(np.complex128).reshape(n_i, n_j),
                                            prep.low_pass, alpha, beta).reshape(-1)
                      for f in range(n_ifg)], axis=1)
    out32 = O.clap_filter_grid_stack(prep, g32, mode=O.REF32)
    err32 = np.abs(out32 - ref32).max() / scale
    assert err32 < 2e-7, (n_i, n_j, alpha, beta, err32)


# --- the pins (each raises AssertionError on a mismatch) ---------------------

def pin_clap_gain_one_window(O):
    # 24 x 24 cells: exactly one window, no clamping; both gain forms
    for alpha in (1.0, 0.7):
        _compare_clap(O, 24, 24, alpha, 0.3, _asym_low_pass(), seed=21)


def pin_clap_gain_edges(O):
    # clamped last window row (27 x 30), clamped last window column (30 x 27),
    # several windows each way, NaN cells (read as 0), the StaMPS low-pass
    from pystamps_amd.prepare import Stage2Options, build_low_pass
    _compare_clap(O, 27, 30, 0.7, 0.3, _asym_low_pass(12), seed=22)
    _compare_clap(O, 30, 27, 1.0, 0.3, _asym_low_pass(13), seed=23)
    _compare_clap(O, 45, 52, 0.7, 0.45, _asym_low_pass(14), seed=24, nan_cells=6)
    _compare_clap(O, 41, 38, 1.0, 0.3, build_low_pass(Stage2Options()), seed=25, n_ifg=3)


def pin_clap_low_pass_convolution(O):
    """beta = 0 leaves gain = low_pass, so each window is the circular
    convolution of its zero-padded block with h = ifft2(low_pass), written as
    a direct sum over the 24 x 24 input cells (no FFT); an asymmetric low_pass
    catches a transposed or mirrored multiply."""
    lp = _asym_low_pass(15)
    prep = _prep(N_WIN, N_WIN, n_ifg=2, low_pass=lp, beta=0.0)
    g = _random_grid(N_WIN, N_WIN, 2, seed=26)
    h = np.fft.ifft2(lp)
    n = np.arange(N_WIN)
    d = (n[:, None] - n[None, :]) % N_EX          # (out index, in index) -> h index
    H4 = h[d[:, None, :, None], d[None, :, None, :]]   # [o1, o2, m1, m2]
    w = _window_weight()
    out = O.clap_filter_grid_stack(prep, g, mode=O.REF64)
    for f in range(2):
        x = g[:, f].reshape(N_WIN, N_WIN)
        y = np.einsum("abcd,cd->ab", H4, x) * w
        err = np.abs(out[:, f].reshape(N_WIN, N_WIN) - y).max() / np.abs(y).max()
        assert err < 1e-12, err


def pin_fill(O):
    """fill_phase_weight with ph_nm = e^{+i bp K} (complex64): the fill's ramp
    e^{-i bp K} cancels it, so every element is the real weight w (to the
    complex64 quantisation of ph_nm); a NaN K gives NaN."""
    base = small_prep(n_ps=3000, n_ifg=20)
    rng = np.random.default_rng(27)
    P, I = base.n_ps, base.n_ifg
    K = rng.uniform(-0.02, 0.02, P)
    W = rng.uniform(0.1, 2.0, P)
    bp32 = base.bperp_mat.astype(np.float32)
    ph32 = np.exp(1j * (bp32 * K.astype(np.float32)[:, None]).astype(np.float64))
    prep = dataclasses.replace(base, ph_nm=ph32.astype(np.complex64))
    for mode in (32, 64):
        out = O.fill_phase_weight(prep, K, W, mode=mode)
        np.testing.assert_allclose(out.real, np.broadcast_to(W[:, None], (P, I)), rtol=1e-6)
        assert np.abs(out.imag).max() < 1e-6 * W.max()
    Kn = K.copy()
    Kn[5] = np.nan
    out = O.fill_phase_weight(prep, Kn, W, mode=64)
    assert np.all(np.isnan(out[5])) and np.all(np.isfinite(out[4]))


def pin_extract(O):
    """extract_patch_phase: each row takes its own cell's filtered value,
    direction kept, magnitude 1; zero cells stay 0."""
    base = small_prep(n_ps=3000, n_ifg=20)
    C, I = base.n_i * base.n_j, base.n_ifg
    rng = np.random.default_rng(28)
    mag = rng.uniform(0.01, 50.0, (C, I))
    ang = rng.uniform(-np.pi, np.pi, (C, I))
    filt = mag * np.exp(1j * ang)
    zero = rng.random((C, I)) < 0.05
    filt[zero] = 0.0
    for mode, tol in ((64, 1e-14), (32, 3e-7)):
        f = filt if mode == 64 else filt.astype(np.complex64)
        out = O.extract_patch_phase(base, f, mode=mode)
        cell = base.grid_lin
        z = zero[cell]
        assert not np.any(out[z])
        assert np.abs(np.abs(out[~z]) - 1.0).max() < tol
        dang = np.angle(out[~z] * np.conj(filt[cell][~z]))
        assert np.abs(dang).max() < 2 * tol


def pin_rms(O):
    rng = np.random.default_rng(29)
    a, b = rng.random(10_001), rng.random(10_001)
    got = O.rms_difference(a, b)
    assert abs(got - np.sqrt(np.mean((a - b) ** 2))) <= 1e-15
    assert O.rms_difference(np.zeros(0), np.zeros(0)) == 0.0


def pin_snr(O):
    """snr_weighting closed forms: g = mean(amp cos ph_res),
    sigma = sqrt((mean amp^2 - g^2) / 2), w = g / sigma (0 when sigma = 0)."""
    for mode in (64, 32):
        rt = np.float64 if mode == 64 else np.float32
        w = O.snr_weighting(np.array([[1.0, 3.0]]), np.zeros((1, 2), rt), mode=mode)
        assert abs(w[0] - 2.0 * math.sqrt(2.0)) < 1e-14          # g 2, sigma 1/sqrt 2
        w = O.snr_weighting(np.ones((1, 2)), np.array([[math.pi / 3, -math.pi / 3]], rt),
                            mode=mode)
        tol = 1e-14 if mode == 64 else 1e-6
        assert abs(w[0] - math.sqrt(2.0 / 3.0)) < tol            # g 1/2, sigma sqrt(3/8)
        w = O.snr_weighting(np.full((1, 4), 2.5), np.zeros((1, 4), rt), mode=mode)
        assert w[0] == 0.0                                        # sigma 0
        amp = np.random.default_rng(30).uniform(0.2, 3.0, (50, 12))
        w = O.snr_weighting(amp, np.zeros((50, 12), rt), mode=mode)
        a32 = amp.astype(np.float32).astype(np.float64)
        g = a32.mean(1)
        sig = np.sqrt(0.5 * ((a32 ** 2).mean(1) - g ** 2))
        np.testing.assert_allclose(w, g / sig, rtol=1e-12)


def pin_loop_wiring(O):
    """The loop driver (native_stage2.rs:242-406) composed in numpy from the
    pinned step functions: initial K = 0 and w = 1/D_A, fill -> grid -> CLAP ->
    extract -> fit, gamma-change RMS, Nr scaled by the low-coherence bin sums
    of hist(coh), P-square weights, second iteration.  The oracle's own loop
    must give the same state after two iterations."""
    prep = dataclasses.replace(small_prep(n_ps=1500, n_ifg=12), gamma_max_iterations=2,
                               gamma_change_convergence=1e-300)
    for mode in (64, 32):
        K = np.zeros(prep.n_ps)
        w = np.where(prep.d_a != 0, 1.0 / np.where(prep.d_a != 0, prep.d_a, 1.0), 0.0)
        nr = prep.nr_base.astype(np.float64).copy()
        coh_prev = np.zeros(prep.n_ps)
        rms_prev = 0.0
        for it in (1, 2):
            pw = O.fill_phase_weight(prep, K, w, mode=mode)
            grid = O.accumulate_grid(prep, pw, mode=mode)
            filt = O.clap_filter_grid_stack(prep, grid, mode=mode)
            patch = O.extract_patch_phase(prep, filt, mode=mode)
            fit = O.fit_rows(prep, patch, mode=mode)
            d = fit["coh_ps"] - coh_prev
            rms = math.sqrt(np.cumsum(d * d)[-1] / prep.n_ps)   # sequential sum, as :2045-2057
            rms_prev, coh_prev = rms, fit["coh_ps"].copy()
            K = fit["K_ps"]
            if it == 1:
                na = O.hist_with_centers(fit["coh_ps"], prep.coh_bins)
                lct = prep.low_coh_thresh
                nr = nr * (na[:lct].sum() / nr[:lct].sum())
                w, _ = O.psquare_weighting(nr, na, lct, prep.nr_max_nz_ix, fit["coh_ps"])
        o = O.run(prep, mode=mode)
        assert o["i_loop"] == 2
        assert np.array_equal(o["K_ps"], K, equal_nan=True)
        assert np.array_equal(o["coh_ps"], fit["coh_ps"])
        assert np.array_equal(o["coh_ps_save"], coh_prev)
        assert o["gamma_change_save"] == rms_prev
        np.testing.assert_array_equal(o["Nr"], nr)
        np.testing.assert_array_equal(o["ph_grid"], grid)
        np.testing.assert_array_equal(o["ph_patch"], patch)
        np.testing.assert_array_equal(o["ph_res"], fit["ph_res"])


PINS = {
    "clap_gain_one_window": pin_clap_gain_one_window,
    "clap_gain_edges": pin_clap_gain_edges,
    "clap_low_pass_convolution": pin_clap_low_pass_convolution,
    "fill": pin_fill,
    "extract": pin_extract,
    "rms": pin_rms,
    "snr": pin_snr,
    "loop_wiring": pin_loop_wiring,
}


@pytest.mark.parametrize("name", sorted(PINS))
def test_oracle_pin(ora, name):
    PINS[name](ora)


# planted mistake (ORA_MUT_* in oracle/stage2_oracle.c) -> the pin that must fail
MUTANTS = {
    1: "clap_gain_one_window",        # smoothing before the fftshift
    2: "clap_gain_one_window",        # median = upper middle value
    3: "clap_low_pass_convolution",   # low_pass transposed
    4: "clap_gain_one_window",        # ifftshift left out
    5: "clap_gain_one_window",        # fftshift by n/2 - 1
    6: "clap_gain_one_window",        # gain not clamped at 0
    7: "clap_gain_edges",             # clamped window weighted without its shift
    8: "fill",                        # ramp sign
    9: "extract",                     # no normalisation
    10: "snr",                        # sigma without 1/2
    11: "rms",                        # sum instead of mean
    12: "clap_gain_one_window",       # alpha ignored
    13: "loop_wiring",                # Nr not rescaled
}


def test_planted_mistakes_fail_their_pins(ora, tmp_path):
    """Every planted mistake is caught by the pin named for it, so the pins
    above do not pass a plausibly wrong oracle."""
    escaped = []
    for m, pin in MUTANTS.items():
        path = ora.build_variant(tmp_path / f"liboracle_mut{m}.so", (f"ORA_MUTANT={m}", ))
        with ora.use_library(path):
            try:
                PINS[pin](ora)
            except AssertionError:
                continue
            escaped.append((m, pin))
    assert not escaped, f"planted mistakes not caught: {escaped}"
    PINS["clap_gain_one_window"](ora)   # the default library again


@pytest.mark.parametrize("clap_win", [16.0, 22.0, 27.0, 64.0])
def test_oracle_clap_other_windows(ora, clap_win):
    """The oracle's CLAP for windows other than 24 of 32 (clap_win 16: 12 of 16;
    22: an odd 17-cell window in a 22-point transform; 27: 20 of 27, an odd
    transform, where the reference's shifts are numpy's swapped; 64: 48 of
    64) against the library composition, with the StaMPS low-pass of that
    size, clamped edge windows and NaN cells."""
    from pystamps_amd.prepare import Stage2Options, build_low_pass
    o = Stage2Options()
    o.clap_win = clap_win
    lp = build_low_pass(o)
    n_win = int(math.floor(clap_win * 0.75 + 0.5))
    n_i, n_j = n_win + 11, n_win + 7
    prep = dataclasses.replace(_prep(n_i, n_j, low_pass=lp, alpha=0.8), clap_window=n_win)
    g = _random_grid(n_i, n_j, 2, seed=31, nan_cells=6)
    ref = np.stack([clap_plane_by_library(g[:, f].reshape(n_i, n_j), lp, 0.8, 0.3, n_win).reshape(-1)
                    for f in range(2)], axis=1)
    out = ora.clap_filter_grid_stack(prep, g, mode=ora.REF64)
    scale = np.abs(ref).max()
    assert scale > 0 and len(_windows(n_i, n_j, n_win)) >= 4
    assert np.abs(out - ref).max() / scale < 1e-12
