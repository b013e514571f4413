"""Pins the oracle (oracle/stage2_oracle.c) against the reference's own golden
vectors and against what the mathematics fixes.  CPU only."""
import dataclasses
import math

import numpy as np
import pytest

from conftest import small_prep

pytestmark = pytest.mark.filterwarnings("ignore::RuntimeWarning")


# --- golden vectors held by the reference ------------------------------------

def test_rng_golden(ora):
    # native_stage2.rs:2463-2484
    v = ora.rng_uniform(1_000_001)
    expected = [(0, 0.092958990583191195), (1, 0.37373775840752277), (2, 0.44877057937117765),
                (9, 0.574897127600281), (10, 0.067322284274237226), (31, 0.43152622807181651),
                (32, 0.88306283014597708), (33, 0.074287163034898504),
                (1000, 0.35920165460545467), (99_999, 0.020806164093088213),
                (1_000_000, 0.4917673475125352)]
    for ix, e in expected:
        assert abs(v[ix] - e) <= 1e-15


def test_psquare_golden(ora):
    # native_stage2.rs:2486-2509
    nr = np.arange(100.0)
    na = np.arange(100.0, 200.0)
    coh = [0.0, 0.0049, 0.005, 0.3144, 0.3145, 0.9999, 1.1]
    w, _ = ora.psquare_weighting(nr, na, 31, 45.0, coh)
    expected = [1.8308507589276466e-06, 0.011860147976383777, 0.011860147976383777,
                4.288734335628797e-06, 8.21030522409814e-06, 1.0, 1.0]
    np.testing.assert_allclose(w, expected, rtol=0, atol=1e-12)


def test_psquare_round_half_away_lookup_and_tail(ora):
    # tests/test_stage2_trial_wraps.py:76-108
    w, hi = ora.psquare_weighting(np.ones(5), np.ones(5), 0, 5.0, [0.0015])
    assert w[0] == (1.0 - hi[2]) ** 2          # round(1.5) = 2, half away from zero
    w, hi = ora.psquare_weighting(np.ones(100), np.ones(100), 0, 100.0, [0.995])
    assert hi.shape == (1001,)
    assert w[0] == (1.0 - hi[995]) ** 2


def test_hist_with_centers_golden(ora):
    # native_stage2.rs:2511-2519
    h = ora.hist_with_centers([0.01, 0.02, 0.03, np.nan], [0.005, 0.015, 0.025, 0.035])
    assert list(h) == [1.0, 1.0, 1.0, 0.0]


NR_GOLDEN = [
    0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 1.0, 1.0, 4.0, 5.0,
    0.0, 2.0, 7.0, 5.0, 14.0, 14.0, 15.0, 10.0, 29.0, 21.0, 26.0, 29.0, 44.0, 42.0, 57.0,
    54.0, 63.0, 71.0, 68.0, 71.0, 87.0, 87.0, 76.0, 98.0, 111.0, 100.0, 137.0, 123.0,
    139.0, 154.0, 148.0, 149.0, 162.0, 195.0, 212.0, 202.0, 200.0, 203.0, 228.0, 209.0,
    222.0, 227.0, 223.0, 239.0, 201.0, 199.0, 231.0, 203.0, 192.0, 184.0, 187.0, 188.0,
    185.0, 180.0, 183.0, 186.0, 163.0, 152.0, 147.0, 155.0, 158.0, 171.0, 169.0, 155.0,
    139.0, 140.0, 144.0, 126.0, 125.0, 125.0, 115.0, 110.0, 116.0, 104.0, 95.0, 75.0, 74.0,
    85.0, 73.0, 56.0, 61.0, 37.0, 26.0,
]


def test_random_coherence_histogram_golden(ora):
    # native_stage2.rs:2521-2559: bperp [-20,-5,0,13,27], wraps 0.5, 10,000 rows
    from pystamps_amd.prepare import coh_bins
    tv = ora.trial_values(0.5)
    nr, mx = ora.random_coherence_histogram([-20.0, -5.0, 0.0, 13.0, 27.0], tv, coh_bins(),
                                            n_rows=10_000)
    assert list(nr) == NR_GOLDEN
    assert mx == 100.0


def test_trial_values_golden(ora):
    # tests/test_stage2_ported.py:703-707
    from pystamps_amd.prepare import trial_values
    for fn in (ora.trial_values, trial_values):
        v = fn(0.725669801235199)
        assert v.shape == (13,)
        assert np.array_equal(v, np.arange(-6.0, 7.0))


def test_grid_index_single_precision_boundary(ora):
    # tests/test_stage2_trial_wraps.py:111-127: f32 arithmetic puts row 2 in
    # column 40 where f64 arithmetic would give 41
    from pystamps_amd.prepare import stage2_grid_indices
    xy = np.array([[1.0, -23497.43359375, 0.0], [2.0, -21497.43359375, 0.0],
                   [3.0, -20497.43359375, 0.0]], dtype=np.float32)
    ij, ni, nj = ora.grid_indices(xy, 50.0)
    assert int(ij[1, 1]) == 40
    assert np.array_equal(stage2_grid_indices(xy, 50.0), ij)


# --- closed forms and invariants ----------------------------------------------

def test_low_pass_closed_form(ora):
    lp = ora.low_pass(32.0, 800.0, 50.0)
    assert lp.shape == (32, 32)
    assert lp[0, 0] == 1.0                        # DC moved to [0, 0]
    np.testing.assert_allclose(lp[2, 0], 0.5, rtol=1e-15)   # half power at 1/800 m^-1
    np.testing.assert_array_equal(lp, lp.T)
    from pystamps_amd.prepare import Stage2Options, build_low_pass
    np.testing.assert_allclose(build_low_pass(Stage2Options()), lp, rtol=0, atol=0)


def test_window_weight_closed_form(ora):
    w = ora.clap_window_weight(24)
    assert w[0, 0] == 1e-6 and w[23, 23] == 1e-6 and w[0, 23] == 1e-6
    assert w[11, 11] == 22.0 + 1e-6 and w[12, 12] == 22.0 + 1e-6
    np.testing.assert_array_equal(w, w[::-1, :])
    np.testing.assert_array_equal(w, w[:, ::-1])
    # every interior cell is covered by 4x4 windows whose weights sum to 176 + 16e-6
    tot = 0.0
    for a in range(4):
        for b in range(4):
            tot += w[3 + 6 * a, 2 + 6 * b]
    assert abs(tot - (176.0 + 16e-6)) < 1e-9


def test_fft_matches_brute_force_dft(ora):
    rng = np.random.default_rng(1)
    x = rng.standard_normal((32, 32)) + 1j * rng.standard_normal((32, 32))
    n = np.arange(32)
    W = np.exp(-2j * np.pi * np.outer(n, n) / 32)
    ref = W @ x @ W.T
    np.testing.assert_allclose(ora.fft32_2d(x), ref, rtol=0, atol=1e-11)
    np.testing.assert_allclose(ora.fft32_2d(ref, inverse=True), 1024 * x, rtol=0, atol=1e-9)


def _clap_prep(ni, nj, n_ifg=2, low_pass=None, beta=0.3):
    p = small_prep(n_ps=400, n_ifg=n_ifg)
    kw = dict(n_i=ni, n_j=nj, clap_beta=beta)
    if low_pass is not None:
        kw["low_pass"] = low_pass
    return dataclasses.replace(p, **kw)


def test_clap_all_pass_is_weight_sum(ora):
    # low_pass == 1 and beta == 0 make each window an identity filter, so the
    # output is the input times the summed window weights (176 + 16e-6 inside)
    prep = _clap_prep(60, 54, low_pass=np.ones((32, 32)), beta=0.0)
    rng = np.random.default_rng(2)
    g = (rng.standard_normal((60 * 54, 2)) + 1j * rng.standard_normal((60 * 54, 2)))
    out = ora.clap_filter_grid_stack(prep, g, mode=ora.REF64).reshape(60, 54, 2)
    gi = g.reshape(60, 54, 2)
    inner = (slice(18, 42), slice(18, 36))
    np.testing.assert_allclose(out[inner], gi[inner] * (176.0 + 16e-6), rtol=1e-12)


def test_clap_shift_equivariance_and_zero(ora):
    prep = _clap_prep(72, 66)
    rng = np.random.default_rng(3)
    g = (rng.standard_normal((72, 66, 2)) + 1j * rng.standard_normal((72, 66, 2)))
    out = ora.clap_filter_grid_stack(prep, g.reshape(-1, 2), mode=ora.REF64).reshape(72, 66, 2)
    g2 = np.roll(g, (6, 12), axis=(0, 1))
    out2 = ora.clap_filter_grid_stack(prep, g2.reshape(-1, 2), mode=ora.REF64).reshape(72, 66, 2)
    # windows tile with stride 6, so interior output shifts with the input
    np.testing.assert_allclose(out2[30:48, 36:48], out[24:42, 24:36], rtol=1e-10, atol=1e-12)
    z = ora.clap_filter_grid_stack(prep, np.zeros((72 * 66, 2), np.complex128), mode=ora.REF64)
    assert not np.any(z)


def test_clap_linear_when_beta_zero(ora):
    prep = _clap_prep(48, 42, beta=0.0)
    rng = np.random.default_rng(4)
    a = rng.standard_normal((48 * 42, 2)) + 1j * rng.standard_normal((48 * 42, 2))
    b = rng.standard_normal((48 * 42, 2)) + 1j * rng.standard_normal((48 * 42, 2))
    fa = ora.clap_filter_grid_stack(prep, a, mode=ora.REF64)
    fb = ora.clap_filter_grid_stack(prep, b, mode=ora.REF64)
    fab = ora.clap_filter_grid_stack(prep, 2.0 * a - 3.0 * b, mode=ora.REF64)
    np.testing.assert_allclose(fab, 2.0 * fa - 3.0 * fb, rtol=1e-10, atol=1e-10)


def test_clap_empty_when_no_windows(ora):
    prep = _clap_prep(18, 40)   # ceil(18/6) - 3 = 0 window rows
    g = np.ones((18 * 40, 2), np.complex64)
    assert not np.any(ora.clap_filter_grid_stack(prep, g))


def test_accumulate_grid_brute_force_and_f32_order(ora):
    prep = small_prep(n_ps=3000, n_ifg=20)
    rng = np.random.default_rng(5)
    pw = (rng.standard_normal((prep.n_ps, prep.n_ifg))
          + 1j * rng.standard_normal((prep.n_ps, prep.n_ifg))).astype(np.complex128)
    g64 = ora.accumulate_grid(prep, pw, mode=ora.REF64)
    ref = np.zeros_like(g64)
    np.add.at(ref, prep.grid_lin, pw)
    np.testing.assert_allclose(g64, ref, rtol=1e-12, atol=1e-12)
    # f32 sequential cancellation (tests/test_stage2_ported.py:864-907 semantics)
    p3 = dataclasses.replace(prep, n_ps=3, n_ifg=1, grid_lin=np.zeros(3, np.int64), n_i=1, n_j=1,
                             ph_nm=np.ones((3, 1), np.complex64), amp=np.ones((3, 1), np.float32),
                             d_a=np.ones(3), bperp_mat=None, row_invariant_bperp=True,
                             row_bperp_nm=np.zeros(1))
    w = np.array([[16777216.0], [1.0], [-16777216.0]])
    assert ora.accumulate_grid(p3, w.astype(np.complex64), mode=ora.REF32)[0, 0] == 0.0
    assert ora.accumulate_grid(p3, w.astype(np.complex128), mode=ora.REF64)[0, 0] == 1.0


def _noise_free_prep(k_true, c_true, tn=6, n_ifg=40, seed=6):
    # baselines come in +/- pairs, so sum e^{i d bp} is real and the refine's
    # offset phase is exactly c: one least-squares step then recovers K
    rng = np.random.default_rng(seed)
    P = len(k_true)
    half = rng.uniform(0, 200, (P, n_ifg // 2))
    half[:, 0] = 200.0                   # range 400 for every row
    bp = np.concatenate([half, -half], axis=1)
    phase = k_true[:, None] * bp + c_true[:, None]
    base = small_prep(n_ps=400, n_ifg=n_ifg)
    prep = dataclasses.replace(
        base, n_ps=P, ph_nm=np.exp(1j * phase).astype(np.complex64),
        amp=np.ones((P, n_ifg), np.float32), bperp_mat=bp, row_invariant_bperp=False,
        row_bperp_nm=bp[0].copy(), grid_lin=np.zeros(P, np.int64), d_a=np.ones(P),
        trial_values=np.arange(-tn, tn + 1, dtype=np.float64))
    return prep


@pytest.mark.parametrize("mode", [32, 64])
def test_topofit_noise_free_on_trial_nodes(ora, mode):
    t = np.array([-6, -3, 0, 2, 5])
    k_true = (math.pi / 4) / 400.0 * t
    c_true = np.array([0.3, -1.0, 2.0, -2.5, 0.0])
    prep = _noise_free_prep(k_true, c_true)
    patch = np.ones((len(t), prep.n_ifg), np.complex64 if mode == 32 else np.complex128)
    o = ora.fit_rows(prep, patch, mode=mode)
    assert np.array_equal(o["trial_ix"], t + 6)
    tol = 2e-5 if mode == 32 else 1e-12
    np.testing.assert_allclose(o["coh_ps"], 1.0, atol=tol)
    np.testing.assert_allclose(o["K_ps"], k_true, atol=tol * 1e-2)
    # the fit removes e^{-iK bp}, leaving the constant phase c; ph_nm is stored
    # as complex64, so even the f64 mode sees ~1e-8 rad input quantisation
    c_tol = 1e-6 if mode == 32 else 1e-7
    np.testing.assert_allclose(np.angle(np.exp(1j * (o["C_ps"] - c_true))), 0.0, atol=c_tol)


def test_topofit_off_node_refines(ora):
    k_true = np.array([0.0013, -0.0071, 0.0042])
    c_true = np.array([0.1, 0.2, -0.3])
    prep = _noise_free_prep(k_true, c_true, seed=7)
    o = ora.fit_rows(prep, np.ones((3, prep.n_ifg), np.complex128), mode=ora.REF64)
    np.testing.assert_allclose(o["K_ps"], k_true, atol=1e-12)
    np.testing.assert_allclose(o["coh_ps"], 1.0, atol=1e-12)


def test_topofit_matches_brute_force_fp64_scan(ora):
    prep = small_prep(n_ps=3000, n_ifg=20)
    rng = np.random.default_rng(8)
    patch = np.exp(1j * rng.uniform(-np.pi, np.pi, (prep.n_ps, prep.n_ifg)))
    o = ora.fit_rows(prep, patch, mode=ora.REF64)
    cpx = np.conj(patch) * prep.ph_nm.astype(np.complex128)
    bp = prep.bperp_mat
    rngb = bp.max(1) - bp.min(1)
    tv = prep.trial_values
    ph = (bp / rngb[:, None])[:, :, None] * (np.pi / 4) * tv[None, None, :]
    coh = np.abs((cpx[:, :, None] * np.exp(-1j * ph)).sum(1)) / np.abs(cpx).sum(1)[:, None]
    srt = np.sort(coh, axis=1)
    clear = (srt[:, -1] - srt[:, -2]) > 1e-9
    assert np.array_equal(o["trial_ix"][clear], np.argmax(coh, axis=1)[clear])


def test_three_ps_fixture_known_outcome(ora):
    # create_stage2_fixture (native_stage2.rs:2698-2769): 3 PS in one 1x1 cell,
    # so CLAP is empty, every row is invalid and the loop stops after one pass
    from pystamps_amd.prepare import RawPatch, prepare_stage2_inputs
    from conftest import oracle_nr
    raw = three_ps_raw()
    prep = prepare_stage2_inputs(raw, nr_fn=oracle_nr, nr_rows=10_000)
    assert (prep.n_i, prep.n_j) == (1, 1)
    for mode in (32, 64):
        o = ora.run(prep, mode=mode)
        assert o["i_loop"] == 1
        assert np.all(np.isnan(o["K_ps"]))
        assert not np.any(o["C_ps"]) and not np.any(o["coh_ps"]) and not np.any(o["N_opt"])
        assert not np.any(o["ph_res"])


def three_ps_raw(bad_bp_shape=False):
    from pystamps_amd.prepare import RawPatch
    ph = np.array([[1.0, 0.8 + 0.2j, 0.6 + 0.4j], [1.0, 0.7 + 0.3j, 0.5 + 0.5j],
                   [1.0, 0.9 + 0.1j, 0.4 + 0.6j]], np.complex64)
    bp = (np.array([[10.0], [20.0], [30.0]]) if bad_bp_shape
          else np.array([[15.0, 30.0]] * 3))
    return RawPatch(ph=ph, bperp=np.array([0.0, 15.0, 30.0]),
                    xy=np.array([[1, 0, 0], [2, 5, 5], [3, 10, 10]], np.float32), master_ix=1,
                    bperp_mat=bp, d_a=np.ones(3), la=np.full(3, math.radians(23.0)),
                    parms={"gamma_max_iterations": 1.0, "filter_grid_size": 50.0})


def test_incompatible_bp1_shape_is_stage2_error():
    # native_stage2.rs:2616-2626
    from pystamps_amd.prepare import Stage2Error, prepare_stage2_inputs
    from conftest import oracle_nr
    with pytest.raises(Stage2Error) as e:
        prepare_stage2_inputs(three_ps_raw(bad_bp_shape=True), nr_fn=oracle_nr, nr_rows=100)
    assert "stage 2 native implementation error" in str(e.value)
    assert "bp1.bperp_mat has incompatible shape" in str(e.value)


def test_ref32_and_ref64_loops_agree(ora):
    prep = small_prep(n_ps=3000, n_ifg=20)
    a = ora.run(prep, mode=ora.REF32)
    b = ora.run(prep, mode=ora.REF64)
    assert a["i_loop"] == b["i_loop"]
    assert np.mean(a["trial_ix"] == b["trial_ix"]) > 0.999
    assert np.median(np.abs(a["coh_ps"] - b["coh_ps"])) < 1e-6


def test_small_baseline_nr_brute_force(ora):
    """The small-baseline Nr composes pinned pieces (RNG, FP64 row fit,
    histogram) on image-difference phases (native_stage2.rs:1321-1334): restate
    the composition with numpy and compare counts exactly."""
    from pystamps_amd.prepare import coh_bins, trial_values
    bp = np.array([-31.0, -12.5, 4.0, 17.0, 40.0, 55.5])
    ia = np.array([0, 0, 1, 2, 3, 1], np.int32)     # 0-based image ids
    ib = np.array([1, 2, 3, 4, 4, 4], np.int32)
    n_image, n_rows = 5, 400
    tv = trial_values(0.9)
    nr, mx = ora.random_coherence_histogram_sb(bp, ia, ib, n_image, tv, coh_bins(), n_rows=n_rows)
    img = ora.rng_uniform(n_rows * n_image).reshape(n_image, n_rows).T * (2.0 * np.pi)  # column-major
    ph = img[:, ib] - img[:, ia]
    coh = np.array([ora.topofit_phase_row_coh(ph[r], bp, tv) for r in range(n_rows)])
    ref = ora.hist_with_centers(coh, coh_bins())
    assert np.array_equal(nr, ref)
    nz = np.nonzero(ref)[0]
    assert mx == (nz[-1] + 1 if nz.size else 1)
    assert nr.sum() == n_rows


def test_small_baseline_indices_and_errors():
    from pystamps_amd.prepare import Stage2Error, small_baseline_random_indices
    a, b, n = small_baseline_random_indices(np.array([[1, 2], [2, 3], [1, 4]]), 3)
    assert list(a) == [0, 1, 0] and list(b) == [1, 2, 3] and n == 4
    a2, b2, n2 = small_baseline_random_indices(np.array([[1, 2, 1], [2, 3, 4]]), 3)  # 2 x n_ifg
    assert list(a2) == [0, 1, 0] and list(b2) == [1, 2, 3] and n2 == 4
    with pytest.raises(Stage2Error, match="positive one-based"):
        small_baseline_random_indices(np.array([[0, 2], [2, 3]]), 2)
    with pytest.raises(Stage2Error, match="incompatible shape"):
        small_baseline_random_indices(np.ones((3, 3)), 2)
    with pytest.raises(Stage2Error, match="missing"):
        small_baseline_random_indices(None, 2)


def test_library_phase_normalisation_equals_numpy_f32():
    """The library's threaded prep routine reproduces prepare_stage2_inputs'
    f32 arithmetic (native_stage2.rs:449-467; numpy float32 hypot is libm
    hypotf, divisions round once) bit for bit, including zeros, infinities,
    NaN and subnormals; and the column selection is a plain copy."""
    from pystamps_amd.prepare import normalize_phase, select_columns
    rng = np.random.default_rng(7)
    P, n = 3000, 9
    re = rng.standard_normal((P, n)).astype(np.float32) * np.float32(10.0) ** rng.integers(-40, 30, (P, n)).astype(np.float32)
    im = rng.standard_normal((P, n)).astype(np.float32)
    re[0, :3] = [0.0, np.inf, np.nan]
    im[0, :3] = [0.0, 1.0, 2.0]
    re[1, :2] = [1e-45, -0.0]
    im[1, :2] = [0.0, 0.0]
    ph = (re + 1j * im).astype(np.complex64)
    cols = [0, 2, 3, 5, 8, 1]
    got_nm, got_amp = normalize_phase(ph, cols)
    sel = ph[:, cols]
    r, i = sel.real.astype(np.float32), sel.imag.astype(np.float32)
    with np.errstate(all="ignore"):
        mag = np.hypot(r, i)
        safe = np.where(mag == np.float32(0.0), np.float32(1.0), mag).astype(np.float32)
        ref_re, ref_im = r / safe, i / safe
    assert np.array_equal(got_amp, safe, equal_nan=True)
    assert np.array_equal(got_nm.real, ref_re, equal_nan=True)
    assert np.array_equal(got_nm.imag, ref_im, equal_nan=True)
    m = rng.standard_normal((P, n))
    assert np.array_equal(select_columns(m, cols), m[:, cols])
