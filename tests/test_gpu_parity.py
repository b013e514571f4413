"""GPU parity: the HIP library against the oracle on the same seeded inputs.

Forms and what each is held to:
  FP64 form vs ref64 (the FP64 oracle): everything to ~1e-9; trial indices and
    iteration counts exact.
  FP32 form vs ref32 (the reference's precision map): the pm1 tolerance
    manifest (pystamps/data/artifact_tolerances.json:45-71: f64 keys atol 1e-6
    rtol 1e-8, f32 keys atol 1e-4 rtol 1e-5) on every row whose selected trial
    is not a near-tie in the reference's own FP32 scan; near-tie rows are
    counted, and must be few.
  FP32 form vs ref64: selected trial indices identical wherever ref64's margin
    exceeds the FP32 input perturbation; coh within 1e-5.
"""
import dataclasses

import numpy as np
import pytest

from conftest import small_prep

pytestmark = pytest.mark.gpu

F64_ATOL, F64_RTOL = 1e-6, 1e-8
F32_ATOL, F32_RTOL = 1e-4, 1e-5


def _close(a, b, atol, rtol):
    a = np.asarray(a)
    b = np.asarray(b)
    both_nan = np.isnan(a) & np.isnan(b) if np.isrealobj(a) else (np.isnan(a) & np.isnan(b))
    ok = np.abs(a - b) <= atol + rtol * np.abs(b)
    return ok | both_nan


def _phase_close(a, b, atol):
    d = (np.asarray(a, np.float64) - np.asarray(b, np.float64) + np.pi) % (2 * np.pi) - np.pi
    return np.abs(d) <= atol


@pytest.fixture(scope="module")
def S():
    from pystamps_amd import stage2
    return stage2


@pytest.fixture(scope="module")
def prep_c1():
    # config 1 at full size: 20,000 candidates, 30 interferograms
    return small_prep(n_ps=20_000, n_ifg=30, extent=5000.0, seed=0x5EED0001)


@pytest.fixture(scope="module")
def prep_small():
    return small_prep(n_ps=6000, n_ifg=24, seed=0x5EED1002)


# --- Nr ------------------------------------------------------------------------

def test_gpu_nr_golden():
    from pystamps_amd.prepare import coh_bins, nr_hip, trial_values
    from test_oracle_golden import NR_GOLDEN
    nr, mx = nr_hip([-20.0, -5.0, 0.0, 13.0, 27.0], trial_values(0.5), coh_bins(), n_rows=10_000)
    assert list(nr) == NR_GOLDEN
    assert mx == 100.0


def test_gpu_nr_matches_oracle(ora, prep_c1):
    from pystamps_amd.prepare import nr_hip
    p = prep_c1
    nr_g, mx_g = nr_hip(p.bperp_nm, p.trial_values, p.coh_bins, n_rows=30_000)
    nr_o, mx_o = ora.random_coherence_histogram(p.bperp_nm, p.trial_values, p.coh_bins, n_rows=30_000)
    assert np.array_equal(nr_g, nr_o)
    assert mx_g == mx_o


def test_gpu_nr_repeated_calls_across_shapes(ora):
    """Repeated calls that alternate row counts, interferogram counts and the
    small-baseline image count each equal the oracle's fresh draw."""
    from pystamps_amd.prepare import coh_bins, nr_hip, nr_hip_sb, trial_values
    tv, cb = trial_values(0.6), coh_bins()
    bp_a, bp_b = np.linspace(-150.0, 170.0, 9), np.linspace(-90.0, 260.0, 14)
    ia = np.arange(9, dtype=np.int32) % 8
    ib = (ia + 1).astype(np.int32)
    for kind, bp, rows in (("sm", bp_a, 7000), ("sm", bp_a, 7000), ("sm", bp_b, 7000),
                           ("sb", bp_a, 7000), ("sm", bp_a, 5000), ("sm", bp_a, 7000)):
        if kind == "sm":
            got = nr_hip(bp, tv, cb, n_rows=rows)
            ref = ora.random_coherence_histogram(bp, tv, cb, n_rows=rows)
        else:
            got = nr_hip_sb(bp, ia, ib, 9, tv, cb, n_rows=rows)
            ref = ora.random_coherence_histogram_sb(bp, ia, ib, 9, tv, cb, n_rows=rows)
        assert np.array_equal(got[0], ref[0]), (kind, bp.size, rows)
        assert got[1] == ref[1]


# --- state-injected steps ----------------------------------------------------------

def _state(prep, seed=11):
    rng = np.random.default_rng(seed)
    k = rng.uniform(-0.008, 0.008, prep.n_ps)
    w = rng.uniform(0.0, 1.0, prep.n_ps)
    k[::997] = np.nan   # invalid rows poison their cells
    return k, w


def test_fill_grid_fp64_vs_ref64(S, ora, prep_small):
    k, w = _state(prep_small)
    with S.Stage2Session(prep_small, form="fp64") as s:
        g = s.fill_grid(k, w)
    r = ora.accumulate_grid(prep_small, ora.fill_phase_weight(prep_small, k, w, mode=ora.REF64),
                            mode=ora.REF64)
    assert np.array_equal(np.isnan(g), np.isnan(r))
    m = ~np.isnan(r)
    np.testing.assert_allclose(g[m], r[m], rtol=1e-12, atol=1e-12)


def test_fill_grid_fp32_vs_ref32(S, ora, prep_small):
    k, w = _state(prep_small)
    with S.Stage2Session(prep_small, form="fp32") as s:
        g = s.fill_grid(k, w)
    r = ora.accumulate_grid(prep_small, ora.fill_phase_weight(prep_small, k, w, mode=ora.REF32),
                            mode=ora.REF32)
    assert np.array_equal(np.isnan(g), np.isnan(r))
    m = ~np.isnan(r)
    assert np.all(_close(g[m].real, r[m].real, F32_ATOL, F32_RTOL))
    assert np.all(_close(g[m].imag, r[m].imag, F32_ATOL, F32_RTOL))
    # same summation order: only the sincosf ulps of the fill can differ
    print("fill_grid fp32 max|diff|", np.abs(g[m] - r[m]).max(),
          "bit-identical fraction", np.mean(g[m] == r[m]))


def _grid_input(prep, seed=12):
    rng = np.random.default_rng(seed)
    C = prep.n_i * prep.n_j
    g = (rng.standard_normal((C, prep.n_ifg)) + 1j * rng.standard_normal((C, prep.n_ifg))) * 2.0
    g[5, :] = np.nan
    return g


def test_clap_fp64_vs_ref64(S, ora, prep_small):
    g = _grid_input(prep_small)
    with S.Stage2Session(prep_small, form="fp64") as s:
        f = s.clap(g)
    r = ora.clap_filter_grid_stack(prep_small, g, mode=ora.REF64)
    scale = np.abs(r).max()
    np.testing.assert_allclose(f, r, rtol=0, atol=1e-12 * scale)


def test_clap_fp32_vs_ref32(S, ora, prep_small):
    g = _grid_input(prep_small).astype(np.complex64)
    with S.Stage2Session(prep_small, form="fp32") as s:
        f = s.clap(g)
    r = ora.clap_filter_grid_stack(prep_small, g, mode=ora.REF32)
    err = np.abs(f - r)
    print("clap fp32 max|diff|", err.max(), "max|ref|", np.abs(r).max())
    assert np.all(_close(f.real, r.real, F32_ATOL, F32_RTOL))
    assert np.all(_close(f.imag, r.imag, F32_ATOL, F32_RTOL))
    # FP32 transforms: absolute error within 1e-6 of the plane's largest value
    assert err.max() <= 1e-6 * np.abs(r).max()


def test_clap_fp32_vs_ref32_full_size_grid(S, ora):
    """CLAP on a 299 x 299 grid with 60 interferograms (the 500k x 60
    workload's grid): the strips cover several windows each, so the partial
    planes alternate between 2 planes and the combine takes its 2-segment
    path; every cell within 1e-6 of the plane's largest value of ref32."""
    prep = small_prep(n_ps=30_000, n_ifg=60, extent=15_000.0, seed=0x5EED1902)
    assert (prep.n_i, prep.n_j) == (299, 299)
    g = _grid_input(prep).astype(np.complex64)
    with S.Stage2Session(prep, form="fp32") as s:
        f = s.clap(g)
    r = ora.clap_filter_grid_stack(prep, g, mode=ora.REF32)
    err = np.abs(f - r)
    print("clap fp32 299x299x60 max|diff|", err.max(), "max|ref|", np.abs(r).max())
    assert np.array_equal(np.isnan(f), np.isnan(r))
    assert err.max() <= 1e-6 * np.abs(r).max()
    # the FP64 form on the same grid (its own segmentation: one wavefront per SIMD)
    g64 = g.astype(np.complex128)
    with S.Stage2Session(prep, form="fp64") as s:
        f64 = s.clap(g64)
    r64 = ora.clap_filter_grid_stack(prep, g64, mode=ora.REF64)
    np.testing.assert_allclose(f64, r64, rtol=0, atol=1e-12 * np.abs(r64).max())


def _filt_input(ora, prep, mode):
    """A realistic filtered grid: the oracle's CLAP of the first fill."""
    k = np.zeros(prep.n_ps)
    w = 1.0 / prep.d_a
    g = ora.accumulate_grid(prep, ora.fill_phase_weight(prep, k, w, mode=mode), mode=mode)
    return ora.clap_filter_grid_stack(prep, g, mode=mode)


def test_fit_fp64_vs_ref64(S, ora, prep_c1):
    f = _filt_input(ora, prep_c1, ora.REF64)
    with S.Stage2Session(prep_c1, form="fp64") as s:
        o = s.fit(f)
    r = ora.fit_rows(prep_c1, ora.extract_patch_phase(prep_c1, f, mode=ora.REF64), mode=ora.REF64)
    assert np.array_equal(o["trial_ix"], r["trial_ix"])
    assert np.array_equal(o["N_opt"], r["N_opt"])
    np.testing.assert_allclose(o["K_ps"], r["K_ps"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(o["coh_ps"], r["coh_ps"], rtol=0, atol=1e-10)
    assert np.all(_phase_close(o["C_ps"], r["C_ps"], 1e-9))
    assert np.all(_phase_close(o["ph_res"], r["ph_res"], 1e-5))


def test_fit_fp32_vs_ref32_and_ref64(S, ora, prep_c1):
    f32 = _filt_input(ora, prep_c1, ora.REF32)
    with S.Stage2Session(prep_c1, form="fp32") as s:
        o = s.fit(f32)
    patch = ora.extract_patch_phase(prep_c1, f32, mode=ora.REF32)
    r32 = ora.fit_rows(prep_c1, patch, mode=ora.REF32)
    r64 = ora.fit_rows(prep_c1, patch.astype(np.complex128), mode=ora.REF64)
    assert np.array_equal(o["N_opt"], r32["N_opt"])
    # ref32's own FP32 margin: below this, its choice is FP32 rounding noise
    bound = 2.0 * (16 + 8 + 2 * prep_c1.n_ifg) * 2.0 ** -24
    clear = r32["trial_margin"] > bound
    print("fit fp32: rows", prep_c1.n_ps, "ref32 near-ties", int((~clear).sum()),
          "rescored", o["n_rescored"])
    assert (~clear).mean() < 0.01
    assert np.array_equal(o["trial_ix"][clear], r32["trial_ix"][clear])
    for key in ("K_ps", "coh_ps"):
        ok = _close(o[key], r32[key], F64_ATOL, F64_RTOL)
        assert np.all(ok[clear]), (key, np.abs(o[key] - r32[key])[clear].max())
    assert np.all(_phase_close(o["C_ps"], r32["C_ps"], F64_ATOL)[clear])
    assert np.all(_phase_close(o["ph_res"], r32["ph_res"], F32_ATOL)[clear])
    np.testing.assert_allclose(o["ph_patch"], patch, rtol=0, atol=1e-6)
    # FP64 re-score: the selected trial is ref64's wherever ref64 is not itself
    # tied at the level of the complex64 input rounding
    agree = o["trial_ix"] == r64["trial_ix"]
    print("fit fp32 vs ref64 trial agreement", agree.mean(), "mismatches", int((~agree).sum()))
    assert agree.mean() >= 0.9999
    np.testing.assert_allclose(o["coh_ps"][agree], r64["coh_ps"][agree], rtol=0, atol=1e-5)


def test_fit_fp32_large_patch_steps_from_bperp(S, ora):
    """80,000 rows with per-pixel bperp and 13 trials: the 8-trial FP32 scan
    computes its unit steps from bperp (the path for patches of 65,536 rows or
    more) instead of reading the step table.  Trial choice and sums as in the
    table path: ref32's trial wherever ref32 is not itself tied, ref64's
    coherence within 1e-5."""
    prep = small_prep(n_ps=80_000, n_ifg=20, extent=12_000.0, seed=0x5EED1901)
    assert prep.n_ps >= 65_536 and len(prep.trial_values) < 16 and not prep.row_invariant_bperp
    f32 = _filt_input(ora, prep, ora.REF32)
    with S.Stage2Session(prep, form="fp32") as s:
        o = s.fit(f32)
    patch = ora.extract_patch_phase(prep, f32, mode=ora.REF32)
    r32 = ora.fit_rows(prep, patch, mode=ora.REF32)
    r64 = ora.fit_rows(prep, patch.astype(np.complex128), mode=ora.REF64)
    bound = 2.0 * (8 + 8 + 2 * prep.n_ifg) * 2.0 ** -24
    clear = r32["trial_margin"] > bound
    print("large-patch fit fp32: near-ties", int((~clear).sum()), "rescored", o["n_rescored"])
    assert np.array_equal(o["N_opt"], r32["N_opt"])
    assert np.array_equal(o["trial_ix"][clear], r32["trial_ix"][clear])
    ok = _close(o["K_ps"], r32["K_ps"], F64_ATOL, F64_RTOL)
    assert np.all(ok[clear])
    agree = o["trial_ix"] == r64["trial_ix"]
    assert agree.mean() >= 0.9999
    np.testing.assert_allclose(o["coh_ps"][agree], r64["coh_ps"][agree], rtol=0, atol=1e-5)


@pytest.mark.parametrize("n_ps", [80_000, 20_000])
def test_fit_fp32_steps_from_bperp_equal_table_bits(S, ora, tmp_path, n_ps):
    """The unit steps the 8-trial FP32 scan can compute from bperp are the
    step table's values bit for bit: the same fit with the steps from the
    table and computed (chosen through the tile cache) gives the same trial,
    K and coherence bits, for a patch of 80,000 rows (one row per thread) and
    one of 20,000 (a quad of lanes per row)."""
    prep = small_prep(n_ps=n_ps, n_ifg=20, extent=12_000.0 * (n_ps / 80_000) ** 0.5, seed=0x5EED1901)
    f32 = _filt_input(ora, prep, ora.REF32)
    probe = tmp_path / "probe.txt"
    with S.Stage2Session(prep, form="fp32", autotune_cache=str(probe)) as s:
        t = s.tiles()
    key = [ln for ln in probe.read_text().splitlines() if ln.strip()][0].split(" fill_cells=")[0]
    fits = {}
    for table in (0, 1):
        path = tmp_path / f"t{table}.txt"
        path.write_text(f"{key} fill_cells={t['fill_cells']} scan_step_table={table} "
                        f"clap_segments={t['clap_segments']}\n")
        with S.Stage2Session(prep, form="fp32", autotune_cache=str(path)) as s:
            assert s.tiles()["scan_step_table"] == bool(table)
            fits[table] = s.fit(f32)
    assert np.array_equal(fits[0]["trial_ix"], fits[1]["trial_ix"])
    for key in ("K_ps", "coh_ps"):
        a = np.ascontiguousarray(fits[0][key], dtype=np.float64)
        b = np.ascontiguousarray(fits[1][key], dtype=np.float64)
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), key


@pytest.mark.parametrize("tn, n_ifg", [(1, 24), (2, 24), (3, 24), (4, 24), (5, 24), (6, 24), (7, 24),
                                       (12, 24), (20, 24), (30, 24), (41, 70), (55, 24)])
def test_fit_trial_chunks_fp64_and_fp32(S, ora, tn, n_ifg):
    """The scan's chunk paths by positive trial count: the 8-trial kernel
    compiled for each trial count 1..7, the 16/24/32/48-trial one-pass kernels
    and, beyond 47, the chunked scan re-anchored per chunk (70 interferograms
    also take the FP64 re-score past one 64-lane slot).
    FP64 form vs ref64 exact; FP32 form: ref32's trial wherever ref32 is not
    itself tied at the form's re-score bound, ref64's coherence within 1e-5."""
    from pystamps_amd.prepare import trial_values
    base = small_prep(n_ps=6000, n_ifg=n_ifg, seed=0x5EED1005)
    wraps = tn / 8.0
    prep = dataclasses.replace(base, n_trial_wraps=wraps, trial_values=trial_values(wraps))
    assert len(prep.trial_values) == 2 * tn + 1
    f64 = _filt_input(ora, prep, ora.REF64)
    with S.Stage2Session(prep, form="fp64") as s:
        o64 = s.fit(f64)
    r64 = ora.fit_rows(prep, ora.extract_patch_phase(prep, f64, mode=ora.REF64), mode=ora.REF64)
    assert np.array_equal(o64["trial_ix"], r64["trial_ix"])
    np.testing.assert_allclose(o64["K_ps"], r64["K_ps"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(o64["coh_ps"], r64["coh_ps"], rtol=0, atol=1e-10)
    f32 = _filt_input(ora, prep, ora.REF32)
    with S.Stage2Session(prep, form="fp32") as s:
        o32 = s.fit(f32)
    patch = ora.extract_patch_phase(prep, f32, mode=ora.REF32)
    r32 = ora.fit_rows(prep, patch, mode=ora.REF32)
    r64b = ora.fit_rows(prep, patch.astype(np.complex128), mode=ora.REF64)
    tc = next((c for c in (8, 16, 24, 32, 48) if tn < c), 48)
    bound = 2.0 * (tc + 8 + 2 * prep.n_ifg) * 2.0 ** -24
    clear = r32["trial_margin"] > bound
    print(f"tn {tn}: ref32 near-ties", int((~clear).sum()), "rescored", o32["n_rescored"])
    assert (~clear).mean() < 0.01
    assert np.array_equal(o32["trial_ix"][clear], r32["trial_ix"][clear])
    assert np.all(_close(o32["K_ps"], r32["K_ps"], F64_ATOL, F64_RTOL)[clear])
    agree = o32["trial_ix"] == r64b["trial_ix"]
    assert agree.mean() >= 0.9995
    np.testing.assert_allclose(o32["coh_ps"][agree], r64b["coh_ps"][agree], rtol=0, atol=1e-5)


def test_fit_real_row_fp64_and_fp32(S, ora):
    """The reference's real-data 75-interferogram row (tests/golden), fitted by
    the native-path kernels: 48 pixels carry the row with different constant
    phase offsets and noise levels, a unit filtered phase makes psdph the row
    itself.  FP64 form vs ref64 exact; FP32 form vs ref32 / ref64."""
    import json
    from pathlib import Path
    from pystamps_amd.prepare import trial_values
    d = json.loads((Path(__file__).resolve().parent / "golden" / "real_row_75ifg.json").read_text())
    row = np.asarray(d["cpx_re"]) + 1j * np.asarray(d["cpx_im"])
    bp = np.asarray(d["bperp"])
    wraps = float(d["n_trial_wraps"])
    base = small_prep(n_ps=48, n_ifg=75, seed=0x5EED1800)
    rng = np.random.default_rng(18)
    P = base.n_ps
    ph = row[None, :] * np.exp(1j * rng.uniform(-np.pi, np.pi, (P, 1)))
    ph = ph * np.exp(1j * rng.normal(0.0, 1.0, (P, 1)) * rng.uniform(0.0, 0.6, (P, 75)))
    ph_nm = (ph / np.abs(ph)).astype(np.complex64)
    prep = dataclasses.replace(base, ph_nm=ph_nm, amp=np.ones((P, 75), np.float32),
                               bperp_mat=np.tile(bp, (P, 1)), row_invariant_bperp=False, row_bperp_nm=bp,
                               bperp_nm=bp,
                               n_trial_wraps=wraps, trial_values=trial_values(wraps))
    unit = np.ones((prep.n_i * prep.n_j, 75), np.complex128)
    with S.Stage2Session(prep, form="fp64") as s:
        o64 = s.fit(unit)
    r64 = ora.fit_rows(prep, ora.extract_patch_phase(prep, unit, mode=ora.REF64), mode=ora.REF64)
    assert np.array_equal(o64["trial_ix"], r64["trial_ix"])
    np.testing.assert_allclose(o64["K_ps"], r64["K_ps"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(o64["coh_ps"], r64["coh_ps"], rtol=0, atol=1e-12)
    with S.Stage2Session(prep, form="fp32") as s:
        o32 = s.fit(unit.astype(np.complex64))
    r32 = ora.fit_rows(prep, ora.extract_patch_phase(prep, unit, mode=ora.REF32), mode=ora.REF32)
    assert np.array_equal(o32["trial_ix"], r64["trial_ix"])
    for key in ("K_ps", "coh_ps"):
        assert np.all(_close(o32[key], r32[key], F64_ATOL, F64_RTOL)), key
    np.testing.assert_allclose(o32["coh_ps"], r64["coh_ps"], rtol=0, atol=1e-5)


# --- whole loop ---------------------------------------------------------------------

@pytest.fixture(scope="module")
def loops(S, ora, prep_c1):
    from conftest import perturb_one_ulp
    return {
        "g64": S.run_stage2_hip(prep_c1, form="fp64"),
        "g32": S.run_stage2_hip(prep_c1, form="fp32"),
        "r64": ora.run(prep_c1, mode=ora.REF64),
        "r64p": ora.run(perturb_one_ulp(prep_c1), mode=ora.REF64),
        "r32": ora.run(prep_c1, mode=ora.REF32),
    }


def test_loop_fp64_vs_ref64(loops):
    g, r = loops["g64"], loops["r64"]
    assert g["i_loop"] == r["i_loop"]
    assert np.array_equal(g["trial_ix"], r["trial_ix"])
    assert np.array_equal(g["N_opt"], r["N_opt"])
    for key in ("K_ps", "coh_ps", "coh_ps_save"):
        np.testing.assert_allclose(g[key], r[key], rtol=0, atol=1e-9)
    assert np.all(_phase_close(g["C_ps"], r["C_ps"], 1e-8))
    np.testing.assert_allclose(g["gamma_change_save"], r["gamma_change_save"], rtol=1e-9)
    np.testing.assert_allclose(g["Nr"], r["Nr"], rtol=1e-12)
    for key in ("ph_grid", "ph_weight", "ph_patch"):
        np.testing.assert_allclose(g[key], r[key].astype(np.complex64), rtol=1e-5, atol=1e-5)
    assert np.all(_phase_close(g["ph_res"], r["ph_res"], 1e-4))


def test_loop_fp32_vs_ref32(loops, prep_c1):
    """The whole loop against the reference's precision map: every pm1 key
    within the tolerance manifest on every row, except the rows whose trial
    the reference's own FP32 scan ties (best - second below the FP32 form's
    re-score bound in some iteration) or whose least-squares step had a
    residual phase within 1e-5 rad of the arg() cut (the FP32 form takes those
    choices in FP64); those are counted and must be few."""
    g, r = loops["g32"], loops["r32"]
    assert g["i_loop"] == r["i_loop"]
    assert np.array_equal(g["N_opt"], r["N_opt"])
    bound = 2.0 * (16 + 8 + 2 * prep_c1.n_ifg) * 2.0 ** -24
    tie = (r["trial_margin_min"] <= bound) | (r["refine_margin_min"] <= 1e-5)
    clear = ~tie
    print("loop fp32 vs ref32: ref32 trial ties", int(tie.sum()), "of", tie.size)
    assert tie.mean() < 0.01
    assert np.array_equal(g["trial_ix"][clear], r["trial_ix"][clear])
    bad = {}
    for key in ("K_ps", "coh_ps", "coh_ps_save"):
        bad[key] = int((~_close(g[key], r[key], F64_ATOL, F64_RTOL))[clear].sum())
    bad["C_ps"] = int((~_phase_close(g["C_ps"], r["C_ps"], F64_ATOL))[clear].sum())
    for key in ("ph_weight", "ph_patch"):
        ok = _close(g[key].real, r[key].real, F32_ATOL, F32_RTOL) & _close(g[key].imag, r[key].imag,
                                                                          F32_ATOL, F32_RTOL)
        bad[key] = int((~ok)[clear].sum())
    bad["ph_res"] = int((~_phase_close(g["ph_res"], r["ph_res"], F32_ATOL))[clear].sum())
    okg = _close(g["ph_grid"].real, r["ph_grid"].real, F32_ATOL, F32_RTOL) & \
        _close(g["ph_grid"].imag, r["ph_grid"].imag, F32_ATOL, F32_RTOL)
    bad["ph_grid"] = int((~okg).sum())
    print("loop fp32 vs ref32 out-of-tolerance elements on untied rows", bad)
    assert not any(bad.values()), bad
    np.testing.assert_allclose(g["gamma_change_save"], r["gamma_change_save"], rtol=1e-6)
    np.testing.assert_allclose(g["Nr"], r["Nr"], rtol=1e-12)


def test_loop_fp32_vs_ref64(loops):
    """The whole loop against the FP64 oracle, judged by the FP64 oracle on
    one-ulp-perturbed inputs (input-quantisation ties); besides, every
    coherence within 1e-6 here."""
    from conftest import floor_check
    g, r = loops["g32"], loops["r64"]
    c = floor_check("C1 loop fp32 vs ref64", g, r, loops["r64p"])
    assert c["trial_mismatch"] == 0 and c["coh_max"] < 1e-6
    print("rescored", g["n_rescored"], "rechecked", g["n_rechecked"])


def test_session_rerun_is_deterministic(S, prep_small):
    with S.Stage2Session(prep_small, form="fp32") as s:
        s.run()
        a = s.download()
        s.run()
        b = s.download()
    for key in ("K_ps", "coh_ps", "ph_grid", "ph_weight", "ph_res"):
        assert np.array_equal(a[key], b[key], equal_nan=True), key


def test_three_ps_fixture_gpu(S):
    from pystamps_amd.prepare import prepare_stage2_inputs
    from conftest import oracle_nr
    from test_oracle_golden import three_ps_raw
    prep = prepare_stage2_inputs(three_ps_raw(), nr_fn=oracle_nr, nr_rows=10_000)
    for form in ("fp32", "fp64"):
        o = S.run_stage2_hip(prep, form=form)
        assert o["i_loop"] == 1
        assert np.all(np.isnan(o["K_ps"]))
        assert not np.any(o["coh_ps"]) and not np.any(o["N_opt"]) and not np.any(o["ph_res"])


def test_narrow_grid_with_windows_is_rejected(S, prep_small):
    from pystamps_amd.prepare import Stage2Error
    bad = dataclasses.replace(prep_small, n_i=20, n_j=400,
                              grid_lin=np.zeros(prep_small.n_ps, np.int64))
    with pytest.raises(Stage2Error, match="narrower than the 24-cell window"):
        S.run_stage2_hip(bad)


def test_row_invariant_and_snr_paths(S, ora):
    base = small_prep(n_ps=5000, n_ifg=16, seed=0x5EED1003)
    prep = dataclasses.replace(base, bperp_mat=None, row_invariant_bperp=True,
                               row_bperp_nm=base.bperp_nm.copy(), filter_weighting="SNR")
    g = S.run_stage2_hip(prep, form="fp64")
    r = ora.run(prep, mode=ora.REF64)
    assert g["i_loop"] == r["i_loop"]
    assert np.array_equal(g["trial_ix"], r["trial_ix"])
    np.testing.assert_allclose(g["coh_ps"], r["coh_ps"], rtol=0, atol=1e-9)
    # the FP32 form with one bperp row: the scan takes its unit steps from the
    # per-interferogram table (a stride slip there shows as wrong trials)
    g32 = S.run_stage2_hip(prep, form="fp32")
    assert g32["i_loop"] == r["i_loop"]
    agree = g32["trial_ix"] == r["trial_ix"]
    dcoh = np.abs(g32["coh_ps"] - r["coh_ps"])
    print("row-invariant fp32 vs ref64: trial agreement", agree.mean(), "coh p99.9",
          np.quantile(dcoh, 0.999))
    assert agree.mean() >= 0.999
    assert np.quantile(dcoh, 0.999) < 1e-5


def test_row_invariant_fit_fp32_vs_ref64(S, ora):
    """One trial fit (no loop) with row-invariant bperp and a realistic filtered
    grid: the FP32 trial choice equals ref64's wherever ref64 is not tied."""
    base = small_prep(n_ps=6000, n_ifg=24, seed=0x5EED1004)
    prep = dataclasses.replace(base, bperp_mat=None, row_invariant_bperp=True,
                               row_bperp_nm=base.bperp_nm.copy())
    f64 = _filt_input(ora, prep, ora.REF64)
    with S.Stage2Session(prep, form="fp32") as s:
        o = s.fit(f64.astype(np.complex64))
    patch = ora.extract_patch_phase(prep, f64.astype(np.complex64), mode=ora.REF32)
    r64 = ora.fit_rows(prep, patch.astype(np.complex128), mode=ora.REF64)
    agree = o["trial_ix"] == r64["trial_ix"]
    print("row-invariant fit fp32 vs ref64 trial agreement", agree.mean())
    assert agree.mean() >= 0.9999
    np.testing.assert_allclose(o["coh_ps"][agree], r64["coh_ps"][agree], rtol=0, atol=1e-5)


# --- batched patches ---------------------------------------------------------------

def test_batch_of_uneven_patches(S, ora):
    """Uneven patches in one batched call (concurrent sessions): each patch's
    result equals its own single-patch run bit for bit, and the FP64 form
    matches ref64 per patch, including per-patch iteration counts."""
    preps = [small_prep(n_ps=n, n_ifg=20, seed=0x5EED1100 + k)
             for k, n in enumerate((2500, 9000, 4000, 6500))]
    for form in ("fp32", "fp64"):
        batch = S.run_stage2_hip_batch(preps, form=form)
        assert len(batch) == len(preps)
        for p, b in zip(preps, batch):
            single = S.run_stage2_hip(p, form=form)
            assert b["i_loop"] == single["i_loop"]
            for key in ("K_ps", "C_ps", "coh_ps", "N_opt", "ph_grid", "ph_weight", "ph_res",
                        "trial_ix"):
                assert np.array_equal(b[key], single[key], equal_nan=True), (form, key)
            if form == "fp64":
                r = ora.run(p, mode=ora.REF64)
                assert b["i_loop"] == r["i_loop"]
                assert np.array_equal(b["trial_ix"], r["trial_ix"])
                np.testing.assert_allclose(b["coh_ps"], r["coh_ps"], rtol=0, atol=1e-9)


def test_batch_more_patches_than_workers_and_failure(S):
    """20 patches (more than the 16 batch workers): every result equals its
    single run.  A batch with one invalid patch fails naming that patch."""
    import dataclasses as dc
    preps = [small_prep(n_ps=1200 + 150 * k, n_ifg=8, seed=0x5EED1400 + k) for k in range(20)]
    batch = S.run_stage2_hip_batch(preps, form="fp32")
    assert len(batch) == 20
    for p, b in zip(preps, batch):
        single = S.run_stage2_hip(p, form="fp32")
        assert b["i_loop"] == single["i_loop"]
        for key in ("K_ps", "coh_ps", "trial_ix", "ph_patch"):
            assert np.array_equal(b[key], single[key], equal_nan=True), key
    bad = dc.replace(preps[5], grid_lin=preps[5].grid_lin.copy())
    bad.grid_lin[3] = bad.n_i * bad.n_j   # outside the grid: rejected on the host
    with pytest.raises(S.Stage2Error, match="patch 5: grid_lin"):
        S.run_stage2_hip_batch(preps[:5] + [bad] + preps[6:8], form="fp32")


def test_calls_from_several_host_threads(S, ora):
    """Sessions' first runs (graph capture), Nr histograms and registry
    topofits issued from different host threads at once: every result equals
    the same call made alone."""
    import threading
    from pystamps_amd import registry_hip
    from pystamps_amd.prepare import coh_bins, nr_hip, trial_values
    preps = [small_prep(n_ps=3000 + 500 * k, n_ifg=12, seed=0x5EED1600 + k) for k in range(4)]
    alone = [S.run_stage2_hip(p, form="fp32") for p in preps]
    rng = np.random.default_rng(9)
    cpx = np.exp(1j * rng.uniform(-np.pi, np.pi, (2000, 15)))
    bp = rng.uniform(-200, 200, (2000, 15))
    fit_alone = registry_hip.stage2_topofit(cpx, bp, 0.6)
    bpn, tv, cb = np.linspace(-120.0, 150.0, 11), trial_values(0.6), coh_bins()
    nr_alone = nr_hip(bpn, tv, cb, n_rows=8000)
    out, errors = {}, []

    def guard(fn):
        def run():
            try:
                fn()
            except Exception as e:  # noqa: BLE001 - reported below
                errors.append(e)
        return run

    def loops(k):
        for _ in range(2):
            with S.Stage2Session(preps[k], form="fp32") as s:
                s.run()
                out[("loop", k)] = s.download()

    def others():
        for q in range(3):
            out[("fit", q)] = registry_hip.stage2_topofit(cpx, bp, 0.6)
            out[("nr", q)] = nr_hip(bpn, tv, cb, n_rows=8000)

    threads = [threading.Thread(target=guard(lambda k=k: loops(k))) for k in range(4)]
    threads.append(threading.Thread(target=guard(others)))
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    for k in range(4):
        for key in ("K_ps", "coh_ps", "trial_ix", "ph_patch"):
            assert np.array_equal(out[("loop", k)][key], alone[k][key], equal_nan=True), (k, key)
    for q in range(3):
        for a, b in zip(out[("fit", q)], fit_alone):
            assert np.array_equal(a, b, equal_nan=True)
        assert np.array_equal(out[("nr", q)][0], nr_alone[0]) and out[("nr", q)][1] == nr_alone[1]


def test_sessions_run_concurrently_matches_sequential(S, prep_small):
    p2 = small_prep(n_ps=3000, n_ifg=24, seed=0x5EED1200)
    ss = [S.Stage2Session(prep_small, form="fp32"), S.Stage2Session(p2, form="fp32")]
    try:
        loops, secs = S.run_sessions(ss)
        conc = [s.download() for s in ss]
        seq = []
        for s in ss:
            il, _ = s.run()
            seq.append(s.download())
            assert il == loops[len(seq) - 1]
        assert all(x > 0 for x in secs)
        for a, b in zip(conc, seq):
            for key in ("K_ps", "coh_ps", "ph_weight", "trial_ix"):
                assert np.array_equal(a[key], b[key], equal_nan=True), key
    finally:
        for s in ss:
            s.close()


def test_merged_sessions_mixed_shapes_forms_and_limits(S):
    """Sessions that differ in shape, form, weighting, bperp layout, iteration
    limit, scan variant and CLAP (one grid has no windows) run as one merged
    launch sequence (ps2_hip_sessions_run): every patch equals its own run bit
    for bit, on the first call (graph built), the second (graph reused) and
    after one session is closed (a new graph for the rest)."""
    base = small_prep(n_ps=5000, n_ifg=16, seed=0x5EED1003)
    runs = [
        (small_prep(n_ps=4000, n_ifg=20, seed=0x5EED1A00), "fp32"),
        (small_prep(n_ps=7000, n_ifg=12, seed=0x5EED1A01), "fp64"),
        (dataclasses.replace(base, bperp_mat=None, row_invariant_bperp=True,
                             row_bperp_nm=base.bperp_nm.copy(), filter_weighting="SNR"), "fp32"),
        (dataclasses.replace(small_prep(n_ps=3000, n_ifg=20, seed=0x5EED1A02), gamma_max_iterations=1),
         "fp32"),
        (small_prep(n_ps=800, n_ifg=10, extent=600.0, seed=0x5EED1A03), "fp32"),
        (small_prep(n_ps=70_000, n_ifg=20, extent=12_000.0, seed=0x5EED1A04), "fp32"),
    ]
    assert runs[4][0].n_i < 19 and runs[5][0].n_ps >= 65536
    single = [S.run_stage2_hip(p, form=f) for p, f in runs]
    assert single[3]["i_loop"] == 1 and max(x["i_loop"] for x in single) > 1
    keys = ("K_ps", "C_ps", "coh_ps", "N_opt", "ph_grid", "ph_weight", "ph_res", "ph_patch", "trial_ix")
    ss = [S.Stage2Session(p, form=f) for p, f in runs]

    def check(idx, tag):
        loops, secs = S.run_sessions([ss[k] for k in idx])
        for q, k in enumerate(idx):
            o = ss[k].download()
            assert loops[q] == single[k]["i_loop"], (tag, k)
            assert o["gamma_change_save"] == single[k]["gamma_change_save"], (tag, k)
            for key in keys:
                assert np.array_equal(o[key], single[k][key], equal_nan=True), (tag, k, key)
        assert all(x > 0 for x in secs)

    try:
        check(list(range(6)), "first")
        check(list(range(6)), "reused")
        ss[1].close()
        check([0, 2, 3, 4, 5], "after close")
        check([5, 0], "two, reordered")
    finally:
        for s in ss:
            s.close()


# --- small-baseline branch ---------------------------------------------------------

def test_gpu_nr_small_baseline_matches_oracle(ora):
    from pystamps_amd.prepare import coh_bins, nr_hip_sb, trial_values
    bp = np.linspace(-180.0, 210.0, 24)
    rng = np.random.default_rng(5)
    ia = rng.integers(0, 12, 24).astype(np.int32)
    ib = ((ia + 1 + rng.integers(0, 6, 24)) % 13).astype(np.int32)
    tv = trial_values(0.8)
    nr_g, mx_g = nr_hip_sb(bp, ia, ib, 13, tv, coh_bins(), n_rows=20_000)
    nr_o, mx_o = ora.random_coherence_histogram_sb(bp, ia, ib, 13, tv, coh_bins(), n_rows=20_000)
    assert np.array_equal(nr_g, nr_o)
    assert mx_g == mx_o


def test_small_baseline_loop_fp64_vs_ref64(S, ora):
    """Small-baseline patch: no master column dropped, image-difference Nr,
    low-coherence threshold 15; the FP64 loop matches ref64."""
    from pystamps_amd import synth
    from pystamps_amd.prepare import prepare_stage2_inputs
    cfg = dataclasses.replace(synth.scaled(synth.CONFIGS[1], 6000, 16), seed=0x5EED1300)
    raw = synth.generate(cfg)
    n_full = raw.ph.shape[1]
    raw.parms = dict(raw.parms, small_baseline_flag="y")
    raw.ifgday_ix = np.stack([np.arange(1, n_full + 1), np.arange(2, n_full + 2)], axis=1)
    prep = prepare_stage2_inputs(
        raw, nr_rows=20_000,
        nr_sb_fn=lambda bp, a, b, ni, tv, cb, n: ora.random_coherence_histogram_sb(bp, a, b, ni, tv, cb, n_rows=n))
    assert prep.small_baseline and prep.low_coh_thresh == 15 and prep.n_ifg == n_full
    g = S.run_stage2_hip(prep, form="fp64")
    r = ora.run(prep, mode=ora.REF64)
    assert g["i_loop"] == r["i_loop"]
    assert np.array_equal(g["trial_ix"], r["trial_ix"])
    np.testing.assert_allclose(g["coh_ps"], r["coh_ps"], rtol=0, atol=1e-9)


# --- device-resident outputs --------------------------------------------------------

def _hip_to_host(ptr, shape, dtype):
    """Copies a device buffer to a new host array with the HIP runtime."""
    import ctypes
    rt = ctypes.CDLL("libamdhip64.so")
    rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    out = np.empty(shape, dtype)
    assert rt.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0   # hipMemcpyDeviceToHost
    return out


def test_export_device_matches_download(S, prep_small):
    with S.Stage2Session(prep_small, form="fp32") as s:
        s.run()
        host = s.download()
        dev = s.export_device()
        P, I = prep_small.n_ps, prep_small.n_ifg
        assert (dev["n_ps"], dev["n_ifg"], dev["i_loop"]) == (P, I, host["i_loop"])
        for key in ("K_ps", "C_ps", "coh_ps", "N_opt"):
            assert np.array_equal(_hip_to_host(dev[key], (P,), np.float64), host[key],
                                  equal_nan=True), key
        assert np.array_equal(_hip_to_host(dev["ph_patch"], (P, I), np.complex64), host["ph_patch"])
        assert np.array_equal(_hip_to_host(dev["ph_res"], (P, I), np.float32), host["ph_res"])
        assert np.array_equal(dev["Nr"], host["Nr"])


def test_large_patch_transfers_one_iteration_fp64_vs_ref64(S, ora):
    """300,000 x 60: the inputs and every plane output cross the bus in several
    pinned chunks.  One FP64 iteration against ref64, and the host download
    against the device export."""
    prep = dataclasses.replace(small_prep(n_ps=300_000, n_ifg=60, extent=25_000.0, seed=0x5EED1700),
                               gamma_max_iterations=1)
    P, I = prep.n_ps, prep.n_ifg
    assert P * I * 8 > 2 * (64 << 20)
    with S.Stage2Session(prep, form="fp64") as s:
        s.run()
        g = s.download()
        dev = s.export_device()
        assert np.array_equal(_hip_to_host(dev["ph_patch"], (P, I), np.complex64), g["ph_patch"])
        assert np.array_equal(_hip_to_host(dev["ph_res"], (P, I), np.float32), g["ph_res"])
        assert np.array_equal(_hip_to_host(dev["coh_ps"], (P,), np.float64), g["coh_ps"], equal_nan=True)
    r = ora.run(prep, mode=ora.REF64)
    assert g["i_loop"] == r["i_loop"] == 1
    assert np.array_equal(g["trial_ix"], r["trial_ix"])
    for key in ("K_ps", "coh_ps", "coh_ps_save"):
        np.testing.assert_allclose(g[key], r[key], rtol=0, atol=1e-9)
    for key in ("ph_grid", "ph_weight", "ph_patch"):
        np.testing.assert_allclose(g[key], r[key].astype(np.complex64), rtol=1e-5, atol=1e-5)
    assert np.all(_phase_close(g["ph_res"], r["ph_res"], 1e-4))


# --- inputs the reference treats specially --------------------------------------------

def _edge_prep(ora, n_ps=5000, n_ifg=12, seed=0x5EED1500, nan_frac=0.0, da_zero_frac=0.0):
    from pystamps_amd import synth
    from pystamps_amd.prepare import prepare_stage2_inputs
    from conftest import oracle_nr
    cfg = dataclasses.replace(synth.scaled(synth.CONFIGS[1], n_ps, n_ifg), seed=seed)
    raw = synth.generate(cfg)
    rng = np.random.default_rng(seed)
    if nan_frac:
        m = rng.random(raw.ph.shape) < nan_frac
        m[:, raw.master_ix - 1] = False
        raw.ph = raw.ph.copy()
        raw.ph[m] = np.complex64(complex(np.nan, np.nan))
    if da_zero_frac:
        d = np.ones(n_ps) if raw.d_a is None else np.array(raw.d_a, dtype=np.float64)
        d[rng.random(n_ps) < da_zero_frac] = 0.0
        raw.d_a = d
    return prepare_stage2_inputs(raw, nr_fn=oracle_nr, nr_rows=5000)


@pytest.mark.parametrize("case", ["nan_phases", "zero_da", "two_ifgs"])
def test_edge_inputs_fp64_vs_ref64(S, ora, case):
    """NaN phases (NaN grid cells, read as 0 by CLAP, native_stage2.rs:852-858),
    D_A = 0 (initial weight 0, :242-246) and a 2-interferogram patch: the FP64
    loop equals ref64; the FP32 loop selects ref64's trials almost everywhere."""
    kw = {"nan_phases": dict(nan_frac=0.002), "zero_da": dict(da_zero_frac=0.05),
          "two_ifgs": dict(n_ifg=2)}[case]
    prep = _edge_prep(ora, **kw)
    g = S.run_stage2_hip(prep, form="fp64")
    r = ora.run(prep, mode=ora.REF64)
    assert g["i_loop"] == r["i_loop"]
    assert np.array_equal(g["trial_ix"], r["trial_ix"])
    assert np.array_equal(g["N_opt"], r["N_opt"])
    np.testing.assert_allclose(g["coh_ps"], r["coh_ps"], rtol=0, atol=1e-9)
    np.testing.assert_allclose(g["K_ps"], r["K_ps"], rtol=0, atol=1e-9, equal_nan=True)
    gw, rw = g["ph_weight"], r["ph_weight"].astype(np.complex64)
    assert np.array_equal(np.isnan(gw), np.isnan(rw))
    # FP32 form: ref64's trials after one iteration; over three, P-square
    # lookup flips may cascade (with 2 interferograms the coherence is coarse)
    one = dataclasses.replace(prep, gamma_max_iterations=1)
    assert np.mean(S.run_stage2_hip(one, form="fp32")["trial_ix"] ==
                   ora.run(one, mode=ora.REF64)["trial_ix"]) >= 0.9995
    g32 = S.run_stage2_hip(prep, form="fp32")
    assert np.mean(g32["trial_ix"] == r["trial_ix"]) > 0.995


@pytest.mark.parametrize("form", ["fp32", "fp64"])
def test_small_patch_fused_loop_equals_separate_steps(S, prep_small, form):
    """A patch of at most 65,536 rows runs its loop with a fused kernel
    (re-score + refine + post step in one workgroup per row group: 64 rows
    with a quad of lanes per row in the FP32 form, 256 rows with a thread per
    row in the FP64 form): a one-iteration loop equals the same iteration done
    by the separate step kernels (fill with K = 0 and the 1/D_A weighting,
    CLAP, fit; a thread per row) bit for bit."""
    assert prep_small.n_ps <= 65536
    one = dataclasses.replace(prep_small, gamma_max_iterations=1)
    with S.Stage2Session(one, form=form) as s:
        il, _ = s.run()
        loop = s.download()
        assert il == 1
        w0 = np.zeros(one.n_ps)
        nz = one.d_a != 0
        w0[nz] = 1.0 / one.d_a[nz]
        fit = s.fit(s.clap(s.fill_grid(np.zeros(one.n_ps), w0)))
    for key in ("K_ps", "C_ps", "coh_ps", "N_opt", "trial_ix", "ph_patch", "ph_res"):
        assert np.array_equal(loop[key], fit[key], equal_nan=True), key


# --- other CLAP windows (clap_win other than 32) -------------------------------

def _prep_clap_win(ora, clap_win, n_ps=6000, n_ifg=16, seed=0x5EED1A40):
    from pystamps_amd import synth
    from pystamps_amd.prepare import prepare_stage2_inputs
    cfg = dataclasses.replace(synth.scaled(synth.CONFIGS[1], n_ps, n_ifg), seed=seed)
    raw = synth.generate(cfg)
    raw.parms = dict(raw.parms, clap_win=clap_win)
    prep = prepare_stage2_inputs(raw, nr_rows=10_000,
                                 nr_fn=lambda bp, tv, cb, n: ora.random_coherence_histogram(bp, tv, cb, n_rows=n))
    assert prep.low_pass.shape[0] == int(np.floor(clap_win + 0.5))
    return prep


@pytest.mark.parametrize("clap_win", [16.0, 22.0, 27.0, 64.0])
def test_clap_other_windows_vs_oracle(S, ora, clap_win):
    """clap_win 16 (12 of 16), 22 (an odd 17-cell window in 22 points), 27
    (20 of 27, an odd transform) and 64 (48 of 64) run the generic CLAP
    kernels: FP64 within 1e-12 of ref64's largest value, FP32 within 1e-6 of
    ref32's."""
    prep = _prep_clap_win(ora, clap_win)
    g = _grid_input(prep)
    with S.Stage2Session(prep, form="fp64") as s:
        f64 = s.clap(g)
    r64 = ora.clap_filter_grid_stack(prep, g, mode=ora.REF64)
    scale = np.abs(r64).max()
    assert np.array_equal(np.isnan(f64), np.isnan(r64))
    np.testing.assert_allclose(f64, r64, rtol=0, atol=1e-12 * scale)
    g32 = g.astype(np.complex64)
    with S.Stage2Session(prep, form="fp32") as s:
        f32 = s.clap(g32)
    r32 = ora.clap_filter_grid_stack(prep, g32, mode=ora.REF32)
    err = np.abs(f32 - r32).max() / np.abs(r32).max()
    print(f"clap_win {clap_win}: fp32 max|diff| / max|ref32| = {err:.2e}")
    assert err <= 1e-6


def test_loop_other_window_vs_ref64(S, ora):
    """A whole loop with clap_win 22 (an odd 17-cell window): the FP64 form
    equals ref64 (iterations and trial indices exact, coherence to 1e-9); the
    FP32 form takes the same iterations and trial indices on all but
    near-tied rows, coherence within 1e-5 where the trials agree."""
    prep = _prep_clap_win(ora, 22.0, n_ps=8000, n_ifg=20, seed=0x5EED1A41)
    r = ora.run(prep, mode=ora.REF64)
    g = S.run_stage2_hip(prep, form="fp64")
    assert g["i_loop"] == r["i_loop"]
    assert np.array_equal(g["trial_ix"], r["trial_ix"])
    np.testing.assert_allclose(g["coh_ps"], r["coh_ps"], rtol=0, atol=1e-9)
    g32 = S.run_stage2_hip(prep, form="fp32")
    assert g32["i_loop"] == r["i_loop"]
    agree = g32["trial_ix"] == r["trial_ix"]
    print("clap_win 22 fp32 loop: trial agreement", agree.mean(), "mismatches", int((~agree).sum()))
    assert agree.mean() >= 0.999
    np.testing.assert_allclose(g32["coh_ps"][agree], r["coh_ps"][agree], rtol=0, atol=1e-5)


def test_fp32_f64_baselines_split_and_plane_agree(S, prep_small):
    """The FP32 form keeps the f64 baselines of its FP64 re-score as the f32
    bulk value plus an int32 residual (exact for every value whose f32 image
    is normal or zero) and falls back to an f64 plane when some value does not
    split.  A baseline of 1e-300 (f32 image 0, not split) puts the patch on
    the plane; the same patch with 0 there runs the split: the outputs agree
    bit for bit (the two f64 values differ by 1e-300, below every sum's
    rounding)."""
    assert prep_small.bperp_mat is not None and not prep_small.row_invariant_bperp
    outs = []
    for v in (0.0, 1e-300):
        bm = np.array(prep_small.bperp_mat, copy=True)
        bm[7, 3] = v
        p = dataclasses.replace(prep_small, bperp_mat=bm)
        with S.Stage2Session(p, form="fp32") as s:
            s.run()
            outs.append(s.download())
    a, b = outs
    assert a["n_rescored"] > 0 and a["i_loop"] == b["i_loop"]
    for key in ("K_ps", "C_ps", "coh_ps", "trial_ix", "ph_patch", "ph_res", "ph_weight", "ph_grid", "Nr"):
        assert np.array_equal(a[key], b[key], equal_nan=True), key


@pytest.mark.parametrize("form", ["fp32", "fp64"])
def test_loop_beyond_one_graph_equals_stepwise(S, form):
    """More than 64 iterations: one captured iteration is launched once per
    iteration after the starting state is set with copies, the outputs after
    it, and the control block is copied down (no graph-held starting state or
    control-block write).  With no convergence stop the loop runs all 70
    iterations; the graph launches, their replay and the timed kernel-by-kernel
    launches give the same outputs bit for bit."""
    from conftest import small_prep
    prep = dataclasses.replace(small_prep(n_ps=2500, n_ifg=12, seed=0x5EED1B70), gamma_max_iterations=70,
                               gamma_change_convergence=0.0)
    keys = ("K_ps", "C_ps", "coh_ps", "ph_grid", "ph_weight", "ph_res", "Nr", "trial_ix", "coh_ps_save")
    with S.Stage2Session(prep, form=form) as s:
        il, _ = s.run()
        a = s.download()
        s.set_timing(True)
        s.run()
        b = s.download()
        kt = s.kernel_times()
        s.set_timing(False)
        s.run()
        c = s.download()
    assert il == 70 and a["i_loop"] == b["i_loop"] == c["i_loop"] == 70
    assert kt["fill_grid"][1] == 70 and kt["outputs"][1] == 1
    for key in keys:
        assert np.array_equal(a[key], b[key], equal_nan=True), key
        assert np.array_equal(a[key], c[key], equal_nan=True), key
    assert a["gamma_change_save"] == b["gamma_change_save"] == c["gamma_change_save"]


def test_iteration_snapshots_with_timing(S, prep_small):
    """The loop with the state after every iteration (run_iterations), with
    per-kernel timing on: one snapshot per iteration, the last equal to the
    graph-launched loop's outputs, and the timers count every iteration's
    launches and the outputs written after each."""
    seen = []
    with S.Stage2Session(prep_small, form="fp32") as s:
        s.set_timing(True)
        n = s.run_iterations(lambda it, st: seen.append((it, st["coh_ps"].copy(), st["K_ps"].copy())))
        kt = s.kernel_times()
        s.set_timing(False)
        s.run()
        full = s.download()
    assert n == full["i_loop"] and [it for it, _, _ in seen] == list(range(1, n + 1))
    assert kt["fill_grid"][1] == n and kt["clap_strip"][1] == n and kt["outputs"][1] == n
    assert np.array_equal(seen[-1][1], full["coh_ps"])
    assert np.array_equal(seen[-1][2], full["K_ps"], equal_nan=True)


def test_graph_replays_equal_stepwise_launches(S, prep_small):
    """The default loop replays the captured iterations; with timing on the
    same kernels are launched one by one between event pairs, every iteration
    queued before one host wait; with verbose output the host reads the
    control block after each iteration.  Every replay, with eager work and
    host synchronisation in between, gives the step-wise result bit for bit,
    and the timers count the launches of the iterations that ran (those
    queued after the stop decision return at once and are not counted)."""
    keys = ("K_ps", "C_ps", "coh_ps", "ph_grid", "ph_weight", "ph_res", "Nr", "trial_ix")
    with S.Stage2Session(prep_small, form="fp32", verbose=True) as s:
        s.run()
        host_read = s.download()
    with S.Stage2Session(prep_small, form="fp32") as s:
        s.run()
        first = s.download()
        s.set_timing(True)
        s.run()
        stepwise = s.download()
        kt = s.kernel_times()
        s.set_timing(False)
        n_it = stepwise["i_loop"]
        assert n_it >= 2
        assert kt["fill_grid"][1] == n_it and kt["clap_strip"][1] == n_it
        # a small patch's P-square weights are computed in the next fill
        fused = prep_small.filter_weighting.lower() == "p-square" and prep_small.n_ps <= 65536
        assert kt["weights"][1] == (0 if fused else n_it)
        assert kt["outputs"][1] == 1
        assert all(ms > 0.0 for ms, n in kt.values() if n)
        for key in keys:
            assert np.array_equal(host_read[key], stepwise[key], equal_nan=True), key
        for _ in range(3):
            _ = s.fill_grid(np.zeros(prep_small.n_ps), np.ones(prep_small.n_ps))  # eager work between replays
            il, _ = s.run()
            again = s.download()
            assert il == stepwise["i_loop"]
            for key in keys:
                assert np.array_equal(again[key], stepwise[key], equal_nan=True), key
        for key in keys:
            assert np.array_equal(first[key], stepwise[key], equal_nan=True), key
        assert first["gamma_change_save"] == stepwise["gamma_change_save"]
