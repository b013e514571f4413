"""The full-size workloads of the stage-2 benchmark configurations against the
oracle and between the two GPU forms.  GPU.

  500,000 x 60 (13 trials, 299 x 299 grid): the whole 3-iteration loop of the
      FP32 form against ref64 (judged by ref64 on one-ulp-perturbed inputs)
      and against ref32 (tolerance manifest on rows ref32 does not tie); the
      FP64 form against ref64 exactly.
  4,000,000 x 120 (9 trials, 399 x 399 grid): one iteration of both forms
      against ref64; the whole 5-iteration loop of the FP32 form against the
      FP64 form, judged by the FP64 form on three perturbed inputs; invariants.
  1,000,000 x 200 (83 trials, 10 m grid, 999 x 999 cells): one iteration of
      the FP32 form against ref32 and ref64; the whole loop against the FP64
      form, judged by the FP64 form on three perturbed inputs.
  100,000 x 200 (the same generator, 316 x 316 cells): the whole loop of the
      FP32 form against ref64, judged by three FP64 loops on perturbed inputs.
  16 uneven patches: one batched call, every patch bit-identical to its own
      run; the two smallest in FP64 against ref64.
  Nr at the reference's 300,000 rows for the 20,000 x 30 and 500,000 x 60
      baselines: counts and Nr_max_nz_ix exact against the oracle.

The FP64 GPU form equals ref64 to ~1e-12 in every test above that runs both,
so at the two largest sizes (where a second FP64 oracle loop would take
minutes) the FP64 form on perturbed inputs stands in for the perturbed ref64.
"""
import dataclasses

import numpy as np
import pytest

from conftest import floor_check, perturb_one_ulp

pytestmark = pytest.mark.gpu

F64_ATOL, F64_RTOL = 1e-6, 1e-8
F32_ATOL, F32_RTOL = 1e-4, 1e-5


def _prep(config, n_ps=None, **kw):
    from pystamps_amd import synth
    from pystamps_amd.prepare import prepare_stage2_inputs
    cfg = synth.CONFIGS[config]
    if n_ps:
        cfg = synth.scaled(cfg, n_ps)
    prep = prepare_stage2_inputs(synth.generate(cfg), nr_rows=300_000)
    return dataclasses.replace(prep, **kw) if kw else prep


def _invariants(tag, o, P):
    coh = o["coh_ps"]
    assert np.all((coh >= 0.0) & (coh <= 1.0 + 1e-12)), tag
    mag = np.abs(o["ph_patch"])
    assert np.all((mag == 0.0) | (np.abs(mag - 1.0) < 1e-6)), tag
    assert np.all(np.isfinite(o["K_ps"]) == (o["N_opt"] == 1.0)), tag
    assert o["K_ps"].shape == (P,)


def _one_iteration_vs_ref64(S, ora, prep, tag):
    """FP64 form: trial indices exact, K / coh to 1e-9; FP32 form: trial
    indices equal and coh within 1e-5 on every row that ref64 does not tie
    (trial margin, or a least-squares residual phase within 1e-6 of the arg()
    cut), and within 1e-6 on all but a 1e-5 fraction of them.  (The FP32
    CLAP's error is absolute, ~2e-7 of a plane's largest value, so a cell whose
    filtered value nearly cancels carries a larger phase error: those rows
    reach a few 1e-6.)"""
    p1 = dataclasses.replace(prep, gamma_max_iterations=1)
    r = ora.run(p1, mode=ora.REF64)
    g64 = S.run_stage2_hip(p1, form="fp64")
    assert g64["i_loop"] == r["i_loop"] == 1
    assert np.array_equal(g64["trial_ix"], r["trial_ix"])
    np.testing.assert_allclose(g64["K_ps"], r["K_ps"], rtol=0, atol=1e-9)
    np.testing.assert_allclose(g64["coh_ps"], r["coh_ps"], rtol=0, atol=1e-9)
    g32 = S.run_stage2_hip(p1, form="fp32")
    tie = (r["trial_margin_min"] < 1e-6) | (r["refine_margin_min"] < 1e-6)
    mism = g32["trial_ix"] != r["trial_ix"]
    d = np.abs(g32["coh_ps"] - r["coh_ps"])
    print(f"{tag} one iteration fp32 vs ref64: ties {int(tie.sum())} mismatches {int(mism.sum())} "
          f"untied {int((mism & ~tie).sum())} coh max {d.max():.2e} p99.9 {np.quantile(d, .999):.2e}",
          flush=True)
    assert not np.any(mism & ~tie)
    du = d[~(mism | tie)]
    print(f"{tag}   untied rows above 1e-6: {int((du >= 1e-6).sum())}, max {du.max():.2e}", flush=True)
    assert np.all(du < 1e-5)
    assert (du >= 1e-6).mean() < 1e-5
    return p1, r, g32


@pytest.fixture(scope="module")
def S():
    from pystamps_amd import stage2
    return stage2


def test_c2_full_loop_vs_ref64_and_ref32(S, ora):
    prep = _prep(2)
    assert (prep.n_ps, prep.n_ifg, prep.n_i, len(prep.trial_values)) == (500_000, 60, 299, 13)
    g32 = S.run_stage2_hip(prep, form="fp32")
    g64 = S.run_stage2_hip(prep, form="fp64")
    r64 = ora.run(prep, mode=ora.REF64)
    assert g64["i_loop"] == r64["i_loop"]
    assert np.array_equal(g64["trial_ix"], r64["trial_ix"])
    np.testing.assert_allclose(g64["coh_ps"], r64["coh_ps"], rtol=0, atol=1e-9)
    r64p = ora.run(perturb_one_ulp(prep), mode=ora.REF64)
    c = floor_check("C2 loop fp32 vs ref64", g32, r64, r64p)
    assert c["coh_above_tol"] == 0
    # the reference's precision map: tolerance manifest on rows ref32 does not tie
    r32 = ora.run(prep, mode=ora.REF32)
    assert g32["i_loop"] == r32["i_loop"]
    bound = 2.0 * (8 + 8 + 2 * prep.n_ifg) * 2.0 ** -24
    clear = (r32["trial_margin_min"] > bound) & (r32["refine_margin_min"] > 1e-5)
    print("C2 ref32 ties", int((~clear).sum()), flush=True)
    assert np.array_equal(g32["trial_ix"][clear], r32["trial_ix"][clear])
    for key in ("K_ps", "coh_ps"):
        bad = (~(np.abs(g32[key] - r32[key]) <= F64_ATOL + F64_RTOL * np.abs(r32[key])) &
               ~(np.isnan(g32[key]) & np.isnan(r32[key])))[clear]
        assert not np.any(bad), (key, int(bad.sum()))
    _invariants("C2 fp32", g32, prep.n_ps)


def test_c4_one_iteration_vs_ref64_and_full_loop(S, ora):
    prep = _prep(4)
    assert (prep.n_ps, prep.n_ifg, prep.n_i, len(prep.trial_values)) == (4_000_000, 120, 399, 9)
    _one_iteration_vs_ref64(S, ora, prep, "C4")
    g32 = S.run_stage2_hip(prep, form="fp32")
    g64 = S.run_stage2_hip(prep, form="fp64")
    g64p = [S.run_stage2_hip(perturb_one_ulp(prep, seed), form="fp64") for seed in (5, 6, 7)]
    floor_check("C4 loop fp32 vs fp64", g32, g64, g64p)
    _invariants("C4 fp32", g32, prep.n_ps)
    _invariants("C4 fp64", g64, prep.n_ps)


def test_c5_one_iteration_vs_ref32_ref64_and_full_loop(S, ora):
    prep = _prep(5)
    assert (prep.n_ps, prep.n_ifg, prep.n_i, len(prep.trial_values)) == (1_000_000, 200, 999, 83)
    p1, r64, g32 = _one_iteration_vs_ref64(S, ora, prep, "C5")
    r32 = ora.run(p1, mode=ora.REF32)
    tc = 48
    bound = 2.0 * (tc + 8 + 2 * prep.n_ifg) * 2.0 ** -24
    # ref32's own ties: trial margin below the FP32 bound, or a residual phase
    # of its least-squares step within its f32 angle error (|k0 bp| up to ~32
    # rad here) of the arg() cut, where the FP32 form decides in FP64
    clear = (r32["trial_margin_min"] > bound) & (r32["refine_margin_min"] > 1e-5)
    print("C5 ref32 ties", int((~clear).sum()), flush=True)
    assert (~clear).mean() < 0.01
    assert np.array_equal(g32["trial_ix"][clear], r32["trial_ix"][clear])
    ok = np.abs(g32["K_ps"] - r32["K_ps"]) <= F64_ATOL + F64_RTOL * np.abs(r32["K_ps"])
    assert np.all((ok | (np.isnan(g32["K_ps"]) & np.isnan(r32["K_ps"])))[clear])
    g32 = S.run_stage2_hip(prep, form="fp32")
    g64 = S.run_stage2_hip(prep, form="fp64")
    g64p = [S.run_stage2_hip(perturb_one_ulp(prep, seed), form="fp64") for seed in (5, 6, 7)]
    floor_check("C5 loop fp32 vs fp64", g32, g64, g64p)
    _invariants("C5 fp32", g32, prep.n_ps)


def test_large_patch_snr_loop_vs_ref64(S, ora):
    """A patch above 65,536 rows with SNR weighting runs the large-patch
    kernels (scan, re-score, refine with its re-decision queue, k_redecide,
    k_post, the SNR weights): the FP64 form equals ref64 (trials exact,
    coherence to 1e-9); the FP32 form is judged against ref64 by the FP64
    loops on one-ulp-perturbed inputs."""
    from conftest import small_prep
    base = small_prep(n_ps=80_000, n_ifg=20, extent=12_000.0, seed=0x5EED1C01)
    prep = dataclasses.replace(base, filter_weighting="SNR")
    assert prep.n_ps > 65_536 and not prep.row_invariant_bperp
    r64 = ora.run(prep, mode=ora.REF64)
    g64 = S.run_stage2_hip(prep, form="fp64")
    assert g64["i_loop"] == r64["i_loop"]
    assert np.array_equal(g64["trial_ix"], r64["trial_ix"])
    np.testing.assert_allclose(g64["coh_ps"], r64["coh_ps"], rtol=0, atol=1e-9)
    g32 = S.run_stage2_hip(prep, form="fp32")
    rps = [ora.run(perturb_one_ulp(prep, seed), mode=ora.REF64) for seed in (5, 6)]
    floor_check("80k SNR loop fp32 vs ref64", g32, r64, rps)
    print("re-scored", g32["n_rescored"], "re-decided", g32["n_rechecked"])


def test_c5_scaled_full_loop_vs_ref64(S, ora):
    prep = _prep(5, n_ps=100_000)
    assert (prep.n_ifg, len(prep.trial_values), prep.n_i) == (200, 83, 316)
    g32 = S.run_stage2_hip(prep, form="fp32")
    r64 = ora.run(prep, mode=ora.REF64)
    # perturbation samples: ref64 on one, the FP64 form (ref64 to 1e-9) on two more
    rps = [ora.run(perturb_one_ulp(prep), mode=ora.REF64)] + \
        [S.run_stage2_hip(perturb_one_ulp(prep, seed), form="fp64") for seed in (6, 7)]
    floor_check("C5 100k loop fp32 vs ref64", g32, r64, rps)
    g64 = S.run_stage2_hip(prep, form="fp64")
    assert g64["i_loop"] == r64["i_loop"]
    assert np.array_equal(g64["trial_ix"], r64["trial_ix"])
    np.testing.assert_allclose(g64["coh_ps"], r64["coh_ps"], rtol=0, atol=1e-9)


def test_c3_batched_patches(S, ora):
    from pystamps_amd import synth
    from pystamps_amd.prepare import prepare_stage2_inputs
    cfgs = synth.c3_patches()
    preps = [prepare_stage2_inputs(synth.generate(c), nr_rows=300_000) for c in cfgs]
    assert len(preps) == 16 and sum(p.n_ps for p in preps) == sum(c.n_ps for c in cfgs)
    batch = S.run_stage2_hip_batch(preps, form="fp32")
    for k, (p, b) in enumerate(zip(preps, batch)):
        single = S.run_stage2_hip(p, form="fp32")
        assert b["i_loop"] == single["i_loop"], k
        for key in ("K_ps", "C_ps", "coh_ps", "N_opt", "trial_ix", "ph_patch", "ph_res", "ph_weight",
                    "ph_grid", "coh_ps_save", "Nr"):
            assert np.array_equal(b[key], single[key], equal_nan=True), (k, key)
    smallest = sorted(range(16), key=lambda k: preps[k].n_ps)[:2]
    b64 = S.run_stage2_hip_batch([preps[k] for k in smallest], form="fp64")
    for k, g in zip(smallest, b64):
        r = ora.run(preps[k], mode=ora.REF64)
        assert g["i_loop"] == r["i_loop"], k
        assert np.array_equal(g["trial_ix"], r["trial_ix"]), k
        np.testing.assert_allclose(g["coh_ps"], r["coh_ps"], rtol=0, atol=1e-9)


@pytest.mark.parametrize("config", [1, 2])
def test_nr_at_300000_rows(ora, config):
    """random_coherence_histogram at the reference's row count
    (native_stage2.rs:25-29, :1310-1356) for the baselines of the 20,000 x 30
    and 500,000 x 60 workloads, as the prep computes it on the GPU: counts and
    Nr_max_nz_ix exact against the oracle."""
    from conftest import oracle_nr
    prep = _prep(config)
    nr_o, mx_o = oracle_nr(prep.bperp_nm, prep.trial_values, prep.coh_bins, 300_000)
    print(f"config {config}: Nr sum {prep.nr_base.sum()} max_nz_ix {prep.nr_max_nz_ix}")
    assert prep.nr_base.sum() == 300_000
    assert np.array_equal(prep.nr_base, nr_o)
    assert prep.nr_max_nz_ix == mx_o


def test_fp64_loop_repeats_bit_identical(S):
    """The FP64 form's post step counts the coherence histogram in LDS: every
    bin is zeroed before any wavefront counts into it, so repeated loops of a
    patch on the k_post path (80,474 rows x 60), on their own and batched with
    a small patch, give the same bits every time (a count lost to the zeroing
    moved Nr and every later weight)."""
    from pystamps_amd import synth
    from pystamps_amd.prepare import prepare_stage2_inputs
    cfgs = synth.c3_patches()
    order = sorted(range(16), key=lambda k: cfgs[k].n_ps)[:2]
    two = [prepare_stage2_inputs(synth.generate(cfgs[k]), nr_rows=300_000) for k in order]
    assert two[1].n_ps > 65_536
    keys = ("K_ps", "coh_ps", "trial_ix", "Nr", "ph_weight")
    first = S.run_stage2_hip(two[1], form="fp64")
    runs = [S.run_stage2_hip(two[1], form="fp64") for _ in range(5)]
    runs += [S.run_stage2_hip_batch(two, form="fp64")[1] for _ in range(4)]
    for r in runs:
        assert r["gamma_change_save"] == first["gamma_change_save"]
        for key in keys:
            assert np.array_equal(r[key], first[key], equal_nan=True), key
