"""PATCH_* MAT v5 I/O (read_patch / write_patch / write_pm1) and the directory
drop-in run_stage2_patch_dir.

CPU tests: a synthetic patch written and read back prepares identically; the
reference's 3-PS fixture (native_stage2.rs:2698-2769) read from files gives its
known degenerate outcome; pm1.mat holds write_pm1's variables with the
reference's MAT classes and shapes (native_stage2.rs:1899-1965); the
reference's input error messages.  GPU test: run_stage2_patch_dir writes a
pm1.mat equal to the in-memory loop's result.
"""
import math

import numpy as np
import pytest
import scipy.io as sio

from conftest import oracle_nr


def _synthetic_raw(n_ps=2500, n_ifg=12, seed=0x5EED1400):
    from dataclasses import replace
    from pystamps_amd import synth
    cfg = replace(synth.scaled(synth.CONFIGS[1], n_ps, n_ifg), seed=seed)
    return synth.generate(cfg)


def test_patch_round_trip_prepares_identically(tmp_path):
    from pystamps_amd.patchio import read_patch, write_patch
    from pystamps_amd.prepare import prepare_stage2_inputs
    raw = _synthetic_raw()
    write_patch(tmp_path / "PATCH_1", raw)
    back = read_patch(tmp_path / "PATCH_1")
    assert np.array_equal(back.ph, raw.ph.astype(np.complex64))
    assert np.array_equal(back.xy, raw.xy.astype(np.float32))
    assert back.master_ix == raw.master_ix
    a = prepare_stage2_inputs(raw, nr_fn=oracle_nr, nr_rows=2000)
    b = prepare_stage2_inputs(back, nr_fn=oracle_nr, nr_rows=2000)
    for key in ("ph_nm", "amp", "row_bperp_nm", "grid_ij", "grid_lin", "d_a", "low_pass",
                "trial_values", "nr_base"):
        va, vb = getattr(a, key), getattr(b, key)
        assert np.array_equal(va, vb), key
    assert (a.n_i, a.n_j, a.n_trial_wraps, a.gamma_max_iterations, a.filter_weighting) == \
           (b.n_i, b.n_j, b.n_trial_wraps, b.gamma_max_iterations, b.filter_weighting)
    if a.bperp_mat is None:
        assert b.bperp_mat is None
    else:
        assert np.array_equal(a.bperp_mat, b.bperp_mat)


def test_transposed_inputs_are_oriented(tmp_path):
    """ph1.ph stored n_ifg x n_ps and ps1.xy stored 3 x n_ps are transposed on
    read, as ps_complex_matrix / ps_dim_f32 do (native_stage2.rs:2300-2360)."""
    from pystamps_amd.patchio import read_patch
    from pystamps_amd.prepare import prepare_stage2_inputs
    rng = np.random.default_rng(3)
    n_ps, n_ifg = 5, 4
    ph = np.exp(1j * rng.uniform(-np.pi, np.pi, (n_ps, n_ifg))).astype(np.complex64)
    xy = np.stack([np.arange(1, n_ps + 1), rng.uniform(0, 900, n_ps),
                   rng.uniform(0, 900, n_ps)], axis=1).astype(np.float32)
    d = tmp_path / "PATCH_T"
    d.mkdir()
    sio.savemat(d / "ps1.mat", {"n_ps": float(n_ps), "master_ix": 1.0,
                                "bperp": np.array([[0.0, 12.0, -30.0, 44.0]]),
                                "xy": xy.T.astype(np.float64)})
    sio.savemat(d / "ph1.mat", {"ph": ph.T})
    back = read_patch(d)
    assert np.array_equal(back.ph, ph)
    prep = prepare_stage2_inputs(back, nr_fn=oracle_nr, nr_rows=100)
    assert prep.n_ps == n_ps and prep.n_ifg == n_ifg - 1


def test_three_ps_fixture_from_files(tmp_path, ora):
    """The reference's end-to-end fixture, as files: empty CLAP, every row
    invalid, i_loop = 1 (native_stage2.rs:2561-2614)."""
    from pystamps_amd.patchio import read_patch, write_patch, write_pm1
    from pystamps_amd.prepare import prepare_stage2_inputs
    from test_oracle_golden import three_ps_raw
    d = write_patch(tmp_path / "PATCH_1", three_ps_raw())
    prep = prepare_stage2_inputs(read_patch(d), nr_fn=oracle_nr, nr_rows=1000)
    o = ora.run(prep, mode=ora.REF32)
    assert o["i_loop"] == 1
    write_pm1(d, prep, o)
    m = sio.loadmat(d / "pm1.mat")
    assert np.all(np.isnan(m["K_ps"])) and not np.any(m["coh_ps"])
    assert float(m["i_loop"][0, 0]) == 1.0


def test_pm1_variables_classes_and_shapes(tmp_path, ora):
    from pystamps_amd.patchio import write_pm1
    from conftest import small_prep
    prep = small_prep(n_ps=2000, n_ifg=10)
    o = ora.run(prep, mode=ora.REF32)
    write_pm1(tmp_path, prep, o)
    m = sio.loadmat(tmp_path / "pm1.mat")
    P, I = prep.n_ps, prep.n_ifg
    expect = {
        "K_ps": ((P, 1), np.float64), "C_ps": ((P, 1), np.float64), "coh_ps": ((P, 1), np.float64),
        "N_opt": ((P, 1), np.float64), "ph_res": ((P, I), np.float32),
        "ph_patch": ((P, I), np.complex64), "step_number": ((1, 1), np.float64),
        "ph_grid": ((prep.n_i, prep.n_j, I), np.complex64), "n_trial_wraps": ((1, 1), np.float32),
        "grid_ij": ((P, 2), np.float32), "grid_size": ((1, 1), np.float64),
        "low_pass": ((32, 32), np.float64), "i_loop": ((1, 1), np.float64),
        "ph_weight": ((P, I), np.complex64), "Nr": ((1, 100), np.float64),
        "Nr_max_nz_ix": ((1, 1), np.float64), "coh_bins": ((1, 100), np.float64),
        "coh_ps_save": ((P, 1), np.float64), "gamma_change_save": ((1, 1), np.float64),
    }
    for k, (shape, dt) in expect.items():
        assert m[k].shape == shape, k
        assert m[k].dtype == dt, k
    # values: cell (i, j) of ph_grid is row i * n_j + j of the loop's [C][I] grid
    g = o["ph_grid"].reshape(prep.n_i, prep.n_j, I)
    assert np.array_equal(m["ph_grid"], g.astype(np.complex64))
    assert np.array_equal(m["K_ps"][:, 0], o["K_ps"], equal_nan=True)
    assert np.array_equal(m["ph_res"], o["ph_res"].astype(np.float32))
    assert float(m["n_trial_wraps"][0, 0]) == np.float32(prep.n_trial_wraps)


def test_reference_input_errors(tmp_path):
    from pystamps_amd.patchio import read_patch, write_patch
    from pystamps_amd.prepare import Stage2Error, prepare_stage2_inputs
    from test_oracle_golden import three_ps_raw
    with pytest.raises(Stage2Error, match="unable to read ps1.mat"):
        read_patch(tmp_path / "missing")
    d = tmp_path / "P0"
    d.mkdir()
    sio.savemat(d / "ps1.mat", {"n_ps": 0.0})
    sio.savemat(d / "ph1.mat", {"ph": np.ones((3, 3), np.complex64)})
    with pytest.raises(Stage2Error, match="ps1.mat missing valid n_ps"):
        read_patch(d)
    d = write_patch(tmp_path / "P1", three_ps_raw(bad_bp_shape=True))
    with pytest.raises(Stage2Error, match="bp1.bperp_mat has incompatible shape"):
        prepare_stage2_inputs(read_patch(d), nr_fn=oracle_nr, nr_rows=100)
    r = three_ps_raw()
    r.master_ix = 7
    d = write_patch(tmp_path / "P2", r)
    with pytest.raises(Stage2Error, match="master_ix must be 1-based"):
        prepare_stage2_inputs(read_patch(d), nr_fn=oracle_nr, nr_rows=100)


def test_parms_text_and_scalars(tmp_path):
    from pystamps_amd.patchio import read_parms
    sio.savemat(tmp_path / "parms.mat", {"filter_weighting": "SNR", "gamma_max_iterations": 7.0,
                                         "small_baseline_flag": "n", "clap_beta": 0.25})
    p = read_parms(tmp_path)
    assert p["filter_weighting"] == "SNR" and p["gamma_max_iterations"] == 7.0
    assert p["small_baseline_flag"] == "n" and math.isclose(p["clap_beta"], 0.25)


@pytest.mark.gpu
def test_run_stage2_patch_dir_matches_in_memory(tmp_path):
    from pystamps_amd.patchio import read_patch, run_stage2_patch_dir, write_patch
    from pystamps_amd.prepare import prepare_stage2_inputs
    from pystamps_amd.stage2 import run_stage2_hip
    raw = _synthetic_raw(n_ps=4000, n_ifg=14, seed=0x5EED1401)
    d = write_patch(tmp_path / "PATCH_1", raw)
    msg = run_stage2_patch_dir(d, form="fp32", nr_rows=20_000)
    assert msg.startswith("Stage 2 computed coherence for 4000 candidates in ")
    m = sio.loadmat(d / "pm1.mat")
    prep = prepare_stage2_inputs(read_patch(d), nr_rows=20_000)
    o = run_stage2_hip(prep, form="fp32")
    assert float(m["i_loop"][0, 0]) == o["i_loop"]
    assert np.array_equal(m["K_ps"][:, 0], o["K_ps"], equal_nan=True)
    assert np.array_equal(m["coh_ps"][:, 0], o["coh_ps"])
    assert np.array_equal(m["ph_grid"], o["ph_grid"].reshape(prep.n_i, prep.n_j, prep.n_ifg))


@pytest.mark.gpu
def test_debug_pm_iteration_snapshots(tmp_path, monkeypatch):
    """With PYSTAMPS_STAGE2_NATIVE_DEBUG_PM set, run_stage2_patch_dir also
    writes pm1_iter_NN.mat after every iteration (native_stage2.rs:259,
    :337-404): one per iteration, the last equal to pm1.mat, the first equal
    to the pm1.mat of a one-iteration loop."""
    import dataclasses
    from pystamps_amd.patchio import read_patch, run_stage2_patch_dir, write_patch
    from pystamps_amd.prepare import prepare_stage2_inputs
    from pystamps_amd.stage2 import run_stage2_hip
    raw = _synthetic_raw(n_ps=4000, n_ifg=14, seed=0x5EED1402)
    d = write_patch(tmp_path / "PATCH_1", raw)
    monkeypatch.setenv("PYSTAMPS_STAGE2_NATIVE_DEBUG_PM", "1")
    run_stage2_patch_dir(d, form="fp32", nr_rows=20_000)
    final = sio.loadmat(d / "pm1.mat")
    n = int(final["i_loop"][0, 0])
    assert n >= 2
    snaps = sorted(p.name for p in d.glob("pm1_iter_*.mat"))
    assert snaps == [f"pm1_iter_{k:02d}.mat" for k in range(1, n + 1)]
    last = sio.loadmat(d / f"pm1_iter_{n:02d}.mat")
    for key in ("K_ps", "C_ps", "coh_ps", "N_opt", "coh_ps_save", "ph_res", "ph_patch", "ph_weight", "ph_grid",
                "Nr", "gamma_change_save", "i_loop"):
        assert np.array_equal(last[key], final[key], equal_nan=True), key
    prep = prepare_stage2_inputs(read_patch(d), nr_rows=20_000)
    one = run_stage2_hip(dataclasses.replace(prep, gamma_max_iterations=1), form="fp32")
    first = sio.loadmat(d / "pm1_iter_01.mat")
    assert float(first["i_loop"][0, 0]) == 1.0
    for key in ("K_ps", "coh_ps", "coh_ps_save"):
        assert np.array_equal(first[key][:, 0], one[key], equal_nan=True), key
    assert np.array_equal(first["ph_weight"], one["ph_weight"])
    assert float(first["gamma_change_save"][0, 0]) == one["gamma_change_save"]
