"""The tile autotuner (ps2_run_opts.autotune / autotune_cache).  GPU.

A first session of a shape measures its tiles and appends them to the cache;
a second session of the same shape reads them.  Every tile choice gives the
same loop: the fill's cell tiles and the scan's unit-step source bit for bit
(the same sums in the same order), the CLAP segmentations to f32 rounding
(the window sums are grouped differently) with the same trial indices.
"""
import numpy as np
import pytest

from conftest import small_prep

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S():
    from pystamps_amd import stage2
    return stage2


@pytest.fixture(scope="module")
def prep():
    # 80,000 rows (the scan may compute its unit steps), 13 trials, 178 x 178 cells
    p = small_prep(n_ps=80_000, n_ifg=20, extent=9_000.0, seed=0x5EED1A00)
    assert p.n_ps >= 65_536 and len(p.trial_values) < 16 and not p.row_invariant_bperp
    return p


def _line(path):
    lines = [ln for ln in open(path).read().splitlines() if ln.strip()]
    assert len(lines) == 1, lines
    return lines[0]


def test_measured_then_cached(S, prep, tmp_path):
    cache = tmp_path / "tiles.txt"
    with S.Stage2Session(prep, form="fp32", autotune_cache=str(cache)) as s:
        t1 = s.tiles()
        s.run()
        a = s.download()
    assert t1["source"] == "measured"
    line = _line(cache)
    assert f"I={prep.n_ifg} T={len(prep.trial_values)}" in line
    assert f"fill_cells={t1['fill_cells']}" in line
    with S.Stage2Session(prep, form="fp32", autotune_cache=str(cache)) as s:
        t2 = s.tiles()
        s.run()
        b = s.download()
    assert t2["source"] == "cache"
    assert {k: v for k, v in t2.items() if k != "source"} == {k: v for k, v in t1.items() if k != "source"}
    assert _line(cache) == line                       # a hit appends nothing
    for key in ("K_ps", "coh_ps", "trial_ix", "ph_weight", "ph_grid"):
        assert np.array_equal(a[key], b[key], equal_nan=True), key
    with S.Stage2Session(prep, form="fp32") as s:   # no autotune: the model's tiles
        assert s.tiles()["source"] == "model"


def test_results_across_tile_choices(S, prep, tmp_path):
    cache = tmp_path / "probe.txt"
    with S.Stage2Session(prep, form="fp32", autotune_cache=str(cache)) as s:
        key = _line(cache).split(" fill_cells=")[0]
    runs = {}
    for fill in (64, 128):
        for table in (0, 1):
            for nseg in (1, 3, 7):
                path = tmp_path / f"t_{fill}_{table}_{nseg}.txt"
                path.write_text(f"{key} fill_cells={fill} scan_step_table={table} clap_segments={nseg}\n")
                with S.Stage2Session(prep, form="fp32", autotune_cache=str(path)) as s:
                    t = s.tiles()
                    assert t["source"] == "cache"
                    assert (t["fill_cells"], t["scan_step_table"], t["clap_segments"]) == (fill, bool(table), nseg)
                    s.run()
                    runs[(fill, table, nseg)] = s.download()
    for nseg in (1, 3, 7):
        base = runs[(128, 1, nseg)]
        for fill in (64, 128):
            for table in (0, 1):
                o = runs[(fill, table, nseg)]
                for key in ("K_ps", "C_ps", "coh_ps", "trial_ix", "ph_grid", "ph_weight", "ph_patch", "ph_res"):
                    assert np.array_equal(o[key], base[key], equal_nan=True), (fill, table, nseg, key)
    ref = runs[(128, 1, 1)]
    for nseg in (3, 7):
        o = runs[(128, 1, nseg)]
        assert o["i_loop"] == ref["i_loop"]
        assert np.array_equal(o["trial_ix"], ref["trial_ix"])
        d = np.abs(o["coh_ps"] - ref["coh_ps"])
        print("segmentation", nseg, "vs 1: coh max diff", d.max())
        assert d.max() < 1e-6


def test_unwritable_cache_is_an_error(S, prep, tmp_path):
    from pystamps_amd.prepare import Stage2Error
    with pytest.raises(Stage2Error, match="autotune cache"):
        S.Stage2Session(prep, form="fp32", autotune_cache=str(tmp_path / "no_such_dir" / "tiles.txt"))
