import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


@pytest.fixture(scope="session")
def ora():
    from oracle import oracle as O
    O.lib()
    return O


def oracle_nr(bp, tv, cb, n):
    from oracle import oracle as O
    return O.random_coherence_histogram(bp, tv, cb, n_rows=n)


_PREP_CACHE = {}


def small_prep(n_ps=3000, n_ifg=20, extent=None, seed=0x5EED1001, nr_rows=10_000, **kw):
    """Seeded synthetic patch (config-1 generator, reduced) prepared with the
    oracle's Nr so CPU tests need no GPU."""
    from pystamps_amd import synth
    from pystamps_amd.prepare import prepare_stage2_inputs
    key = (n_ps, n_ifg, extent, seed, nr_rows, tuple(sorted(kw.items())))
    if key not in _PREP_CACHE:
        base = synth.CONFIGS[1]
        cfg = synth.scaled(base, n_ps, n_ifg)
        from dataclasses import replace
        cfg = replace(cfg, seed=seed, **kw)
        if extent is not None:
            cfg = replace(cfg, extent_m=extent)
        raw = synth.generate(cfg)
        _PREP_CACHE[key] = prepare_stage2_inputs(raw, nr_fn=oracle_nr, nr_rows=nr_rows)
    return _PREP_CACHE[key]


def gpu_available():
    try:
        from pystamps_amd import _native
        return _native.device_count() > 0
    except Exception:
        return False


def perturb_one_ulp(prep, seed=5):
    """The same patch with every ph_nm component moved by one complex64 ulp
    (up or down at random): the smallest change of the stored inputs.  An FP64
    loop on it shows which rows are input-quantisation ties: rows whose
    discrete choices (trial index, P-square lookup and bin) that change
    alone flips, and the rows those flips move in later iterations."""
    import dataclasses
    rng = np.random.default_rng(seed)
    ph = np.ascontiguousarray(prep.ph_nm, np.complex64)
    up = np.float32(np.inf)
    down = np.float32(-np.inf)
    re = np.where(rng.integers(0, 2, ph.shape).astype(bool), np.nextafter(ph.real, up),
                  np.nextafter(ph.real, down))
    im = np.where(rng.integers(0, 2, ph.shape).astype(bool), np.nextafter(ph.imag, up),
                  np.nextafter(ph.imag, down))
    return dataclasses.replace(prep, ph_nm=(re + 1j * im).astype(np.complex64))


TIE_MARGIN = 1e-6   # coherence margins below this count as input-quantisation ties


def floor_check(tag, g, r, rps, coh_tol=1e-5):
    """FP32 form g against an FP64 loop r, judged by FP64 loops rps (one or
    several) on one-ulp-perturbed inputs: every trial index of g differs from r
    only on a row that r itself ties (margins below TIE_MARGIN, when r reports
    them) or that some perturbation moves, and every coherence of g off by more
    than coh_tol lies on such a row.  A row whose FP64 margin is far below the
    input quantisation flips in about half of the perturbed runs, hence several
    samples where r reports no margins.  Returns the counts."""
    if isinstance(rps, dict):
        rps = [rps]
    assert g["i_loop"] == r["i_loop"], (tag, g["i_loop"], r["i_loop"])
    dg = np.abs(g["coh_ps"] - r["coh_ps"])
    moved = np.zeros(dg.shape, bool)
    n_pert = 0
    for rp in rps:
        dp = np.abs(rp["coh_ps"] - r["coh_ps"])
        moved |= (rp["trial_ix"] != r["trial_ix"]) | (dp > TIE_MARGIN)
        n_pert = max(n_pert, int((dp > coh_tol).sum()))
    if "trial_margin_min" in r:
        moved |= r["trial_margin_min"] < TIE_MARGIN
    if "refine_margin_min" in r:
        moved |= r["refine_margin_min"] < TIE_MARGIN
    mism = g["trial_ix"] != r["trial_ix"]
    counts = {"trial_mismatch": int(mism.sum()), "trial_mismatch_untied": int((mism & ~moved).sum()),
              "coh_above_tol": int((dg > coh_tol).sum()),
              "coh_above_tol_unmoved": int(((dg > coh_tol) & ~moved).sum()),
              "perturbed_coh_above_tol": n_pert, "perturbation_samples": len(rps),
              "coh_max": float(dg.max()), "coh_p999": float(np.quantile(dg, 0.999))}
    print(tag, counts, flush=True)
    assert counts["trial_mismatch_untied"] == 0, (tag, counts)
    assert counts["coh_above_tol_unmoved"] == 0, (tag, counts)
    assert counts["coh_above_tol"] <= 2 * counts["perturbed_coh_above_tol"] + 10, (tag, counts)
    return counts
