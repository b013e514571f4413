"""The `hip` provider of the Python pipeline's kernel registry.

CPU tests pin the oracle's restatement of the Python-path topofit
(pystamps/pipeline/ported.py:2355-2513): the near-max candidate and selection
rules on hand-made coherence vectors, the noise-free on-node fit, the subset
fit of rows with zero elements; the grid accumulation against numpy's
``np.add.at`` (the routine the reference calls); the library's histogram
(host code) against the oracle's golden-pinned rule; and the registry glue.
GPU tests compare the provider's kernels with the oracle.
"""
import json
import math
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"


def _real_row():
    """The reference's real-data 75-interferogram row (input of its
    near-max refinement test, tests/test_stage2_ported.py:376-537)."""
    d = json.loads((GOLDEN / "real_row_75ifg.json").read_text())
    cpx = np.asarray(d["cpx_re"]) + 1j * np.asarray(d["cpx_im"])
    return cpx, np.asarray(d["bperp"], dtype=np.float64), float(d["n_trial_wraps"])


# --- rules of the Python-path topofit ------------------------------------------

def test_near_max_candidate_rule(ora):
    f = ora.py_near_max_candidates
    # interior local maxima within 2e-4 of the max; 0.4999 is not a local max
    assert list(f([0.1, 0.5, 0.4999, 0.2, 0.49995, 0.3])) == [1, 4]
    # ends compare with their one neighbour (>=)
    assert list(f([0.5, 0.3, 0.49985])) == [0, 2]
    # a plateau: both cells are local maxima (>=)
    assert list(f([0.3, 0.5, 0.5, 0.3])) == [1, 2]
    # outside the tolerance
    assert list(f([0.5, 0.2, 0.4997])) == [0]
    assert list(f([0.7])) == [0]


def test_select_candidate_rule(ora):
    g = ora.py_select_candidate
    # one candidate: itself
    assert g([3], [0.4], [0.9], 7) == 3
    # the two endpoints: the coarse argmax even when refinement prefers the other
    assert g([0, 6], [0.5, 0.49995], [0.7, 0.8], 7) == 0
    # otherwise the refined argmax, first on ties
    assert g([1, 4], [0.5, 0.49995], [0.7, 0.8], 7) == 4
    assert g([1, 4, 5], [0.5, 0.49995, 0.4999], [0.8, 0.8, 0.1], 7) == 1
    # endpoints but three candidates: refined
    assert g([0, 3, 6], [0.5, 0.4, 0.49995], [0.1, 0.2, 0.3], 7) == 6


def _on_node_rows(n_rows=40, n_ifg=25, wraps=0.7, seed=1):
    rng = np.random.default_rng(seed)
    tn = math.ceil(8 * wraps)
    bp = rng.uniform(-200, 200, (n_rows, n_ifg))
    t = rng.integers(-tn, tn + 1, n_rows)
    c0 = rng.uniform(-np.pi, np.pi, n_rows)
    rng_b = bp.max(1) - bp.min(1)
    k = (np.pi / 4) / rng_b * t
    cpx = np.exp(1j * (k[:, None] * bp + c0[:, None]))
    return cpx, bp, k, c0, wraps


def test_py_topofit_noise_free_on_node(ora):
    cpx, bp, k, c0, wraps = _on_node_rows()
    K, Cc, coh, res = ora.py_topofit_rows(cpx, bp, wraps)
    np.testing.assert_allclose(K, k, rtol=0, atol=1e-12)
    np.testing.assert_allclose(coh, 1.0, rtol=0, atol=1e-12)
    d = np.angle(np.exp(1j * (Cc - c0)))
    np.testing.assert_allclose(d, 0.0, atol=1e-9)
    np.testing.assert_allclose(np.abs(res), 1.0, atol=1e-12)


def test_py_topofit_zero_elements_fit_the_subset(ora):
    rng = np.random.default_rng(2)
    cpx = np.exp(1j * rng.uniform(-np.pi, np.pi, (30, 20)))
    bp = rng.uniform(-150, 150, (30, 20))
    holes = cpx.copy()
    holes[:, [3, 11, 17]] = 0.0
    K, Cc, coh, res = ora.py_topofit_rows(holes, bp, 0.6)
    keep = [j for j in range(20) if j not in (3, 11, 17)]
    K2, C2, coh2, res2 = ora.py_topofit_rows(cpx[:, keep], bp[:, keep], 0.6)
    np.testing.assert_array_equal(K, K2)
    np.testing.assert_array_equal(coh, coh2)
    assert not np.any(res[:, [3, 11, 17]])
    np.testing.assert_array_equal(res[:, keep], res2)
    Kz, Cz, cz, rz = ora.py_topofit_rows(np.zeros((2, 5)), np.ones((2, 5)), 0.5)
    assert np.all(np.isnan(Kz)) and np.all(np.isnan(cz)) and not np.any(rz)


def test_real_row_takes_the_near_max_refinement_path(ora):
    """The real row has several near-max trials (the reference's test exists
    for that path); the oracle's fit refines from one of them."""
    from pystamps_amd.prepare import trial_values
    cpx, bp, wraps = _real_row()
    t = trial_values(wraps)
    assert list(t) == list(range(-6, 7))
    step = np.pi / 4 / (bp.max() - bp.min())
    coh_t = np.abs(np.exp(-1j * np.outer(t, bp * step)) @ cpx) / np.abs(cpx).sum()
    cand = ora.py_near_max_candidates(coh_t)
    assert len(cand) >= 2
    K, Cc, coh, res = ora.py_topofit_rows(cpx[None, :], bp[None, :], wraps)
    assert 0.0 < coh[0] <= 1.0 and -np.pi <= Cc[0] <= np.pi
    # the refined K stays within one trial step of a candidate's coarse K
    assert min(abs(K[0] - step * t[c]) for c in cand) <= step
    np.testing.assert_allclose(np.abs(res[0]), np.abs(cpx), atol=1e-6)   # a rotation (complex64 out)


def _coh_rule_rows(seed=8, P=400, I=24):
    """Rows against one bperp vector, symmetric about 0; a third of them have
    zero elements at the vector's extremes (the columns that set its range),
    some random holes.  The last eighth are real rows cos(7 s b) (s the trial
    step of the whole vector), whose trial coherence is symmetric in the trial,
    so its peaks at the grid's two ends tie; half of those miss both extremes."""
    rng = np.random.default_rng(seed)
    u = np.sort(rng.uniform(5, 220, I // 2))
    bp = np.concatenate([-u[::-1], np.zeros(I % 2), u])
    k = rng.uniform(-0.008, 0.008, P)
    cpx = np.exp(1j * (k[:, None] * bp + rng.uniform(0.0, 1.2, P)[:, None] * rng.standard_normal((P, I))))
    ext = [int(np.argmin(bp)), int(np.argmax(bp))]
    cpx[: P // 3, ext[0]] = 0.0
    cpx[P // 6: P // 2, ext[1]] = 0.0
    cpx[rng.random((P, I)) < 0.02] = 0.0
    tie = P - P // 8
    step = (np.pi / 4) / (bp.max() - bp.min())
    cpx[tie:] = np.cos(7 * step * bp) * rng.uniform(0.5, 2.0, (P - tie, 1))
    cpx[tie::2, ext] = 0.0
    return cpx, bp, ext


def test_coh_row_invariant_rule_pins(ora):
    """_ps_topofit_batch_row_invariant_coh (ported.py:2555-2616).  Rows whose
    whole-vector trial grid has one near-maximum peak equal the subset fit of
    the same row with a 1e-200 phasor put in every zero element (which makes
    the subset's range the whole vector's and changes no sum); the other rows
    equal the subset fit of the row as it is; rows without zeros equal it
    exactly either way."""
    cpx, bp, ext = _coh_rule_rows()
    coh, single = ora.py_topofit_coh_row_invariant(cpx, bp, 0.7)
    _, _, coh_sub, _ = ora.py_topofit_rows(cpx, bp, 0.7)
    filled = np.where(cpx == 0, 1e-200 * np.exp(1j * 0.3), cpx)
    _, _, coh_fill, _ = ora.py_topofit_rows(filled, bp, 0.7)
    assert single.sum() > 200 and (~single).sum() > 0
    np.testing.assert_allclose(coh[single], coh_fill[single], rtol=0, atol=1e-13)
    np.testing.assert_array_equal(coh[~single], coh_sub[~single])
    whole = np.all(cpx != 0, axis=1)
    np.testing.assert_array_equal(coh[whole], coh_sub[whole])
    # the rule is seen: rows missing an extreme element differ from the subset fit
    gap = np.abs(coh - coh_sub)[single & ~np.all(cpx[:, ext] != 0, axis=1)]
    assert gap.size > 50 and np.count_nonzero(gap > 1e-6) > gap.size // 4
    # all-zero rows: NaN on a grid of several trials, 0 on a single-trial grid
    z = np.zeros((2, 6))
    assert np.all(np.isnan(ora.py_topofit_coh_row_invariant(z, bp[:6], 0.7)[0]))
    c0, s0 = ora.py_topofit_coh_row_invariant(z, bp[:6], 0.0)
    assert np.array_equal(c0, [0.0, 0.0]) and s0.all()


def test_py_grid_accumulate_matches_numpy_add_at(ora):
    rng = np.random.default_rng(3)
    P, I, ni, nj = 700, 9, 6, 7
    ph = (rng.standard_normal((P, I)) + 1j * rng.standard_normal((P, I))).astype(np.complex64)
    gl = rng.integers(0, ni * nj, P)
    ref = np.zeros((ni, nj, I), np.complex64)
    for i in range(I):
        flat = ref[:, :, i].reshape(-1)
        np.add.at(flat, gl, ph[:, i])
    out = ora.py_grid_accumulate(ph, gl, ni, nj)
    assert np.array_equal(out, ref)


# --- library host code and registry glue (no GPU) ---------------------------------------

def test_library_histogram_matches_oracle(ora):
    from pystamps_amd import registry_hip
    from pystamps_amd.prepare import coh_bins
    rng = np.random.default_rng(4)
    v = np.concatenate([rng.uniform(-0.2, 1.2, 5000), [np.nan, np.inf, -np.inf, 0.005, 0.995]])
    for centers in (coh_bins(), np.array([0.0, 0.1, 0.35, 0.4, 1.0]), np.array([0.5])):
        got = registry_hip.stage2_histogram(v, centers)
        ref = ora.hist_with_centers(v[np.isfinite(v)], centers)
        assert np.array_equal(got, ref), centers
    assert registry_hip.stage2_histogram(v, np.array([])).size == 0


def test_register_adds_provider_and_kernels():
    from pystamps_amd import registry_hip

    class FakeRegistry:
        def __init__(self):
            self.providers, self.kernels = {}, {}

        def register_provider(self, backend_id, *, description, aliases=(), availability_probe=None,
                              unavailable_reason=""):
            self.providers[backend_id] = (description, availability_probe, unavailable_reason)

        def register(self, name, cpu=None, gpu=None, native=None, **impls):
            self.kernels.setdefault(name, {}).update(impls)

    reg = registry_hip.register(FakeRegistry())
    assert "hip" in reg.providers and callable(reg.providers["hip"][1])
    assert set(reg.kernels) == set(registry_hip.KERNELS)
    for name in registry_hip.KERNELS:
        assert reg.kernels[name]["hip"] is registry_hip.IMPLEMENTATIONS[name]


# --- GPU -------------------------------------------------------------------------

@pytest.mark.gpu
def test_gpu_topofit_matches_oracle(ora):
    from pystamps_amd import registry_hip
    rng = np.random.default_rng(5)
    P, I = 3000, 33
    bp = rng.uniform(-250, 250, (P, I))
    k = rng.uniform(-0.01, 0.01, P)
    noise = rng.uniform(0.0, 1.5, P)
    cpx = np.exp(1j * (k[:, None] * bp + noise[:, None] * rng.standard_normal((P, I))))
    cpx[rng.random((P, I)) < 0.03] = 0.0                     # missing elements
    cpx[7] = 0.0                                             # an all-zero row
    # two near-equal peaks: +t and -t mixtures on a symmetric baseline
    sym = np.linspace(-200, 200, I)
    bp[11] = sym
    kk = (np.pi / 4) / 400 * 3
    cpx[11] = np.exp(1j * kk * sym) + 0.9998 * np.exp(-1j * kk * sym)
    K, Cc, coh, res = registry_hip.stage2_topofit(cpx, bp, 0.7)
    Ko, Co, co, ro = ora.py_topofit_rows(cpx, bp, 0.7)
    np.testing.assert_allclose(K, Ko, rtol=0, atol=1e-9, equal_nan=True)
    np.testing.assert_allclose(coh, co, rtol=0, atol=1e-9, equal_nan=True)
    dC = np.angle(np.exp(1j * (Cc - Co)))
    assert np.nanmax(np.abs(dC)) < 1e-8
    np.testing.assert_allclose(res, ro.astype(np.complex64), rtol=0, atol=2e-6)
    assert np.isnan(K[7]) and not np.any(res[7])
    # row invariant, and the coherence-only kernel
    Kr, Cr, cr, rr = registry_hip.stage2_topofit_row_invariant(cpx, np.tile(sym, (P, 1)), 0.7)
    Ko2, Co2, co2, ro2 = ora.py_topofit_rows(cpx, sym, 0.7)
    np.testing.assert_allclose(Kr, Ko2, rtol=0, atol=1e-9, equal_nan=True)
    np.testing.assert_allclose(cr, co2, rtol=0, atol=1e-9, equal_nan=True)


@pytest.mark.gpu
def test_gpu_topofit_coh_row_invariant_matches_oracle(ora):
    from pystamps_amd import registry_hip
    cpx, bp, ext = _coh_rule_rows(seed=9, P=3000, I=33)
    cpx[5] = 0.0
    got = registry_hip.stage2_topofit_coh_row_invariant(cpx, bp, 0.7)
    want, single = ora.py_topofit_coh_row_invariant(cpx, bp, 0.7)
    assert single.sum() > 1000 and np.isnan(got[5])
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-9, equal_nan=True)
    # a [n_row][n_ifg] bperp matrix of equal rows is the same call
    got2 = registry_hip.stage2_topofit_coh_row_invariant(cpx, np.tile(bp, (cpx.shape[0], 1)), 0.7)
    np.testing.assert_array_equal(got2, got)
    z = registry_hip.stage2_topofit_coh_row_invariant(np.zeros((3, 6)), bp[:6], 0.0)
    assert np.array_equal(z, np.zeros(3))


@pytest.mark.gpu
def test_gpu_topofit_real_row_matches_oracle(ora):
    from pystamps_amd import registry_hip
    cpx, bp, wraps = _real_row()
    rows = np.tile(cpx, (3, 1))
    bps = np.tile(bp, (3, 1))
    K, Cc, coh, res = registry_hip.stage2_topofit(rows, bps, wraps)
    Ko, Co, co, ro = ora.py_topofit_rows(rows, bps, wraps)
    np.testing.assert_allclose(K, Ko, rtol=0, atol=1e-12)
    np.testing.assert_allclose(coh, co, rtol=0, atol=1e-12)
    assert np.max(np.abs(np.angle(np.exp(1j * (Cc - Co))))) < 1e-10
    Kr, Cr, cr, rr = registry_hip.stage2_topofit_row_invariant(rows, bp, wraps)
    np.testing.assert_array_equal(Kr, K)
    np.testing.assert_array_equal(cr, coh)


@pytest.mark.gpu
def test_gpu_grid_accumulate_matches_add_at(ora):
    from pystamps_amd import registry_hip
    rng = np.random.default_rng(6)
    P, I, ni, nj = 20000, 17, 40, 55
    ph = (rng.standard_normal((P, I)) + 1j * rng.standard_normal((P, I))).astype(np.complex64)
    gl = rng.integers(0, ni * nj, P)
    out = registry_hip.stage2_grid_accumulate(ph, gl, ni, nj)
    assert np.array_equal(out, ora.py_grid_accumulate(ph, gl, ni, nj))
    with pytest.raises(registry_hip.BackendError, match="out of range"):
        registry_hip.stage2_grid_accumulate(ph[:3], np.array([0, 1, ni * nj]), ni, nj)
