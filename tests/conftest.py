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
