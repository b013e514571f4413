"""Seeded synthetic PS patches for the stage-2 workloads (own generator).

Each patch: PS candidates uniform in a square extent; single-master stack with
per-interferogram perpendicular baselines uniform in [-B, B]; per-pixel
bperp = bperp_ifg * (1 + 0.02 * normalised range position); DEM error K
uniform in [-E, E] m mapped to phase by 4 pi bperp K / (lambda R sin(theta))
with lambda = 0.0555 m, R = 850 km, theta = 35 deg; a per-interferogram
atmosphere drawn as a Gaussian random field (2 km correlation length, 1 rad
std) on a 100 m lattice and bilinearly sampled; a coherent fraction with
N(0, sigma) phase noise and the rest uniform phase.  A master column (index 0,
bperp 0, phase 0) is prepended so the stage-2 interferogram count after
dropping the master equals the configured one.
"""
from __future__ import annotations

import math
import os
import threading
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, replace

import numpy as np

from .prepare import RawPatch

LAMBDA_M = 0.0555
SLANT_RANGE_M = 850_000.0
INCIDENCE_DEG = 35.0


@dataclass(frozen=True)
class SynthConfig:
    name: str
    n_ps: int
    n_ifg: int
    extent_m: float
    bperp_max: float
    dem_err_max: float
    coherent_frac: float
    noise_std: float = 0.3
    max_topo_err: float = 20.0
    filter_grid_size: float = 50.0
    gamma_max_iterations: int = 3
    gamma_change_convergence: float = 0.005
    seed: int = 0x5EED0000
    atmo_corr_m: float = 2000.0
    atmo_std: float = 1.0


# The five stage-2 workloads.  config 3 is a list of 16 patches (see c3_patches).
CONFIGS = {
    1: SynthConfig("plain-cpu-20k", 20_000, 30, 5_000.0, 300.0, 20.0, 0.30, seed=0x5EED0001),
    2: SynthConfig("plain-vs-fp32-500k", 500_000, 60, 15_000.0, 250.0, 20.0, 0.40, seed=0x5EED0002),
    4: SynthConfig("large-4m-120ifg", 4_000_000, 120, 20_000.0, 150.0, 20.0, 0.25,
                   gamma_max_iterations=5, seed=0x5EED0004),
    5: SynthConfig("wide-trial-1m-200ifg", 1_000_000, 200, 10_000.0, 600.0, 60.0, 0.20,
                   noise_std=0.5, max_topo_err=60.0, filter_grid_size=10.0, seed=0x5EED0005),
}


def c3_patches() -> list[SynthConfig]:
    """16 patches, candidate counts log-uniform in [50k, 1.2M], ~2000 PS/km^2."""
    rng = np.random.default_rng(0x5EED0300)
    counts = np.exp(rng.uniform(math.log(50_000), math.log(1_200_000), 16)).astype(np.int64)
    out = []
    for p, n in enumerate(counts):
        side_km = math.sqrt(float(n) / 2000.0)
        out.append(SynthConfig(f"batched-{p:02d}", int(n), 60, side_km * 1000.0, 250.0, 20.0, 0.40,
                               seed=0x5EED0300 + p))
    return out


def scaled(cfg: SynthConfig, n_ps: int, n_ifg: int | None = None) -> SynthConfig:
    """Same generator at a smaller candidate count, keeping the PS density."""
    f = math.sqrt(n_ps / cfg.n_ps)
    return replace(cfg, name=f"{cfg.name}-scaled-{n_ps}", n_ps=n_ps,
                   n_ifg=cfg.n_ifg if n_ifg is None else n_ifg, extent_m=cfg.extent_m * f)


def _atmosphere(rng, extent_m, corr_m, std, spacing=100.0):
    n = int(math.ceil(extent_m / spacing)) + 3
    m = 1 << max(4, int(math.ceil(math.log2(n + int(4 * corr_m / spacing)))))
    white = rng.standard_normal((m, m))
    f = np.fft.fftfreq(m, d=spacing)
    s = corr_m / math.sqrt(2.0)  # smoothing std whose self-convolution gives exp(-r^2/2L^2)
    kern = np.exp(-2.0 * math.pi ** 2 * s * s * (f[:, None] ** 2 + f[None, :] ** 2))
    field = np.fft.ifft2(np.fft.fft2(white) * kern).real[:n, :n]
    field -= field.mean()
    sd = field.std()
    return field * (std / sd if sd > 0 else 0.0)


def _bilinear_add(out, cells, n_side, x, y, tmp, spacing=100.0):
    """out[p, f] += bilinear sample of interferogram f's field at (x_p, y_p)
    for every f at once; cells [n*n][I] holds the fields lattice-point-major,
    so each corner is one contiguous row gather per pixel."""
    gx = x / spacing
    gy = y / spacing
    i0 = np.floor(gy).astype(np.int64)
    j0 = np.floor(gx).astype(np.int64)
    fy = gy - i0
    fx = gx - j0
    q = i0 * n_side + j0
    for w, off in (((1 - fy) * (1 - fx), 0), ((1 - fy) * fx, 1), (fy * (1 - fx), n_side),
                   (fy * fx, n_side + 1)):
        np.take(cells, q + off, axis=0, out=tmp)
        tmp *= w[:, None]
        out += tmp


def _bilinear(field, x, y, spacing=100.0):
    gx = x / spacing
    gy = y / spacing
    i0 = np.floor(gy).astype(np.int64)
    j0 = np.floor(gx).astype(np.int64)
    fy = gy - i0
    fx = gx - j0
    return ((1 - fy) * ((1 - fx) * field[i0, j0] + fx * field[i0, j0 + 1])
            + fy * ((1 - fx) * field[i0 + 1, j0] + fx * field[i0 + 1, j0 + 1]))


GEN_CHUNK = 32_768   # rows per noise stream (fixed: the data does not depend on the thread count)


def _threads() -> int:
    n = len(os.sched_getaffinity(0))
    share = os.environ.get("OMP_NUM_THREADS", "")
    if share.isdigit() and int(share) > 0:
        n = min(n, int(share))
    return max(1, min(n, 16))


def generate(cfg: SynthConfig) -> RawPatch:
    """Raw stage-2 inputs for one patch (ph1/ps1/bp1/parms content).  Row
    chunks of GEN_CHUNK pixels draw their phase noise from their own child
    stream of the seed and are built on host threads."""
    rng = np.random.default_rng(cfg.seed)
    P, I = cfg.n_ps, cfg.n_ifg
    x = rng.uniform(0.0, cfg.extent_m, P)
    y = rng.uniform(0.0, cfg.extent_m, P)
    bperp_ifg = rng.uniform(-cfg.bperp_max, cfg.bperp_max, I)
    k_dem = rng.uniform(-cfg.dem_err_max, cfg.dem_err_max, P)
    coherent = rng.random(P) < cfg.coherent_frac
    atmo = [_atmosphere(rng, cfg.extent_m, cfg.atmo_corr_m, cfg.atmo_std) for _ in range(I)]
    span = max(x.max() - x.min(), 1e-9)
    xn = (x - x.min()) / span
    scale = 4.0 * math.pi / (LAMBDA_M * SLANT_RANGE_M * math.sin(math.radians(INCIDENCE_DEG)))

    # the master column (index 0) is written by the chunks too: first touches
    # of these large arrays then happen on every thread, not on one
    ph = np.empty((P, I + 1), dtype=np.complex64)
    bperp_mat = np.empty((P, I + 1), dtype=np.float64)
    n_side = atmo[0].shape[1]
    cells = np.ascontiguousarray(np.stack(atmo).reshape(I, -1).T)   # [n*n][I]
    n_chunks = (P + GEN_CHUNK - 1) // GEN_CHUNK
    streams = np.random.SeedSequence(cfg.seed).spawn(n_chunks)
    local = threading.local()

    def buffers(m):
        # per-thread work arrays, reused by every chunk of the thread
        if getattr(local, "m", 0) < m:
            local.m = m
            local.phase = np.empty((m, I))
            local.tmp = np.empty((m, I))
        return local.phase[:m], local.tmp[:m]

    def chunk(c):
        a, b = c * GEN_CHUNK, min(P, (c + 1) * GEN_CHUNK)
        crng = np.random.default_rng(streams[c])
        phase, tmp = buffers(b - a)
        bp = bperp_ifg[None, :] * (1.0 + 0.02 * xn[a:b, None])
        np.multiply(scale * bp, k_dem[a:b, None], out=phase)
        _bilinear_add(phase, cells, n_side, x[a:b], y[a:b], tmp)
        coh = coherent[a:b]
        crng.random(out=tmp)                              # noise: uniform [-pi, pi),
        tmp *= 2.0 * math.pi                              # N(0, sigma) where coherent
        tmp -= math.pi
        nc = int(coh.sum())
        if nc:
            tmp[coh] = crng.normal(0.0, cfg.noise_std, (nc, I))
        phase += tmp
        ph[a:b, 0] = 1.0
        bperp_mat[a:b, 0] = 0.0
        out = ph[a:b, 1:]
        np.cos(phase, out=tmp)
        out.real = tmp
        np.sin(phase, out=tmp)
        out.imag = tmp
        bperp_mat[a:b, 1:] = bp

    with ThreadPoolExecutor(max_workers=_threads()) as ex:
        list(ex.map(chunk, range(n_chunks)))
    xy = np.empty((P, 3), dtype=np.float32)
    xy[:, 0] = np.arange(1, P + 1, dtype=np.float32)
    xy[:, 1] = x.astype(np.float32)
    xy[:, 2] = y.astype(np.float32)
    parms = {
        "max_topo_err": cfg.max_topo_err,
        "filter_grid_size": cfg.filter_grid_size,
        "clap_win": 32.0,
        "clap_alpha": 1.0,
        "clap_beta": 0.3,
        "clap_low_pass_wavelength": 800.0,
        "lambda": LAMBDA_M,
        "gamma_change_convergence": cfg.gamma_change_convergence,
        "gamma_max_iterations": float(cfg.gamma_max_iterations),
        "small_baseline_flag": "n",
        "filter_weighting": "P-square",
    }
    return RawPatch(
        ph=ph,
        bperp=np.concatenate([[0.0], bperp_ifg]),
        xy=xy,
        master_ix=1,
        bperp_mat=bperp_mat,
        d_a=None,
        mean_incidence=math.radians(INCIDENCE_DEG),
        parms=parms,
    )
