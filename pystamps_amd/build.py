"""Build recipe of the stage-2 HIP library (gfx950) into pystamps_amd/.

    python -m pystamps_amd.build        # or pystamps_amd.build.build()

Produces pystamps_amd/libpystamps_hip.so from pystamps_amd/csrc/.  Flags:
  -ffp-contract=off     products and sums round separately, as the reference
                        (kernels that want FMAs call fma()/fmaf() explicitly)
  -packed-fp32-ops off  no packed FP32 VALU instructions
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libpystamps_hip.so"
ARCH = os.environ.get("PS2_OFFLOAD_ARCH", "gfx950")

SOURCES = ["ps2_stage2.hip", "ps2_nr.hip", "ps2_registry.hip", "ps2_host_util.cpp", "ps2_prep.cpp"]
HEADERS = ["ps2_device.hpp", "ps2_kernels.hpp", "ps2_host_util.hpp", "ps2_staging.hpp"]


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build the stage-2 library)")


def command(out: Path = LIB, defines: tuple[str, ...] = ()) -> list[str]:
    return [
        hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
        "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall", "-Wno-unused-result",
        "-I", str(PKG.parent / "include"),
        *[f"-D{d}" for d in defines],
        "-o", str(out),
        *[str(CSRC / s) for s in SOURCES],
    ]


def build(verbose: bool = False, out: Path = LIB, defines: tuple[str, ...] = ()) -> Path:
    tmp = out.with_suffix(".so.tmp")
    cmd = command(tmp, defines)
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libpystamps_hip.so")
    os.replace(tmp, out)
    return out


# Stand-alone measurement programs behind two design choices:
#   ps2_probe_sync: a kernel boundary in a captured graph against a
#     device-wide barrier inside one persistent launch;
#   ps2_probe_dft: the CLAP window transform as FP32 MFMA DFT-matrix products
#     against the radix-2 FFT in registers.
PROBES = ["ps2_probe_sync", "ps2_probe_dft"]


def build_probe(verbose: bool = False) -> list[Path]:
    out = []
    for name in PROBES:
        exe = PKG / name
        cmd = [hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-ffp-contract=off",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall",
               "-I", str(CSRC), "-o", str(exe), str(CSRC / f"{name}.hip")]
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError(f"hipcc failed building {name}")
        out.append(exe)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_probe(verbose=True))
