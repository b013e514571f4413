"""pystamps_amd — pySTAMPS stage-2 PS gamma estimation on AMD MI355X (gfx950).

The gamma loop runs in the HIP library ``libpystamps_hip.so`` (C ABI:
include/pystamps_hip.h); this package holds its Python host side: the prep
mirror of the reference's ``prepare_stage2_inputs``, the ctypes binding, the
run/session API and the synthetic workload generator.
"""
from .prepare import (RawPatch, Stage2Error, Stage2Options, Stage2Parms, Stage2Prepared,
                      prepare_stage2_inputs)
from .stage2 import Stage2Session, run_stage2_hip, run_stage2_patch

__all__ = [
    "RawPatch", "Stage2Error", "Stage2Options", "Stage2Parms", "Stage2Prepared",
    "prepare_stage2_inputs", "Stage2Session", "run_stage2_hip", "run_stage2_patch",
]
