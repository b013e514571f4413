"""MAT v5 patch I/O for stage 2.

``read_patch`` loads a ``PATCH_*`` directory's stage-2 inputs into a
``RawPatch`` with the reference's files, variables and shape rules
(prepare_stage2_inputs / load_bperp_mat / load_da / the parameter and
incidence loaders, crates/pystamps-core/src/native_stage2.rs:415-627,
2058-2153); ``write_pm1`` writes ``pm1.mat`` with write_pm1's variables,
classes and shapes (:1899-1965); ``run_stage2_patch_dir`` is the drop-in for
``run_stage2_native(patch_dir)`` (:201-206, loop :225-413) on the HIP backend;
``write_patch`` writes a synthetic ``PATCH_*`` directory in the same format.

MAT v5 files go through scipy.io.  MAT v7.3 (HDF5) files need h5py, which is
not installed here: reading one raises ``Stage2Error`` naming the file.
"""
from __future__ import annotations

import os
from pathlib import Path
from typing import Optional

import numpy as np

from .prepare import (RawPatch, Stage2Error, Stage2Prepared, load_stage2_parms,
                      prepare_stage2_inputs)

try:  # scipy is part of this environment; keep the import error readable
    import scipy.io as _sio
except ImportError as _e:  # pragma: no cover
    _sio = None
    _SIO_ERROR = _e


def _loadmat(path: Path, label: str) -> dict:
    if _sio is None:  # pragma: no cover
        raise Stage2Error(f"unable to read {label}: scipy.io is unavailable ({_SIO_ERROR})")
    try:
        return _sio.loadmat(str(path), squeeze_me=False, mat_dtype=False)
    except NotImplementedError as e:  # v7.3
        raise Stage2Error(f"unable to read {label}: MAT v7.3 (HDF5) needs h5py, which is not "
                          f"installed ({e})") from e
    except Exception as e:  # noqa: BLE001 - any reader failure is a stage-2 input error
        raise Stage2Error(f"unable to read {label}: {e}") from e


def _vars(d: dict) -> dict:
    return {k: v for k, v in d.items() if not k.startswith("__")}


def _scalar(d: dict, name: str, default: Optional[float]) -> Optional[float]:
    """First element of a numeric variable (scalar_from_mat, :2260-2268)."""
    v = d.get(name)
    if v is None:
        return default
    a = np.asarray(v)
    if a.size == 0 or a.dtype.kind not in "biuf":
        return default
    return float(a.ravel(order="F")[0])


def _vector(d: dict, name: str) -> Optional[np.ndarray]:
    v = d.get(name)
    if v is None:
        return None
    a = np.asarray(v)
    if a.dtype.kind not in "biuf":
        return None
    return a.astype(np.float64).ravel(order="F")


def _text(d: dict, name: str) -> str:
    v = d.get(name)
    if v is None:
        return ""
    a = np.asarray(v)
    if a.dtype.kind in "US":
        return "".join(str(x) for x in a.ravel()).strip()
    return ""


def read_parms(patch_dir: Path) -> dict:
    """parms.mat from the patch directory or its parent (resolve_file_optional)
    as a plain dict of the scalars and texts stage 2 reads (:2058-2128)."""
    for cand in (patch_dir / "parms.mat", patch_dir.parent / "parms.mat"):
        if cand.exists():
            d = _vars(_loadmat(cand, "parms.mat"))
            out = {}
            for k, v in d.items():
                t = _text(d, k)
                if t:
                    out[k] = t
                else:
                    s = _scalar(d, k, None)
                    if s is not None:
                        out[k] = s
            return out
    return {}


def read_patch(patch_dir) -> RawPatch:
    """The stage-2 inputs of a PATCH_* directory (ps1, ph1, bp1, da1, inc1/la1,
    parms)."""
    patch_dir = Path(patch_dir)
    ps = _vars(_loadmat(patch_dir / "ps1.mat", "ps1.mat"))
    ph = _vars(_loadmat(patch_dir / "ph1.mat", "ph1.mat"))
    n_ps = _scalar(ps, "n_ps", 0.0)
    n_ps = int(np.floor(n_ps + 0.5)) if n_ps is not None else 0
    if n_ps <= 0:
        raise Stage2Error("ps1.mat missing valid n_ps")
    if "ph" not in ph:
        raise Stage2Error("ph1.ph is missing or invalid: variable not found")
    phm = np.asarray(ph["ph"])
    if phm.ndim != 2:
        raise Stage2Error(f"ph1.ph is missing or invalid: expected a matrix, got {phm.ndim} dims")
    if phm.shape[0] != n_ps:
        if phm.shape[1] == n_ps:
            phm = phm.T
        else:
            raise Stage2Error(f"ph1.ph has incompatible shape {phm.shape[0]}x{phm.shape[1]} "
                              f"for n_ps={n_ps}")
    phm = np.ascontiguousarray(phm.astype(np.complex64))
    bperp = _vector(ps, "bperp")
    if bperp is None:
        raise Stage2Error("ps1.bperp is missing")
    if "xy" not in ps:
        raise Stage2Error("ps1.xy is missing or invalid: variable not found")
    xy = np.asarray(ps["xy"], dtype=np.float32)
    master_ix = _scalar(ps, "master_ix", 1.0)
    raw = RawPatch(ph=phm, bperp=bperp, xy=xy, master_ix=int(np.floor(master_ix + 0.5)))
    mi = _vector(ps, "mean_incidence")
    if mi is not None and mi.size:
        raw.mean_incidence = float(mi[0])
    if "ifgday_ix" in ps:
        raw.ifgday_ix = np.asarray(ps["ifgday_ix"], dtype=np.float64)
    bp_path = patch_dir / "bp1.mat"
    if bp_path.exists():
        bp = _vars(_loadmat(bp_path, "bp1.mat"))
        if "bperp_mat" not in bp or np.asarray(bp["bperp_mat"]).dtype.kind not in "biuf":
            raise Stage2Error("bp1.bperp_mat is missing or invalid: variable not found")
        raw.bperp_mat = np.asarray(bp["bperp_mat"], dtype=np.float64)
    da_path = patch_dir / "da1.mat"
    if da_path.exists():
        da = _vars(_loadmat(da_path, "da1.mat"))
        v = _vector(da, "D_A")
        raw.d_a = v if (v is not None and v.size == n_ps) else None
    for fname, var, attr in (("inc1.mat", "inc", "inc"), ("la1.mat", "la", "la")):
        path = patch_dir / fname
        if path.exists():
            try:
                d = _vars(_loadmat(path, fname))
            except Stage2Error:
                continue  # the reference skips unreadable incidence files
            v = _vector(d, var)
            if v is not None:
                setattr(raw, attr, v)
    raw.parms = read_parms(patch_dir)
    return raw


def write_pm1(patch_dir, prep: Stage2Prepared, pm1: dict, filename: str = "pm1.mat") -> Path:
    """pm1.mat with write_pm1's variables (native_stage2.rs:1899-1965):
    column vectors K_ps, C_ps, coh_ps, N_opt, coh_ps_save (double); ph_res
    (single n_ps x n_ifg); ph_patch, ph_weight (complex single n_ps x n_ifg);
    ph_grid (complex single n_i x n_j x n_ifg); grid_ij (single n_ps x 2);
    low_pass (double); row vectors Nr, coh_bins; scalars step_number, grid_size,
    i_loop, Nr_max_nz_ix, gamma_change_save (double) and n_trial_wraps
    (single)."""
    path = Path(patch_dir) / filename
    P, I = prep.n_ps, prep.n_ifg
    col = lambda a: np.asarray(a, dtype=np.float64).reshape(-1, 1)  # noqa: E731
    row = lambda a: np.asarray(a, dtype=np.float64).reshape(1, -1)  # noqa: E731
    grid = np.asarray(pm1["ph_grid"], dtype=np.complex64).reshape(prep.n_i, prep.n_j, I)
    out = {
        "K_ps": col(pm1["K_ps"]),
        "C_ps": col(pm1["C_ps"]),
        "coh_ps": col(pm1["coh_ps"]),
        "N_opt": col(pm1["N_opt"]),
        "ph_res": np.asarray(pm1["ph_res"], dtype=np.float32).reshape(P, I),
        "ph_patch": np.asarray(pm1["ph_patch"], dtype=np.complex64).reshape(P, I),
        "step_number": np.float64(1.0),
        "ph_grid": grid,
        "n_trial_wraps": np.float32(prep.n_trial_wraps),
        "grid_ij": np.asarray(prep.grid_ij, dtype=np.float32).reshape(P, 2),
        "grid_size": np.float64(prep.grid_size),
        "low_pass": np.asarray(prep.low_pass, dtype=np.float64),
        "i_loop": np.float64(pm1["i_loop"]),
        "ph_weight": np.asarray(pm1["ph_weight"], dtype=np.complex64).reshape(P, I),
        "Nr": row(pm1["Nr"]),
        "Nr_max_nz_ix": np.float64(prep.nr_max_nz_ix),
        "coh_bins": row(prep.coh_bins),
        "coh_ps_save": col(pm1["coh_ps_save"]),
        "gamma_change_save": np.float64(pm1["gamma_change_save"]),
    }
    tmp = path.with_name(path.name + ".tmp")
    with open(tmp, "wb") as fh:
        _sio.savemat(fh, out, format="5", do_compression=False, oned_as="column")
    os.replace(tmp, path)
    return path


def run_stage2_patch_dir(patch_dir, form: str = "fp32", device: int = 0,
                         nr_rows: Optional[int] = None) -> str:
    """Drop-in for run_stage2_native(patch_dir): read the patch, prepare (the
    Nr histogram on the GPU), run the gamma loop on the GPU, write pm1.mat, and
    return the reference's summary line with the backend named."""
    import os
    from .stage2 import Stage2Session, run_stage2_hip
    patch_dir = Path(patch_dir)
    raw = read_patch(patch_dir)
    kw = {} if nr_rows is None else {"nr_rows": nr_rows}
    prep = prepare_stage2_inputs(raw, device=device, **kw)
    if os.environ.get("PYSTAMPS_STAGE2_NATIVE_DEBUG_PM") is not None:
        # the state after every iteration as pm1_iter_NN.mat, as the
        # reference's loop writes them (native_stage2.rs:259, :337-404)
        with Stage2Session(prep, form=form, device=device) as s:
            s.run_iterations(lambda it, st: write_pm1(patch_dir, prep, st, f"pm1_iter_{it:02d}.mat"))
            pm1 = s.download()
    else:
        pm1 = run_stage2_hip(prep, form=form, device=device)
    write_pm1(patch_dir, prep, pm1)
    return (f"Stage 2 computed coherence for {prep.n_ps} candidates in {pm1['i_loop']} iterations "
            f"using the hip backend ({form} form)")


def write_patch(patch_dir, raw: RawPatch) -> Path:
    """Writes a RawPatch as a PATCH_* directory (ps1, ph1, bp1, da1, parms) in
    MAT v5, in the layouts read_patch and the reference's loader accept."""
    patch_dir = Path(patch_dir)
    patch_dir.mkdir(parents=True, exist_ok=True)
    n_ps = int(raw.ph.shape[0])
    ps = {
        "n_ps": np.float64(n_ps),
        "master_ix": np.float64(raw.master_ix),
        "bperp": np.asarray(raw.bperp, dtype=np.float64).reshape(-1, 1),
        "xy": np.asarray(raw.xy, dtype=np.float32),
    }
    if raw.mean_incidence is not None:
        ps["mean_incidence"] = np.float64(raw.mean_incidence)
    if raw.ifgday_ix is not None:
        ps["ifgday_ix"] = np.asarray(raw.ifgday_ix, dtype=np.float64)
    _sio.savemat(str(patch_dir / "ps1.mat"), ps, format="5")
    _sio.savemat(str(patch_dir / "ph1.mat"), {"ph": np.asarray(raw.ph, dtype=np.complex64)}, format="5")
    if raw.bperp_mat is not None:
        _sio.savemat(str(patch_dir / "bp1.mat"),
                     {"bperp_mat": np.asarray(raw.bperp_mat, dtype=np.float64)}, format="5")
    if raw.d_a is not None:
        _sio.savemat(str(patch_dir / "da1.mat"),
                     {"D_A": np.asarray(raw.d_a, dtype=np.float64).reshape(-1, 1)}, format="5")
    if raw.inc is not None:
        _sio.savemat(str(patch_dir / "inc1.mat"), {"inc": np.asarray(raw.inc, dtype=np.float64)}, format="5")
    if raw.la is not None:
        _sio.savemat(str(patch_dir / "la1.mat"), {"la": np.asarray(raw.la, dtype=np.float64)}, format="5")
    if raw.parms:
        parms = {}
        for k, v in raw.parms.items():
            parms[k] = v if isinstance(v, str) else np.float64(v)
        _sio.savemat(str(patch_dir / "parms.mat"), parms, format="5")
    return patch_dir
