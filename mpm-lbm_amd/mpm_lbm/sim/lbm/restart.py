"""LBM restart files in the reference's format (lbm/restart.py:1-147).

A restart is a compressed NPZ with the arrays `RESTART_ARRAY_KEYS` in the
reference's shapes - rho (n,n,n), v (n,n,n,3), f and F (n,n,n,19), solid and
static_solid (n,n,n) int8 - plus the metadata dict as a JSON string under
`metadata_json`.  Loading checks that metadata against the run's config
(n_grid, boundary-condition mode, target velocity) and the array shapes,
then writes the arrays through the solver's field views, so the device state
and every derived buffer (cell flags, lazy moments) follow.

On this solver `f` and `F` are the same view: the fused substep keeps only
the post-stream populations F, from which every step recomputes the
post-collision f, so the `f` of a restart is written as F and read back into
F.  A reference restart loads its F (written last) as the state.
"""
from __future__ import annotations

import json
from pathlib import Path
from typing import Any

import numpy as np

RESTART_ARRAY_KEYS = ("rho", "v", "f", "F", "solid", "static_solid")
RESTART_METADATA_KEY = "metadata_json"


def write_lbm_restart_npz(path: str | Path, arrays: dict, metadata: dict) -> dict:
    """restart.py:14-23: every array key must be present; returns the payload stats."""
    path = Path(path)
    missing = [k for k in RESTART_ARRAY_KEYS if k not in arrays]
    if missing:
        raise ValueError("missing LBM restart arrays: " + ", ".join(missing))
    path.parent.mkdir(parents=True, exist_ok=True)
    payload = {k: np.asarray(arrays[k]) for k in RESTART_ARRAY_KEYS}
    payload[RESTART_METADATA_KEY] = np.asarray(json.dumps(metadata, sort_keys=True))
    np.savez_compressed(path, **payload)
    return restart_payload_stats(payload, metadata)


def read_lbm_restart_npz(path: str | Path) -> tuple[dict, dict]:
    """restart.py:26-33 (no pickled objects are loaded)."""
    path = Path(path)
    if not path.is_file():
        raise FileNotFoundError(f"LBM restart not found: {path}")
    with np.load(path, allow_pickle=False) as data:
        arrays = {k: np.asarray(data[k]) for k in RESTART_ARRAY_KEYS}
        meta = str(np.asarray(data[RESTART_METADATA_KEY]).item())
    return arrays, json.loads(meta)


def validate_lbm_restart_metadata(metadata: dict, expected: dict) -> dict:
    """restart.py:36-46: which of the guarded metadata entries match."""
    def floats(values):
        return [float(x) for x in values]

    return {
        "restart_lbm_boundary_condition_mode_matches": bool(
            str(metadata.get("lbm_boundary_condition_mode", "")) == str(expected.get("lbm_boundary_condition_mode", ""))),
        "restart_n_grid_matches": bool(int(metadata.get("n_grid", -1)) == int(expected.get("n_grid", -2))),
        "restart_target_u_lbm_matches": bool(
            floats(metadata.get("target_u_lbm", [])) == floats(expected.get("target_u_lbm", []))),
    }


def load_lbm_restart_to_lbm(lbm, path: str | Path, expected_metadata: dict, required: bool = False) -> dict:
    """restart.py:49-76.  A missing file is an error only when `required`."""
    try:
        arrays, metadata = read_lbm_restart_npz(path)
    except FileNotFoundError:
        if required:
            raise
        return {"restart_loaded": False, "restart_missing": True}
    guard = validate_lbm_restart_metadata(metadata, expected_metadata)
    if not all(guard.values()):
        raise ValueError(f"LBM restart metadata mismatch: {guard}")
    _validate_restart_shapes(arrays, int(expected_metadata["n_grid"]))
    # populations first (they mark the moments stale), then the moments
    # (fresh), then the flags (the cell words are rebuilt before the next step)
    lbm.f.from_numpy(arrays["f"].astype(np.float32, copy=False))
    lbm.F.from_numpy(arrays["F"].astype(np.float32, copy=False))
    lbm.rho.from_numpy(arrays["rho"].astype(np.float32, copy=False))
    lbm.v.from_numpy(arrays["v"].astype(np.float32, copy=False))
    lbm.solid.from_numpy(arrays["solid"].astype(np.int8, copy=False))
    lbm.static_solid.from_numpy(arrays["static_solid"].astype(np.int8, copy=False))
    report = restart_payload_stats(arrays, metadata)
    report.update(guard)
    report.update({"restart_loaded": True, "restart_missing": False, "restart_path": str(path)})
    return report


def save_lbm_restart_from_lbm(lbm, path: str | Path, metadata: dict) -> dict:
    """restart.py:79-88."""
    arrays = {"rho": lbm.rho.to_numpy(), "v": lbm.v.to_numpy(), "f": lbm.f.to_numpy(), "F": lbm.F.to_numpy(),
              "solid": lbm.solid.to_numpy(), "static_solid": lbm.static_solid.to_numpy()}
    return write_lbm_restart_npz(path, arrays, metadata)


def expected_restart_metadata_from_config(config) -> dict:
    """restart.py:91-97."""
    return {
        "lbm_boundary_condition_mode": config.lbm_boundary_condition_mode,
        "lbm_restart_scope": config.lbm_restart_scope,
        "n_grid": int(config.n_grid),
        "target_u_lbm": [float(x) for x in config.target_u_lbm],
    }


def restart_payload_stats(arrays: dict, metadata: dict) -> dict:
    """restart.py:100-129: fluid rho range and largest |v|, NaN/Inf flags."""
    rho = np.asarray(arrays["rho"], dtype=np.float64)
    v = np.asarray(arrays["v"], dtype=np.float64)
    fluid = np.asarray(arrays["solid"]) == 0
    if np.any(fluid):
        rho_min, rho_max = float(rho[fluid].min()), float(rho[fluid].max())
        max_v = float(np.linalg.norm(v[fluid], axis=1).max())
    else:
        rho_min = rho_max = max_v = 0.0
    numeric = [np.asarray(arrays[k]) for k in RESTART_ARRAY_KEYS]
    floating = [a for a in numeric if np.issubdtype(a.dtype, np.floating)]
    return {
        "array_keys": list(RESTART_ARRAY_KEYS),
        "has_inf": bool(any(np.isinf(a).any() for a in floating)),
        "has_nan": bool(any(np.isnan(a).any() for a in floating)),
        "metadata": dict(metadata),
        "rho_max": rho_max,
        "rho_min": rho_min,
        "velocity_max_norm": max_v,
    }


def _validate_restart_shapes(arrays: dict, n_grid: int) -> None:
    """restart.py:132-143."""
    n = n_grid
    expected = {"F": (n, n, n, 19), "f": (n, n, n, 19), "rho": (n, n, n), "solid": (n, n, n),
                "static_solid": (n, n, n), "v": (n, n, n, 3)}
    bad = {k: arrays[k].shape for k, shape in expected.items() if arrays[k].shape != shape}
    if bad:
        raise ValueError(f"LBM restart array shape mismatch: {bad}")
