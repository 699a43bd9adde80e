"""Checkpoint and resume of a whole FSI run (FSIDriver3D).

The reference saves and loads only the LBM state (lbm/restart.py); its
checkpointed campaign rows add a tmp-file-then-rename write and keep the last
N files (step120's checkpoint contract, tests/test_step120_checkpoint_restart_
contract.py:17-66).  A checkpoint here holds the state a run carries from one
FSI step to the next, so that a run resumed from it continues exactly as the
uninterrupted run would:

* LBM: the reference's restart arrays (rho, v, f, F, solid, static_solid, in
  its shapes) plus old_solid, the projected fields, the cell force and the
  hydrodynamic force of the last step;
* MPM: particle positions, velocities, APIC C, deformation gradient F, Jp,
  mass, rest volume, the fixed-particle mask and pinned positions;
* the driver's step counters, the open-boundary limiter's run counters and
  its diagnostics and monitor rows.

What is not saved is rebuilt or recomputed before it is read: the cell words
(link masks), the MPM grid (cleared every substep), the moving-wall link
buffers and accumulator (reset every FSI step), and the captured step graph.
Lazy device state (moments, the hydro force of the last links) is settled
when the checkpoint is written; settled and pending forms give the same
numbers.

Files are `ckpt_<lbm step, 8 digits>.npz` (uncompressed: a checkpoint is
written mid-run) with a JSON metadata record; no pickled objects are read.
"""
from __future__ import annotations

import json
import os
import tempfile
from pathlib import Path
from typing import Optional

import numpy as np

CHECKPOINT_FORMAT = "fsi_checkpoint_v1"
CHECKPOINT_METADATA_KEY = "metadata_json"
_PREFIX = "ckpt_"

# Config entries that must match between the run that wrote a checkpoint and
# the run that loads it.
_GUARDED = ("coupling_mode", "fsi_exchange_mode", "n_grid", "lattice_shape", "n_particles",
            "lbm_substeps_per_fsi_step", "mpm_substeps_per_lbm_step", "lbm_population_storage",
            "lbm_collision", "lbm_boundary_condition_mode", "lbm_open_boundary_semantics", "target_u_lbm",
            "mpm_dt", "lbm_dt_phys_override_s", "gravity", "dynamic_solid_threshold", "beta_lbm",
            "penalty_force_cap_lbm", "mb_reaction_scale", "mb_force_cap_norm", "reaction_transfer_mode",
            "lbm_viscosity_semantics")

_LBM_FIELDS = ("rho", "v", "F", "solid", "static_solid", "old_solid", "solid_phi", "solid_mass", "solid_vel",
               "cell_force", "hydro_force")
_MPM_FIELDS = ("x", "v", "C", "F", "Jp", "mass", "vol0", "fixed_mask", "fixed_x")


def _resolved(driver) -> dict:
    """Parameters the driver resolved from its config and geometry files (the
    material may come from a geometry file): they must match too."""
    sp, lp = driver.solid._params, driver.lbm._params
    return {"mpm_mu": float(sp.mu), "mpm_lambda": float(sp.la), "mpm_dt": float(sp.dt), "lbm_s_v": float(lp.s_v),
            "lbm_s_other": float(lp.s_other)}


def _guard(config) -> dict:
    d = config.to_dict()
    out = {}
    for k in _GUARDED:
        v = d.get(k)
        out[k] = list(v) if isinstance(v, (list, tuple)) else v
    return out


def checkpoint_path(directory: str | Path, step: int) -> Path:
    return Path(directory) / f"{_PREFIX}{int(step):08d}.npz"


def list_fsi_checkpoints(directory: str | Path) -> list:
    """Checkpoint files in `directory`, oldest first."""
    d = Path(directory)
    if not d.is_dir():
        return []
    return sorted(p for p in d.glob(f"{_PREFIX}*.npz") if p.name[len(_PREFIX):-4].isdigit())


def latest_fsi_checkpoint(directory: str | Path) -> Optional[Path]:
    found = list_fsi_checkpoints(directory)
    return found[-1] if found else None


def save_fsi_checkpoint(driver, directory: str | Path, keep_last: int = 3) -> Path:
    """Write the driver's state after its current FSI step; keep the newest
    `keep_last` checkpoints in `directory` (0 keeps all)."""
    if not driver.initialized:
        raise RuntimeError("checkpoint of a driver that has not been initialized")
    lbm, solid = driver.lbm, driver.solid
    arrays = {f"lbm_{k}": getattr(lbm, k).to_numpy() for k in _LBM_FIELDS}
    arrays["lbm_f"] = arrays["lbm_F"]  # the reference's restart keys (f is F on this solver)
    arrays.update({f"mpm_{k}": getattr(solid, k).to_numpy() for k in _MPM_FIELDS})
    meta = {
        "format": CHECKPOINT_FORMAT,
        "config": _guard(driver.config),
        "resolved": _resolved(driver),
        "current_lbm_step": int(driver.current_lbm_step),
        "total_lbm_substeps": int(driver.total_lbm_substeps),
        "total_mpm_substeps": int(driver.total_mpm_substeps),
        "lbm_cell_force_active": bool(lbm._cell_force_active),
        "lbm_mb_sample_count": int(lbm.mb_sample_count),
        "lbm_open_boundary_run_counters": lbm._ob_counters()[1],
        "diagnostics_rows": driver.diagnostics_rows,
        "flap_tip_monitor_rows": driver.flap_tip_monitor_rows,
        "fluent_like_monitor_rows": driver.fluent_like_monitor_rows,
    }
    # key order kept: the rows' column order is the order of the output CSVs
    arrays[CHECKPOINT_METADATA_KEY] = np.asarray(json.dumps(meta, default=float))
    d = Path(directory)
    d.mkdir(parents=True, exist_ok=True)
    final = checkpoint_path(d, driver.current_lbm_step)
    fd, tmp = tempfile.mkstemp(prefix=".tmp_ckpt_", suffix=".npz", dir=d)
    try:
        with os.fdopen(fd, "wb") as f:
            np.savez(f, **arrays)
            f.flush()
            os.fsync(f.fileno())
        os.replace(tmp, final)  # readers see the old file or the whole new one
    except BaseException:
        if os.path.exists(tmp):
            os.unlink(tmp)
        raise
    if keep_last > 0:
        for old in list_fsi_checkpoints(d)[:-keep_last]:
            old.unlink()
    return final


def load_fsi_checkpoint(driver, path: str | Path) -> dict:
    """Load a checkpoint into `driver` (initialized here if it is not yet);
    the run's config must match the guarded entries of the writer's."""
    path = Path(path)
    if not path.is_file():
        raise FileNotFoundError(f"FSI checkpoint not found: {path}")
    with np.load(path, allow_pickle=False) as data:
        arrays = {k: np.asarray(data[k]) for k in data.files if k != CHECKPOINT_METADATA_KEY}
        meta = json.loads(str(np.asarray(data[CHECKPOINT_METADATA_KEY]).item()))
    if meta.get("format") != CHECKPOINT_FORMAT:
        raise ValueError(f"not an FSI checkpoint of format {CHECKPOINT_FORMAT}: {path}")
    expected = _guard(driver.config)
    mismatch = {k: (meta["config"].get(k), v) for k, v in expected.items() if meta["config"].get(k) != v}
    if mismatch:
        raise ValueError(f"FSI checkpoint config mismatch (checkpoint, run): {mismatch}")
    if not driver.initialized:
        driver.initialize()
    resolved = _resolved(driver)
    mismatch = {k: (meta.get("resolved", {}).get(k), v) for k, v in resolved.items()
                if meta.get("resolved", {}).get(k) != v}
    if mismatch:
        raise ValueError(f"FSI checkpoint material/time-step mismatch (checkpoint, run): {mismatch}")
    lbm, solid = driver.lbm, driver.solid
    # populations first (they mark the moments stale), then the moments
    # (fresh); the flag writes make the cell words rebuild before the next step
    lbm.F.from_numpy(arrays["lbm_F"])
    lbm.rho.from_numpy(arrays["lbm_rho"])
    lbm.v.from_numpy(arrays["lbm_v"])
    for k in ("solid", "static_solid", "old_solid", "solid_phi", "solid_mass", "solid_vel"):
        getattr(lbm, k).from_numpy(arrays[f"lbm_{k}"])
    # always: a driver that already ran penalty steps must not keep its own force
    lbm.cell_force.from_numpy(arrays["lbm_cell_force"])
    lbm._cell_force_active = bool(meta["lbm_cell_force_active"])
    # the hydrodynamic force of the last step (recomputed by the next
    # moving-boundary step; read by diagnostics and reports until then)
    lbm.drop_pending_links()
    lbm.hydro_force.from_numpy(arrays["lbm_hydro_force"])
    lbm.mb_sample_count = int(meta["lbm_mb_sample_count"])
    lbm.set_open_boundary_limiter_run_counters(*meta.get("lbm_open_boundary_run_counters", [0] * 5))
    for k in _MPM_FIELDS:
        getattr(solid, k).from_numpy(arrays[f"mpm_{k}"])
    driver.current_lbm_step = int(meta["current_lbm_step"])
    driver.total_lbm_substeps = int(meta["total_lbm_substeps"])
    driver.total_mpm_substeps = int(meta["total_mpm_substeps"])
    driver.diagnostics_rows = list(meta["diagnostics_rows"])
    driver.flap_tip_monitor_rows = list(meta["flap_tip_monitor_rows"])
    driver.fluent_like_monitor_rows = list(meta["fluent_like_monitor_rows"])
    driver._drop_step_graph()
    return {"checkpoint_loaded": True, "checkpoint_path": str(path), "current_lbm_step": driver.current_lbm_step}
