"""The reference's Step-10 mode matrix (32^3 periodic box, 4096 box particles,
20 LBM steps x 10 MPM substeps, initial solid velocity 0.15625 normalised)
rebuilt from its committed driver_config.json, plus diagnostics-row
computation from oracle fields (fsi_driver.py:597-669, fsi_diagnostics.py)."""
from __future__ import annotations

import csv
import json
import math
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden" / "step10_mode_matrix"
INT_COLUMNS = ("bb_link_count", "active_cell_count", "active_reaction_particle_count", "total_mpm_substeps")
FLOAT_COLUMNS = ("rho_min", "rho_max", "lbm_max_v", "fluid_mean_ux", "projection_zone_fluid_mean_ux",
                 "far_field_fluid_mean_ux", "solid_mean_vx_norm", "mpm_min_J", "mpm_max_speed", "projected_mass",
                 "cell_force_max_norm", "hydro_force_max_norm", "bb_max_correction", "max_grid_reaction_norm")


def golden_rows(mode):
    with open(GOLDEN / mode / "diagnostics_timeseries.csv", newline="") as f:
        return {int(r["step"]): r for r in csv.DictReader(f)}


def config(mode):
    from mpm_lbm.sim.drivers.fsi_config import FSIDriverConfig

    data = json.loads((GOLDEN / mode / "driver_config.json").read_text())
    data.update(write_vtk=False, write_particles=False)
    # Pre-Step104 drivers started the solid at the inlet velocity mapped to
    # normalised units: 0.02 * dx / (10 * mpm_dt) = 0.15625.
    data["initial_solid_velocity_norm"] = (0.15625, 0.0, 0.0)
    return FSIDriverConfig(**data)


def box_cloud(cfg):
    """MPMSolid3D.init_box (solid.py:127-159) in float32."""
    n = cfg.n_particles
    ppa = max(1, math.ceil(n ** (1.0 / 3.0)))
    p = np.arange(n)
    idx = np.stack([p % ppa, (p // ppa) % ppa, (p // (ppa * ppa)) % ppa], axis=1).astype(np.float32)
    local = (idx + np.float32(0.5)) / np.float32(max(1, ppa - 1) + 1)
    lo = np.asarray(cfg.box_min, dtype=np.float32)
    span = np.asarray([b - a for a, b in zip(cfg.box_min, cfg.box_max)], dtype=np.float32)
    vol = (cfg.box_max[0] - cfg.box_min[0]) * (cfg.box_max[1] - cfg.box_min[1]) * (cfg.box_max[2] - cfg.box_min[2]) / n
    vol0 = np.full(n, vol, dtype=np.float32)
    return {"x": (lo + span * local).astype(np.float32), "vol0": vol0, "mass": vol0 * np.float32(1.0)}


MATERIAL = (400.0, 0.2, 1.0)  # MPMConfig defaults (mpm/config.py)


def oracle_row(o, mode):
    f = o.fields()
    fluid = f["solid"] == 0
    v = f["v"].astype(np.float64)
    phi = f["solid_phi"].astype(np.float64)
    act = phi > 1.0e-6
    far = fluid & (phi <= 1.0e-6)
    m = o.mass.astype(np.float64)
    pv = f["pv"].astype(np.float64)
    row = {
        "rho_min": float(f["rho"][fluid].min()),
        "rho_max": float(f["rho"][fluid].max()),
        "lbm_max_v": float(np.linalg.norm(v[fluid], axis=1).max()),
        "fluid_mean_ux": float(v[fluid][:, 0].mean()),
        "projection_zone_fluid_mean_ux": float((phi[act] * v[act][:, 0]).sum() / phi[act].sum()) if act.any() else 0.0,
        "far_field_fluid_mean_ux": float(v[far][:, 0].mean()) if far.any() else 0.0,
        "solid_mean_vx_norm": float((m * pv[:, 0]).sum() / m.sum()),
        "mpm_min_J": float(o.Jp.min()),
        "mpm_max_speed": float(np.linalg.norm(pv, axis=1).max()),
        "projected_mass": float(o.proj_stats.projected_mass),
        "active_cell_count": int(o.proj_stats.active_cell_count),
        "cell_force_max_norm": float(np.linalg.norm(f["cell_force"].reshape(-1, 3), axis=1).max()),
        "hydro_force_max_norm": float(np.linalg.norm(f["hydro_force"].reshape(-1, 3), axis=1).max()),
        "bb_link_count": int(o.mbstats.link_count) if mode == "moving_boundary" else 0,
        "bb_max_correction": float(o.mbstats.max_correction) if mode == "moving_boundary" else 0.0,
        "active_reaction_particle_count": int(o.reaction_stats[0]) if mode == "moving_boundary" else 0,
        "max_grid_reaction_norm": float(o.reaction_stats[2]) if mode != "none" else 0.0,
        "total_mpm_substeps": o.total_mpm_substeps,
        # the subcycle accumulator columns (fsi_driver.py:640-652)
        "mb_subcycle_force_accumulation_mode": ("mean_force_accumulator" if o.mb_sample_count > 0 and
                                                o.config.fsi_exchange_mode == "lbm_subcycled_per_fsi_step"
                                                else "not_active"),
        "mb_subcycle_force_sample_count": int(o.mb_sample_count),
        "mb_subcycle_force_accum_norm_max": float(o.accum_norm_max),
        "mb_subcycle_force_mean_norm_max": float(o.mean_norm_max),
    }
    return row


def compare_to_golden(row, golden, float_rtol):
    problems = []
    for k in INT_COLUMNS:
        if int(float(golden[k])) != int(row[k]):
            problems.append(f"{k}: {row[k]} != golden {golden[k]}")
    for k in FLOAT_COLUMNS:
        g = float(golden[k])
        scale = max(abs(g), 1e-12)
        if abs(row[k] - g) > float_rtol * scale + 1e-9:
            problems.append(f"{k}: {row[k]:.9g} vs golden {g:.9g} (rel {abs(row[k] - g) / scale:.2e})")
    return problems
