"""Diagnostics-row comparison shared by the CPU and GPU tests: a driver row
(FSIDriver3D.collect_diagnostics, reduced on the device) against the row the
oracle's fields give (tests/step10.py:oracle_row), column by column.

Each float column is measured against the natural scale of what it reduces
(a mean of signed velocities against the largest speed, not its own value,
which may sit near zero), and held to the parity bar of tests/parity.py:
within `tol` of the FP32 oracle, or of the FP64 oracle, or, for the columns
that reduce the velocity-level particle fields and the forces linear in them,
within the FP32 oracle's own distance from FP64.  Integer and text columns
must be equal."""
from __future__ import annotations

from tests import step10

INT_COLUMNS = ("step", "total_mpm_substeps", "active_cell_count", "bb_link_count", "mb_subcycle_force_sample_count",
               "active_reaction_particle_count")
TEXT_COLUMNS = ("coupling_mode", "mb_subcycle_force_accumulation_mode")
# column -> the oracle-row column whose magnitude is its scale
SCALE_OF = {
    "rho_min": "rho_min", "rho_max": "rho_max",
    "lbm_max_v": "lbm_max_v", "fluid_mean_ux": "lbm_max_v", "projection_zone_fluid_mean_ux": "lbm_max_v",
    "far_field_fluid_mean_ux": "lbm_max_v",
    "solid_mean_vx_norm": "mpm_max_speed", "mpm_max_speed": "mpm_max_speed",
    "mpm_min_J": "mpm_min_J", "projected_mass": "projected_mass",
    "cell_force_max_norm": "cell_force_max_norm", "hydro_force_max_norm": "hydro_force_max_norm",
    "bb_max_correction": "bb_max_correction",
    "mb_subcycle_force_accum_norm_max": "mb_subcycle_force_accum_norm_max",
    "mb_subcycle_force_mean_norm_max": "mb_subcycle_force_mean_norm_max",
    "max_grid_reaction_norm": "max_grid_reaction_norm",
}
FLOAT_COLUMNS = tuple(SCALE_OF)
# Columns held to the FP32 method error when they miss 1e-5: reductions of
# particle velocity and of the forces linear in it (the conditioned fields of
# tests/parity.py), and the lattice velocity columns.  In a box without inflow
# (the Step-10 penalty run) |u| stays near 1e-5 while the populations it is
# summed from are ~0.05: one FP32 rounding of that momentum sum is already
# ~1e-4 of |u|, and the FP32 oracle itself sits up to 5e-3 of |u| from FP64.
CONDITIONED_COLUMNS = ("solid_mean_vx_norm", "mpm_max_speed", "cell_force_max_norm", "hydro_force_max_norm",
                       "mb_subcycle_force_accum_norm_max", "mb_subcycle_force_mean_norm_max",
                       "max_grid_reaction_norm", "lbm_max_v", "fluid_mean_ux", "projection_zone_fluid_mean_ux",
                       "far_field_fluid_mean_ux")


# The penalty force beta phi rho (u_s - u), the hydrodynamic force and the
# reaction they give are linear in the projected solid velocity u_s, itself a
# weighted mean of particle velocities (|u| << |u_s| in the box runs): their
# relative rounding follows the particle velocity's, so they also take the
# particle speed columns' FP32 method error as theirs.
LINKED_METHOD = {k: ("mpm_max_speed", "solid_mean_vx_norm")
                 for k in ("cell_force_max_norm", "hydro_force_max_norm", "max_grid_reaction_norm")}


def column_error(row, ref, k):
    scale = abs(float(ref[SCALE_OF[k]]))
    d = abs(float(row[k]) - float(ref[k]))
    if scale == 0.0:
        return d
    return d / scale


def row_problems(row, ref32, ref64=None, tol=1.0e-5, method_so_far=None):
    """Columns of `row` that miss the bar against the oracle rows (FP32, and
    FP64 when given); returns {column: message}.  `method_so_far` holds, per
    column, the largest FP32 method error of the run up to this row (updated
    here): the FP32 oracle's rounding of the tiny velocities is noise of that
    size from step to step, so one step's value is a poor yardstick."""
    bad = {}
    for k in INT_COLUMNS + TEXT_COLUMNS:
        if row[k] != ref32[k]:
            bad[k] = f"{row[k]!r} != oracle {ref32[k]!r}"
    if ref64 is not None and method_so_far is not None:
        for k in FLOAT_COLUMNS:
            method_so_far[k] = max(column_error(ref32, ref64, k), method_so_far.get(k, 0.0))
    for k in FLOAT_COLUMNS:
        e32 = column_error(row, ref32, k)
        if e32 <= tol:
            continue
        msg = f"{row[k]:.9g} vs oracle {ref32[k]:.9g} (error {e32:.2e} of scale)"
        if ref64 is not None:
            e64 = column_error(row, ref64, k)
            if e64 <= tol:
                continue
            m = column_error(ref32, ref64, k)
            if method_so_far is not None:
                m = max([method_so_far[k]] + [method_so_far[j] for j in LINKED_METHOD.get(k, ())])
            if k in CONDITIONED_COLUMNS and (e32 <= m or e64 <= 2.0 * m):
                continue
            msg += f"; vs FP64 {e64:.2e}, FP32 method error {m:.2e}"
        bad[k] = msg
    return bad


def oracle_rows(o, mode, step):
    """Oracle rows (FP32 and, when the oracle carries one, FP64) for `step`."""
    r32 = dict(step10.oracle_row(o, mode), step=step, coupling_mode=mode)
    r64 = None
    if getattr(o, "fp64", None) is not None:
        r64 = dict(step10.oracle_row(o.fp64, mode), step=step, coupling_mode=mode)
    return r32, r64


GOLDEN_RTOL = 2.0e-3  # the reference run's fast-math FP32 with atomics (test_oracle_vs_reference_artifacts.py)


def run_step10(mode, native, device, out_dir, with_fp64=False):
    """The reference's Step-10 mode matrix (32^3 box, 4096 box particles, 20
    steps x 10 MPM substeps) through FSIDriver3D on `native`, beside the
    oracle.  Returns (driver, oracle, {step: problems}) where the problems are
    the row's misses against the oracle rows (every step) and against the
    reference's committed rows (steps 0, 10, 20)."""
    import numpy as np

    from mpm_lbm.sim.drivers.fsi_driver import FSIDriver3D
    from oracle.oracle import OracleFSI

    cfg = step10.config(mode)
    cloud = step10.box_cloud(cfg)
    d = FSIDriver3D(cfg, str(out_dir), particle_cloud=cloud, native=native, device=device, write_geometry_file=False)
    d.initialize()
    flags = d.lbm.solid.to_numpy()
    o = OracleFSI(cfg, flags, cloud, step10.MATERIAL)
    if with_fp64:
        o.fp64 = OracleFSI(cfg, flags, cloud, step10.MATERIAL, dtype=np.float64)
    golden = step10.golden_rows(mode)
    problems = {}
    method = {}
    for step in range(0, cfg.n_lbm_steps + 1):
        if step:
            d.step_once()
            o.step_once()
            if with_fp64:
                o.fp64.step_once()
        row = d.collect_diagnostics(step)
        r32, r64 = oracle_rows(o, mode, step)
        p = row_problems(row, r32, r64, method_so_far=method)
        if step in golden:
            p.update({f"golden:{m}": "" for m in step10.compare_to_golden(row, golden[step], GOLDEN_RTOL)})
        if p:
            problems[step] = p
    return d, o, problems


def golden_particle_problems(d, cloud_x0):
    """Final particle positions and J against the reference's committed
    particles_x/J.npy of the moving-boundary Step-10 run: positions within
    GOLDEN_RTOL of the largest displacement, J within GOLDEN_RTOL of max |J - 1|."""
    import numpy as np

    g = step10.GOLDEN / "moving_boundary"
    gx = np.load(g / "particles_x.npy")
    gJ = np.load(g / "particles_J.npy")
    x = d.solid.x.to_numpy()
    J = d.solid.Jp.to_numpy()
    out = []
    ex = float(np.abs(x - gx).max()) / float(np.abs(gx - cloud_x0).max())
    eJ = float(np.abs(J - gJ).max()) / float(np.abs(gJ - 1).max())
    if ex > GOLDEN_RTOL:
        out.append(f"particles_x: {ex:.2e} of the largest displacement")
    if eJ > GOLDEN_RTOL:
        out.append(f"particles_J: {eJ:.2e} of max |J - 1|")
    return out, ex, eJ


def run_workload_rows(w, native, device, out_dir, n_steps):
    """A workload (mpm_lbm.workloads) through FSIDriver3D beside the FP32 and
    FP64 oracles; returns (driver, oracle, {step: problems}) for the rows of
    steps 0..n_steps against the oracle rows."""
    from tests.parity import make_pair

    mode = w.config.coupling_mode
    d, o = make_pair(w, out_dir, native, device, with_fp64=True)
    problems, method = {}, {}
    for step in range(0, n_steps + 1):
        if step:
            d.step_once()
            o.step_once()
            o.fp64.step_once()
        row = d.collect_diagnostics(step)
        r32, r64 = oracle_rows(o, mode, step)
        p = row_problems(row, r32, r64, method_so_far=method)
        if p:
            problems[step] = p
    return d, o, problems
