"""FSIDriver3D: the drop-in FSI driver running on gfx950.

Same constructor, attributes, step paths, diagnostics rows, CSV/NPZ outputs
and timing keys as the reference (sim/drivers/fsi_driver.py:62-858).  The
solver objects it builds (LBMFluid3D, MPMSolid3D, projector, couplers) run
hand-written HIP kernels; the moving-boundary subcycle
(fsi_driver.py:533-550) is issued as one fused sequence
(`LBMFluid3D.run_moving_boundary_subcycle`) without per-substep host work.

Extension: `particle_cloud=` (dict with x, vol0, mass and optional
fixed_base_mask / free_tip_proxy_mask / v) replaces the reference's lattice
sampling with caller-supplied particles, e.g. seeded jittered clouds.
"""
from __future__ import annotations

import json
import os
import time
from dataclasses import replace
from pathlib import Path

import numpy as np

from ...diagnostics.fsi_diagnostics import FSIDiagnostics3D
from ..coupling.link_area import LinkAreaMovingBoundaryCoupler3D
from ..coupling.moving_boundary import MovingBoundaryFSICoupler3D
from ..coupling.penalty import PenaltyFSICoupler3D
from ..coupling.projection import MPMToLBMProjector3D
from ..geometry.config import GeometryConfig
from ..geometry.duct_flap_proxy import duct_flap_proxy_static_geometry
from ..geometry.sampler import GeometrySampler3D
from ..io.run_utils import assert_no_nan_inf_array, ensure_output_dir, save_csv_rows, save_json_config
from ..lbm.fluid import LBMFluid3D
from ..lbm.restart import expected_restart_metadata_from_config, load_lbm_restart_to_lbm
from ..monitoring.fluent_like import official_point_like_displacement
from ..mpm.solid import MPMSolid3D
from ..units.mapper import GridUnitMapper
from .fsi_config import FSIDriverConfig

DIAGNOSTIC_FIELDS = [
    "step", "total_mpm_substeps", "coupling_mode", "rho_min", "rho_max", "lbm_max_v", "fluid_mean_ux",
    "projection_zone_fluid_mean_ux", "far_field_fluid_mean_ux", "solid_mean_vx_norm", "mpm_min_J",
    "mpm_max_speed", "projected_mass", "active_cell_count", "cell_force_max_norm", "hydro_force_max_norm",
    "bb_link_count", "bb_max_correction", "mb_subcycle_force_accumulation_mode", "mb_subcycle_force_sample_count",
    "mb_subcycle_force_accum_norm_max", "mb_subcycle_force_mean_norm_max", "active_reaction_particle_count",
    "max_grid_reaction_norm", "elapsed_seconds",
]
DIAGNOSTIC_TEXT_FIELDS = {"coupling_mode", "mb_subcycle_force_accumulation_mode"}
TIMING_KEYS = ("init_time", "projection_time", "coupling_time", "lbm_step_time", "mpm_substep_time",
               "diagnostics_time", "export_time", "wall_velocity_application_time",
               "geometry_motion_interface_time", "total_time")
FLAP_TIP_FIELDS = ["step", "time_s", "flap_tip_total_displacement_m", "flap_tip_x_displacement_m",
                   "flap_tip_y_displacement_m"]
FLUENT_LIKE_FIELDS = [
    "step", "time_s", "free_tip_particle_count", "free_tip_mean_total_displacement_m",
    "free_tip_mean_x_displacement_m", "free_tip_mean_y_displacement_m", "free_tip_mean_z_displacement_m",
    "free_tip_max_total_displacement_m", "official_point_like_particle_count",
    "official_point_like_total_displacement_m", "official_point_like_x_displacement_m",
    "official_point_like_y_displacement_m", "official_point_like_z_displacement_m",
    "official_point_like_target_x_norm", "official_point_like_target_y_norm", "official_point_like_target_z_norm",
    "monitor_is_direct_fluent_equivalent",
]


class FSIDriver3D:
    def __init__(self, config: FSIDriverConfig, out_dir: str, particle_cloud=None, native=None, device=None,
                 write_geometry_file: bool = True, use_graph: bool = True):
        self.config = config
        self.out_dir = out_dir
        self.sim = config.make_unified_sim_config()
        self.mapper = GridUnitMapper.from_sim_config(self.sim)
        self.current_lbm_step = 0
        self.total_lbm_substeps = 0
        self.total_mpm_substeps = 0
        self.diagnostics_rows = []
        self.timing = {key: 0.0 for key in TIMING_KEYS}
        self.geo_path = os.path.join(self.out_dir, f"geo_all_fluid_{self.config.n_grid}.dat")
        self.initialized = False
        self.start_time = None
        self.lbm = None
        self.solid = None
        self.projector = None
        self.penalty_coupler = None
        self.mb_coupler = None
        self.link_area_coupler = None
        self.geometry_quality_report = None
        self.boundary_motion_interface_report = None
        self.duct_boundary_condition_report = None
        self.duct_static_geometry_report = None
        self.wall_velocity_application_report = None
        self.wall_velocity_application_reports = []
        self.geometry_motion_interface_report = None
        self.sampling_stats = None
        self.initial_particle_positions = None
        self.free_tip_proxy_mask = None
        self.flap_tip_monitor_rows = []
        self.fluent_like_monitor_rows = []
        self.material_reference_used_for_mpm_config = False
        self.lbm_restart_load_report = None
        self._particle_cloud = particle_cloud
        if native is None and config.mpm_transfer_order == "particle":
            # the particle-ordered mode runs the kernels built without fma
            # fusion: the same arithmetic as the serial CPU build and the oracle
            from ... import _native as _nat

            native = _nat.load_hip(exact=True)
        self._native = native
        self._device = device
        self._write_geometry_file = write_geometry_file
        self._geometry_config_cache = None
        # The subcycled FSI step is replayed from a captured HIP graph after
        # one eager step (allocations happen then); see _step_moving_boundary_subcycled.
        self.use_graph = bool(use_graph)
        self._fsi_graph = None
        self._eager_subcycled_steps = 0

    # ------------------------------------------------------------- set-up
    def _unsupported_features(self):
        c = self.config
        if c.wall_velocity_application_mode != "disabled" or c.geometry_motion_application_mode != "disabled":
            raise NotImplementedError("prescribed wall-velocity / geometry-motion application is not implemented on the GPU path")
        if c.boundary_motion_report_enabled:
            raise NotImplementedError("boundary-motion interface reports are not implemented on the GPU path")
        if c.quality_check_enabled:
            raise NotImplementedError("geometry quality checks are not implemented on the GPU path")
        if c.write_vtk:
            raise NotImplementedError("write_vtk is not implemented on the GPU path")

    def initialize(self):
        """fsi_driver.py:110-190."""
        t0 = time.perf_counter()
        self._unsupported_features()
        ensure_output_dir(self.out_dir)
        geometry_config = None
        if not (self.config.geometry_type == "box" and self.config.geometry_config_path is None):
            geometry_config = self._make_geometry_config()
        solid_flags = self._static_lbm_geometry(geometry_config)
        save_json_config(self.config, os.path.join(self.out_dir, "driver_config.json"))
        lbm_config = self._make_lbm_config()
        self._write_lbm_boundary_condition_report(lbm_config)
        self.lbm = LBMFluid3D(lbm_config, native=self._native, device=self._device)
        self.lbm.set_solid_flags(solid_flags)
        self.lbm.init_simulation()
        self._load_lbm_restart_if_configured()

        overrides = {"gravity": self.config.gravity, "box_min": self.config.box_min, "box_max": self.config.box_max}
        overrides.update(self.config.make_mpm_control_overrides())
        overrides.update(self._mpm_material_overrides(geometry_config))
        self.solid = MPMSolid3D(self.sim.make_mpm_config(**overrides), n_particles=self.config.n_particles,
                                native=self._native, device=self._device, domain_extent=self.sim.domain_extent)
        cloud = self._particle_cloud
        if cloud is None and not (self.config.geometry_type == "box" and self.config.geometry_config_path is None):
            cloud = GeometrySampler3D(geometry_config).sample_particles()
            self.sampling_stats = dict(cloud["sampling_stats"])
        if cloud is None:
            self.solid.init_box()
        else:
            self.initial_particle_positions = np.asarray(cloud["x"], dtype=np.float32).copy()
            tip = cloud.get("free_tip_proxy_mask")
            self.free_tip_proxy_mask = (np.zeros((self.config.n_particles,), dtype=bool) if tip is None
                                        else np.asarray(tip, dtype=bool))
            self.solid.init_from_numpy(cloud["x"], cloud["vol0"], cloud["mass"], v_np=cloud.get("v"),
                                       fixed_mask_np=cloud.get("fixed_base_mask"))
        v0 = self.config.initial_solid_velocity_norm
        self.solid.set_uniform_velocity(float(v0[0]), float(v0[1]), float(v0[2]))

        self.solid.transfer_order = self.config.mpm_transfer_order
        self.projector = MPMToLBMProjector3D(self.sim)
        self.projector.transfer_order = self.config.mpm_transfer_order
        if self.config.coupling_mode == "penalty":
            self.penalty_coupler = PenaltyFSICoupler3D(self.sim, beta_lbm=self.config.beta_lbm,
                                                       force_cap_lbm=self.config.penalty_force_cap_lbm)
        if self.config.coupling_mode == "moving_boundary" and self.config.reaction_transfer_mode == "engineering":
            self.mb_coupler = MovingBoundaryFSICoupler3D(self.sim, reaction_scale=self.config.mb_reaction_scale,
                                                         force_cap_norm=self.config.mb_force_cap_norm, phi_min=1.0e-6)
        if (self.config.coupling_mode == "moving_boundary"
                and self.config.reaction_transfer_mode == "link_area_experimental"):
            self.link_area_coupler = LinkAreaMovingBoundaryCoupler3D(
                self.sim, area_policy=self.config.link_area_policy, reaction_scale=self.config.mb_reaction_scale,
                force_cap_norm=self.config.mb_force_cap_norm, phi_min=1.0e-6,
                area_scale_min=self.config.link_area_scale_min, area_scale_max=self.config.link_area_scale_max)
        self.initialized = True
        self.start_time = time.perf_counter()
        self.timing["init_time"] += self.start_time - t0

    def _load_lbm_restart_if_configured(self):
        """fsi_driver.py:192-216.  A relative lbm_restart_path is taken from the
        current working directory (the reference's is from its repository root)."""
        if not self.config.lbm_restart_path:
            self.lbm_restart_load_report = {"restart_loaded": False, "restart_path": "",
                                            "restart_required": bool(self.config.lbm_restart_required)}
            return self.lbm_restart_load_report
        path = Path(self.config.lbm_restart_path)
        if not path.is_absolute():
            path = Path.cwd() / path
        report = load_lbm_restart_to_lbm(self.lbm, path, expected_restart_metadata_from_config(self.config),
                                         required=bool(self.config.lbm_restart_required))
        report["restart_required"] = bool(self.config.lbm_restart_required)
        report["restart_scope"] = self.config.lbm_restart_scope
        self.lbm_restart_load_report = report
        with open(os.path.join(self.out_dir, "lbm_restart_load_report.json"), "w", encoding="utf-8") as f:
            json.dump(report, f, indent=2, sort_keys=True)
            f.write("\n")
        return report

    def _static_lbm_geometry(self, geometry_config):
        """fsi_driver.py:219-236: duct flags (written as a Fortran-order .dat like the reference) or all fluid."""
        shape = (self.sim.nx, self.sim.ny, self.sim.nz)
        if self.config.lbm_boundary_condition_mode == "duct_velocity_inlet_pressure_outlet":
            if geometry_config is None or geometry_config.geometry_type != "duct_flap_proxy":
                raise ValueError("duct_velocity_inlet_pressure_outlet requires duct_flap_proxy geometry_config")
            flags, report = duct_flap_proxy_static_geometry(shape, geometry_config)
            self.geo_path = os.path.join(self.out_dir, report["geometry_name"])
            if self._write_geometry_file:
                np.savetxt(self.geo_path, flags.reshape(-1, order="F"), fmt="%d")
            report = dict(report)
            report["geo_path"] = os.path.basename(self.geo_path)
            report["lbm_boundary_condition_mode"] = self.config.lbm_boundary_condition_mode
            self.duct_static_geometry_report = report
            with open(os.path.join(self.out_dir, "duct_static_geometry_report.json"), "w", encoding="utf-8") as f:
                json.dump(report, f, indent=2, sort_keys=True)
                f.write("\n")
            return flags
        flags = np.zeros(shape, dtype=np.int8)
        if self._write_geometry_file:
            np.savetxt(self.geo_path, flags.reshape(-1, order="F"), fmt="%d")
        return flags

    def _make_lbm_config(self):
        """fsi_driver.py:238-257."""
        if self.config.lbm_boundary_condition_mode == "default_periodic":
            return self.sim.make_lbm_config()
        if self.config.lbm_boundary_condition_mode != "duct_velocity_inlet_pressure_outlet":
            raise ValueError(f"unsupported lbm_boundary_condition_mode: {self.config.lbm_boundary_condition_mode}")
        target = tuple(float(v) for v in self.config.target_u_lbm)
        if self.config.velocity_inlet_side == "min":
            return self.sim.make_lbm_config(bc_x_left=2, bc_x_right=1, vel_bc_x_left=target, rho_bc_x_right=1.0)
        return self.sim.make_lbm_config(bc_x_left=1, bc_x_right=2, rho_bc_x_left=1.0, vel_bc_x_right=target)

    def _write_lbm_boundary_condition_report(self, lbm_config):
        """fsi_driver.py:259-307."""
        if self.config.lbm_boundary_condition_mode == "default_periodic":
            return None
        st = self.duct_static_geometry_report or {}
        inlet = int(st.get("inlet_fluid_cell_count", 0))
        outlet = int(st.get("pressure_outlet_fluid_cell_count", 0))
        if self.config.velocity_inlet_side != "min":
            inlet, outlet = outlet, inlet
        c = self.config
        report = {
            "all_population_equilibrium_reset_used": c.lbm_open_boundary_semantics == "equilibrium_all_population_reset",
            "bc_x_left": int(lbm_config.bc_x_left),
            "bc_x_right": int(lbm_config.bc_x_right),
            "boundary_condition_equivalence_claim_allowed": False,
            "duct_wall_cell_count": int(st.get("duct_wall_cell_count", 0)),
            "flow_dimensionality_mode": c.flow_dimensionality_mode,
            "implemented_axis": "x",
            "lbm_boundary_condition_mode": c.lbm_boundary_condition_mode,
            "lbm_open_boundary_semantics": c.lbm_open_boundary_semantics,
            "open_boundary_limiter_enabled": bool(c.open_boundary_limiter_enabled),
            "periodic_boundary_used": False,
            "pressure_outlet_density": 1.0,
            "pressure_outlet_cell_count": outlet,
            "pressure_outlet_side": c.pressure_outlet_side,
            "reaction_transfer_mode": c.reaction_transfer_mode,
            "solid_dimensionality": c.solid_dimensionality,
            "solid_model": c.solid_model,
            "target_u_lbm": list(c.target_u_lbm),
            "target_u_lbm_applied_to_inlet": True,
            "target_u_lbm_applied_to_solid_initial_velocity": False,
            "unknown_population_reconstruction_used": False,
            "validation_claim_allowed": False,
            "velocity_inlet_axis": c.velocity_inlet_axis,
            "velocity_inlet_cell_count": inlet,
            "velocity_inlet_side": c.velocity_inlet_side,
            "velocity_inlet_target": list(c.target_u_lbm),
            "vel_bc_x_left": list(lbm_config.vel_bc_x_left),
            "vel_bc_x_right": list(lbm_config.vel_bc_x_right),
            "viscosity_mapping": c.lbm_viscosity_mapping_report(),
        }
        self.duct_boundary_condition_report = report
        with open(os.path.join(self.out_dir, "duct_boundary_condition_report.json"), "w", encoding="utf-8") as f:
            json.dump(report, f, indent=2, sort_keys=True)
            f.write("\n")
        return report

    def _mpm_material_overrides(self, geometry_config):
        """fsi_driver.py:309-322."""
        if geometry_config is None or geometry_config.geometry_type != "duct_flap_proxy":
            self.material_reference_used_for_mpm_config = False
            return {}
        material = geometry_config.material_reference or {}
        if not bool(material.get("used_for_mpm_config", False)):
            self.material_reference_used_for_mpm_config = False
            return {}
        self.material_reference_used_for_mpm_config = True
        return {"p_rho": float(material["density"]), "young_modulus": float(material["youngs_modulus"]),
                "poisson_ratio": float(material["poisson_ratio"])}

    def _make_geometry_config(self):
        """fsi_driver.py:324-351."""
        if self._geometry_config_cache is not None:
            return self._geometry_config_cache
        if self.config.geometry_config_path is None:
            gc = GeometryConfig(geometry_type=self.config.geometry_type, n_particles=self.config.n_particles,
                                box_min=self.config.box_min, box_max=self.config.box_max)
        else:
            gc = GeometryConfig.from_json(self.config.geometry_config_path)
            if gc.geometry_type != self.config.geometry_type:
                raise ValueError(
                    "geometry_config_path geometry_type does not match FSIDriverConfig.geometry_type: "
                    f"{gc.geometry_type} != {self.config.geometry_type}")
            if gc.n_particles != self.config.n_particles:
                raise ValueError(
                    "geometry_config_path n_particles does not match FSIDriverConfig.n_particles: "
                    f"{gc.n_particles} != {self.config.n_particles}")
            if self.config.quality_check_enabled or self.config.quality_check_strict or self.config.quality_report_path:
                gc = replace(gc, quality_check_enabled=bool(self.config.quality_check_enabled),
                             quality_check_strict=bool(self.config.quality_check_strict),
                             quality_report_path=self.config.quality_report_path)
        self._geometry_config_cache = gc
        return gc

    # ---------------------------------------------------------------- steps
    def step_once(self):
        """fsi_driver.py:425-445."""
        if not self.initialized:
            self.initialize()
        if self.config.fsi_exchange_mode == "lbm_subcycled_per_fsi_step":
            if self.config.coupling_mode != "moving_boundary":
                raise ValueError("subcycled FSI exchange currently supports only moving_boundary coupling")
            self._step_moving_boundary_subcycled()
            self.current_lbm_step += 1
            return
        mode = self.config.coupling_mode
        if mode == "none":
            self._step_none()
        elif mode == "penalty":
            self._step_penalty()
        elif mode == "moving_boundary":
            self._step_moving_boundary()
        else:
            raise ValueError(f"unsupported coupling_mode: {mode}")
        self.current_lbm_step += 1
        self.total_lbm_substeps += 1

    def _project(self):
        t0 = time.perf_counter()
        self.projector.project(self.solid, self.lbm)
        self.timing["projection_time"] += time.perf_counter() - t0

    def _step_none(self):
        """fsi_driver.py:484-495."""
        self._project()
        t0 = time.perf_counter()
        self.lbm.step()
        self.timing["lbm_step_time"] += time.perf_counter() - t0
        t0 = time.perf_counter()
        self.solid.subcycle_coupled(self.config.mpm_substeps_per_lbm_step)
        self.total_mpm_substeps += self.config.mpm_substeps_per_lbm_step
        self.timing["mpm_substep_time"] += time.perf_counter() - t0

    def _step_penalty(self):
        """fsi_driver.py:497-517."""
        self._project()
        t0 = time.perf_counter()
        self.penalty_coupler.clear_force_fields(self.lbm)
        self.penalty_coupler.build_penalty_force(self.lbm)
        self.timing["coupling_time"] += time.perf_counter() - t0
        t0 = time.perf_counter()
        self.lbm.step()
        self.timing["lbm_step_time"] += time.perf_counter() - t0
        t0 = time.perf_counter()
        # clear_grid, p2g, add_lbm_reaction_to_mpm_grid, grid_update, g2p - fused per substep
        self.solid.subcycle_coupled(self.config.mpm_substeps_per_lbm_step, self.lbm, self.penalty_coupler, 1)
        self.total_mpm_substeps += self.config.mpm_substeps_per_lbm_step
        self.timing["mpm_substep_time"] += time.perf_counter() - t0

    def _step_moving_boundary(self):
        """fsi_driver.py:519-531 (no accumulator: hydro_force is the step's link force)."""
        self._project()
        t0 = time.perf_counter()
        self.lbm.update_dynamic_solid(self.config.dynamic_solid_threshold)
        self.lbm.reinitialize_new_fluid_cells()
        self.timing["coupling_time"] += time.perf_counter() - t0
        t0 = time.perf_counter()
        self.lbm.step_moving_bounceback()
        self.timing["lbm_step_time"] += time.perf_counter() - t0
        self._advance_mpm_with_moving_boundary_reaction()

    def _step_moving_boundary_subcycled(self):
        """fsi_driver.py:533-550, issued as one fused device sequence.

        Every FSI step launches the same kernels with the same arguments (the
        population buffers swap an even number of times per step), so after
        one eager step the whole step - projection, flag rebuild, the fused
        substeps, finalize, moments and the MPM substeps - is captured once
        as a HIP graph and replayed."""
        n = self.config.lbm_substeps_per_fsi_step
        if self._graph_ready():
            t0 = time.perf_counter()
            if self._fsi_graph is None:
                self._capture_subcycled_step()
            self._fsi_graph.replay()
            self.total_lbm_substeps += n
            self.total_mpm_substeps += self.config.mpm_substeps_per_lbm_step
            self.timing["lbm_step_time"] += time.perf_counter() - t0
            return
        self._project()
        t0 = time.perf_counter()
        self.lbm.run_moving_boundary_subcycle(n, self.config.dynamic_solid_threshold, accumulate=True)
        self.total_lbm_substeps += n
        self.timing["lbm_step_time"] += time.perf_counter() - t0
        self._advance_mpm_with_moving_boundary_reaction()
        self._eager_subcycled_steps += 1

    def tune_launch(self, cache_path=None, launches: int = 20):
        """Pick the fused substep's launch shape for this lattice by timing the
        candidates on its live state (mpm_lbm.tuning.tune_substep; results do
        not change).  Call after a first step, so the flags and walls are
        those of the run; a captured step graph is dropped and captured again
        with the new shape.  No reference counterpart."""
        from ... import tuning

        if not self.initialized:
            self.initialize()
        rec = tuning.tune_substep(self.lbm, tuning.driver_substep_flags(self), launches=launches,
                                  cache_path=cache_path)
        self._drop_step_graph()
        return rec

    def _drop_step_graph(self):
        """Forget the captured FSI-step graph (its launches hold the state the
        driver had at capture); the next step runs eagerly and recaptures."""
        self._fsi_graph = None
        self._eager_subcycled_steps = 0

    def _graph_ready(self) -> bool:
        # the link-area transfer sets its scale on the host from the step's
        # link statistics, between the LBM and MPM parts: no graph
        return (self.use_graph and self.lbm.device.type == "cuda" and self._eager_subcycled_steps >= 1
                and self.config.lbm_substeps_per_fsi_step % 2 == 0 and self.link_area_coupler is None)

    def _capture_subcycled_step(self):
        import torch

        torch.cuda.synchronize(self.lbm.device)
        graph = torch.cuda.CUDAGraph()
        counted = self.total_mpm_substeps
        with torch.cuda.graph(graph):
            self.projector.project(self.solid, self.lbm)
            self.lbm.run_moving_boundary_subcycle(self.config.lbm_substeps_per_fsi_step,
                                                  self.config.dynamic_solid_threshold, accumulate=True)
            self._advance_mpm_with_moving_boundary_reaction()
        self.total_mpm_substeps = counted  # capture records the kernels without running them
        self._fsi_graph = graph

    def _advance_mpm_with_moving_boundary_reaction(self):
        """fsi_driver.py:552-575 (engineering or link-area reaction transfer)."""
        t0 = time.perf_counter()
        coupler = self.mb_coupler
        if self.link_area_coupler is not None:
            self.link_area_coupler.update_area_scale_from_lbm(self.lbm)
            coupler = self.link_area_coupler
        # clear_grid, p2g, clear_reaction_diagnostics, add_*_reaction_to_mpm_grid,
        # grid_update, g2p - as one fused device sequence per substep
        self.solid.subcycle_coupled(self.config.mpm_substeps_per_lbm_step, self.lbm, coupler, 2)
        self.total_mpm_substeps += self.config.mpm_substeps_per_lbm_step
        self.timing["mpm_substep_time"] += time.perf_counter() - t0

    def run(self, checkpoint_dir=None, checkpoint_every: int = 0, resume: bool = False,
            stop_after_steps=None, keep_last: int = 3):
        """fsi_driver.py:577-595, with checkpoint and resume (no reference
        counterpart in the driver; the contract of the reference's
        checkpointed rows, tests/test_step120_checkpoint_restart_contract.py):

        * checkpoint_every > 0 writes a checkpoint to `checkpoint_dir` every
          that many FSI steps (keeping the newest `keep_last`);
        * resume=True continues from the newest checkpoint in `checkpoint_dir`
          (a fresh run when there is none) up to n_lbm_steps;
        * stop_after_steps ends the run after that many FSI steps of this
          call, without the final outputs (an interrupted run), unless those
          were its last steps."""
        from ...checkpoint import latest_fsi_checkpoint, load_fsi_checkpoint, save_fsi_checkpoint

        total_start = time.perf_counter()
        if not self.initialized:
            self.initialize()
        resumed = False
        if resume:
            if checkpoint_dir is None:
                raise ValueError("resume=True needs checkpoint_dir")
            latest = latest_fsi_checkpoint(checkpoint_dir)
            if latest is not None:
                load_fsi_checkpoint(self, latest)
                resumed = True
        if checkpoint_every and checkpoint_dir is None:
            raise ValueError("checkpoint_every needs checkpoint_dir")
        if not resumed:
            self.collect_diagnostics(0)
        # a fresh run takes n_lbm_steps steps (as the reference's run); a
        # resumed one the steps its checkpoint had left
        steps = (max(0, self.config.n_lbm_steps - self.current_lbm_step) if resumed
                 else self.config.n_lbm_steps)
        for done in range(1, steps + 1):
            self.step_once()
            if (self.current_lbm_step % self.config.output_interval == 0
                    or self.current_lbm_step == self.config.n_lbm_steps):
                self.collect_diagnostics(self.current_lbm_step)
            if checkpoint_every and self.current_lbm_step % int(checkpoint_every) == 0:
                save_fsi_checkpoint(self, checkpoint_dir, keep_last=keep_last)
            if stop_after_steps is not None and done >= int(stop_after_steps) and done < steps:
                self.timing["total_time"] = time.perf_counter() - total_start
                return self.diagnostics_rows
        self.export_outputs(self.current_lbm_step)
        self.save_timeseries()
        self.timing["total_time"] = time.perf_counter() - total_start
        return self.diagnostics_rows

    def save_checkpoint(self, directory, keep_last: int = 3):
        """Checkpoint after the current FSI step (mpm_lbm.checkpoint)."""
        from ...checkpoint import save_fsi_checkpoint

        return save_fsi_checkpoint(self, directory, keep_last=keep_last)

    def load_checkpoint(self, path):
        """Resume from a checkpoint (mpm_lbm.checkpoint)."""
        from ...checkpoint import load_fsi_checkpoint

        return load_fsi_checkpoint(self, path)

    # ----------------------------------------------------------- diagnostics
    def collect_diagnostics(self, step: int):
        """fsi_driver.py:597-669: the same 25 columns, reduced on the device."""
        t0 = time.perf_counter()
        self.lbm.check_resident_waits()
        self.solid.check_resident_waits()
        ls = self.lbm.reduce_stats()
        ms = self.solid.reduce_stats()
        fluid = FSIDiagnostics3D.lbm_fluid_stats(self.lbm, stats=ls)
        zone = FSIDiagnostics3D.projection_zone_fluid_mean_velocity(self.lbm, stats=ls)
        far = FSIDiagnostics3D.far_field_fluid_mean_velocity(self.lbm, stats=ls)
        solid_v = FSIDiagnostics3D.solid_mean_velocity_norm(self.solid, stats=ms)
        proj = self.projector.get_stats()
        forces = FSIDiagnostics3D.force_stats(self.lbm, stats=ls)
        bb_link_count, bb_max_correction = 0, 0.0
        if self.config.coupling_mode == "moving_boundary":
            mbs = self.lbm.get_moving_boundary_stats()
            bb_link_count, bb_max_correction = mbs["bb_link_count"], mbs["bb_max_correction"]
        acc = self.lbm.get_moving_boundary_accumulation_stats()
        samples = acc["mb_subcycle_force_sample_count"]
        acc_mode = ("mean_force_accumulator"
                    if self.config.fsi_exchange_mode == "lbm_subcycled_per_fsi_step" and samples > 0 else "not_active")
        active_reaction, max_grid_reaction = 0, 0.0
        if self.config.coupling_mode == "moving_boundary" and (self.mb_coupler or self.link_area_coupler) is not None:
            st = (self.mb_coupler or self.link_area_coupler).get_stats()
            active_reaction, max_grid_reaction = st["active_reaction_particle_count"], st["max_grid_reaction_norm"]
        elif self.config.coupling_mode == "penalty" and self.penalty_coupler is not None:
            max_grid_reaction = self.penalty_coupler.get_stats()["max_reaction_grid_force_norm"]
        elapsed = 0.0 if self.start_time is None else time.perf_counter() - self.start_time
        row = {
            "step": int(step),
            "total_mpm_substeps": int(self.total_mpm_substeps),
            "coupling_mode": self.config.coupling_mode,
            "rho_min": fluid["rho_min"],
            "rho_max": fluid["rho_max"],
            "lbm_max_v": fluid["max_v"],
            "fluid_mean_ux": float(fluid["fluid_mean_velocity"][0]),
            "projection_zone_fluid_mean_ux": float(zone[0]),
            "far_field_fluid_mean_ux": float(far[0]),
            "solid_mean_vx_norm": float(solid_v[0]),
            "mpm_min_J": ms["min_J"],
            "mpm_max_speed": ms["max_speed"],
            "projected_mass": proj["projected_mass"],
            "active_cell_count": proj["active_cell_count"],
            "cell_force_max_norm": forces["max_cell_force_norm"],
            "hydro_force_max_norm": forces["max_hydro_force_norm"],
            "bb_link_count": int(bb_link_count),
            "bb_max_correction": float(bb_max_correction),
            "mb_subcycle_force_accumulation_mode": acc_mode,
            "mb_subcycle_force_sample_count": int(samples),
            "mb_subcycle_force_accum_norm_max": float(acc["mb_subcycle_force_accum_norm_max"]),
            "mb_subcycle_force_mean_norm_max": float(acc["mb_subcycle_force_mean_norm_max"]),
            "active_reaction_particle_count": int(active_reaction),
            "max_grid_reaction_norm": float(max_grid_reaction),
            "elapsed_seconds": elapsed,
        }
        assert_no_nan_inf_array("driver diagnostics", [v for v in row.values() if isinstance(v, (int, float))])
        self.diagnostics_rows.append(row)
        self.collect_flap_tip_monitor(step)
        self.collect_fluent_like_monitor(step)
        self.timing["diagnostics_time"] += time.perf_counter() - t0
        return row

    def collect_flap_tip_monitor(self, step: int):
        """fsi_driver.py:671-691."""
        if self.initial_particle_positions is None or self.free_tip_proxy_mask is None:
            return None
        if not np.any(self.free_tip_proxy_mask):
            return None
        x = self.solid.x.to_numpy()
        m = self.free_tip_proxy_mask
        disp = np.mean(x[m] - self.initial_particle_positions[m], axis=0) * self._duct_length_scale_m()
        row = {
            "step": int(step),
            "time_s": float(step) * self._monitor_time_step_s(),
            "flap_tip_total_displacement_m": float(np.linalg.norm(disp)),
            "flap_tip_x_displacement_m": float(disp[0]),
            "flap_tip_y_displacement_m": float(disp[1]),
        }
        self.flap_tip_monitor_rows.append(row)
        return row

    def collect_fluent_like_monitor(self, step: int):
        """fsi_driver.py:693-741."""
        if not self.config.fluent_like_monitor_enabled or self.initial_particle_positions is None:
            return None
        x = self.solid.x.to_numpy()
        gc = self._make_geometry_config() if self.config.geometry_config_path else None
        target = self._fluent_like_monitor_target_norm(gc)
        radius = self._fluent_like_monitor_radius_norm(gc)
        scale = self._duct_length_scale_m()
        official = official_point_like_displacement(self.initial_particle_positions, x, target,
                                                    nearest_count=self.config.fluent_like_monitor_nearest_count,
                                                    radius_norm=radius, scale_m=scale)
        m = self.free_tip_proxy_mask
        if m is not None and np.any(m):
            disp = (x[m] - self.initial_particle_positions[m]) * scale
            tip_mean = np.mean(disp, axis=0)
            tip_max = float(np.max(np.linalg.norm(disp, axis=1)))
            tip_count = int(np.count_nonzero(m))
        else:
            tip_mean, tip_max, tip_count = np.zeros(3), 0.0, 0
        row = {
            "step": int(step),
            "time_s": float(step) * self._monitor_time_step_s(),
            "free_tip_particle_count": tip_count,
            "free_tip_mean_total_displacement_m": float(np.linalg.norm(tip_mean)),
            "free_tip_mean_x_displacement_m": float(tip_mean[0]),
            "free_tip_mean_y_displacement_m": float(tip_mean[1]),
            "free_tip_mean_z_displacement_m": float(tip_mean[2]),
            "free_tip_max_total_displacement_m": tip_max,
            "official_point_like_particle_count": int(official["selected_particle_count"]),
            "official_point_like_total_displacement_m": official["official_point_like_total_displacement_m"],
            "official_point_like_x_displacement_m": official["official_point_like_x_displacement_m"],
            "official_point_like_y_displacement_m": official["official_point_like_y_displacement_m"],
            "official_point_like_z_displacement_m": official["official_point_like_z_displacement_m"],
            "official_point_like_target_x_norm": float(target[0]),
            "official_point_like_target_y_norm": float(target[1]),
            "official_point_like_target_z_norm": float(target[2]),
            "monitor_is_direct_fluent_equivalent": False,
        }
        self.fluent_like_monitor_rows.append(row)
        return row

    def _fluent_like_monitor_target_norm(self, gc) -> np.ndarray:
        px, py = self.config.fluent_like_monitor_physical_point_m
        if gc is None:
            return np.asarray([px, py, 0.5], dtype=np.float64)
        duct = gc.duct or {}
        meta = gc.dimensional_reference or {}
        dx, dy, dz = duct.get("x", [0.0, 1.0]), duct.get("y", [0.0, 1.0]), duct.get("z", [0.0, 1.0])
        length, height = float(meta.get("duct_length_m", 1.0)), float(meta.get("duct_height_m", 1.0))
        return np.asarray([float(dx[0]) + (float(px) / length) * (float(dx[1]) - float(dx[0])),
                           float(dy[0]) + (float(py) / height) * (float(dy[1]) - float(dy[0])),
                           0.5 * (float(dz[0]) + float(dz[1]))], dtype=np.float64)

    def _fluent_like_monitor_radius_norm(self, gc):
        if self.config.fluent_like_monitor_radius_m is None:
            return None
        if gc is None:
            return float(self.config.fluent_like_monitor_radius_m)
        meta = gc.dimensional_reference or {}
        return float(self.config.fluent_like_monitor_radius_m) / float(meta.get("duct_length_m", 1.0))

    def _monitor_time_step_s(self) -> float:
        gc = self._make_geometry_config() if self.config.geometry_config_path else None
        meta = gc.dimensional_reference if gc is not None else None
        if meta and "transient_dt_s" in meta:
            return float(meta["transient_dt_s"])
        return float(self.sim.lbm_dt_phys)

    def _duct_length_scale_m(self) -> float:
        gc = self._make_geometry_config() if self.config.geometry_config_path else None
        if gc is None:
            return 1.0
        duct = gc.duct or {}
        meta = gc.dimensional_reference or {}
        dx = duct.get("x", [0.0, 1.0])
        return float(meta.get("duct_length_m", 1.0)) / max(float(dx[1]) - float(dx[0]), 1.0e-12)

    def export_outputs(self, step: int):
        t0 = time.perf_counter()
        if self.config.write_particles:
            self.solid.export_particles(self.out_dir)
        self.timing["export_time"] += time.perf_counter() - t0

    def save_timeseries(self):
        """fsi_driver.py:797-848."""
        save_csv_rows(self.diagnostics_rows, os.path.join(self.out_dir, "diagnostics_timeseries.csv"),
                      fieldnames=DIAGNOSTIC_FIELDS)
        payload = {"columns": np.asarray(DIAGNOSTIC_FIELDS),
                   "modes": np.asarray([r["coupling_mode"] for r in self.diagnostics_rows])}
        for field in DIAGNOSTIC_FIELDS:
            values = [r[field] for r in self.diagnostics_rows]
            payload[field] = np.asarray(values) if field in DIAGNOSTIC_TEXT_FIELDS else np.asarray(values, dtype=np.float64)
        np.savez(os.path.join(self.out_dir, "diagnostics_timeseries.npz"), **payload)
        if self.flap_tip_monitor_rows:
            save_csv_rows(self.flap_tip_monitor_rows, os.path.join(self.out_dir, "flap_tip_displacement_timeseries.csv"),
                          fieldnames=FLAP_TIP_FIELDS)
        if self.fluent_like_monitor_rows:
            save_csv_rows(self.fluent_like_monitor_rows, os.path.join(self.out_dir, "fluent_like_monitor_timeseries.csv"),
                          fieldnames=FLUENT_LIKE_FIELDS)

    def final_diagnostics(self):
        return self.diagnostics_rows[-1] if self.diagnostics_rows else None

    def performance_row(self):
        row = {"mode": self.config.coupling_mode}
        row.update(self.timing)
        return row
