"""FSIDriverConfig: the driver's frozen, validated configuration.

Keys, defaults, validation order and error messages follow the reference
(sim/drivers/fsi_config.py:67-497) so existing JSON configs and callers work
unchanged.  Extensions: `lattice_shape` (non-cubic lattices),
`lbm_population_storage` ("f32" | "f16" | "bf16" deviation storage) and
`lbm_collision` ("mrt" | "regularized") and `mpm_transfer_order` ("atomic" |
"particle": the order of the projection's and p2g's node sums, "particle"
being the deterministic order of a sequential scatter).
"""
from dataclasses import asdict, dataclass
import json
import math
from pathlib import Path
from typing import Optional, Tuple

from ..geometry.config import VALID_GEOMETRY_TYPES
from ..lbm.relaxation_semantics import (
    LEGACY_EXTERNAL_SOLVER_RELAXATION_PARAMETER,
    STANDARD_LATTICE_KINEMATIC_VISCOSITY,
    tau_from_lattice_kinematic_viscosity,
    tau_from_legacy_external_solver_parameter,
)
from .sim_config import UnifiedSimConfig

VALID_COUPLING_MODES = ("none", "penalty", "moving_boundary")
VALID_REACTION_TRANSFER_MODES = ("engineering", "link_area_experimental", "interface_traction_conservative")
VALID_LINK_AREA_POLICIES = ("uniform", "inverse_length", "length")
VALID_BOUNDARY_MOTION_MODES = ("static", "prescribed_kinematic")
VALID_WALL_VELOCITY_APPLICATION_MODES = ("disabled", "solid_vel_experimental")
VALID_GEOMETRY_MOTION_MODES = ("static", "prescribed_kinematic")
VALID_GEOMETRY_MOTION_APPLICATION_MODES = ("disabled", "diagnostic_only")
VALID_LBM_BOUNDARY_CONDITION_MODES = ("default_periodic", "duct_velocity_inlet_pressure_outlet")
VALID_LBM_OPEN_BOUNDARY_SEMANTICS = (
    "equilibrium_all_population_reset",
    "zou_he_reconstruct_unknowns",
    "regularized_velocity_pressure",
    "regularized_velocity_pressure_limited",
    "convective_pressure_outlet_experimental",
)
IMPLEMENTED_LBM_OPEN_BOUNDARY_SEMANTICS = (
    "equilibrium_all_population_reset",
    "regularized_velocity_pressure",
    "regularized_velocity_pressure_limited",
    "convective_pressure_outlet_experimental",
)
VALID_LBM_VISCOSITY_SEMANTICS = ("legacy_external", "physical_nu_mapping")
VALID_LBM_TAU_STABILITY_POLICIES = ("report_only", "strict")
VALID_FSI_EXCHANGE_MODES = ("one_lbm_step_per_fsi_step", "lbm_subcycled_per_fsi_step")
VALID_LBM_RESTART_SCOPES = ("rho_velocity_populations",)
VALID_MPM_PLANAR_CONSTRAINT_MODES = ("disabled", "lock_z")
VALID_MPM_PLANAR_CONSTRAINT_AXES = ("z",)
VALID_MPM_DAMPING_APPLICATIONS = ("disabled", "particle_velocity_post_g2p")
VALID_SOLID_MODELS = ("finite_deformation_mpm", "small_strain_linear_elastic")
IMPLEMENTED_SOLID_MODELS = ("finite_deformation_mpm",)
VALID_SOLID_DIMENSIONALITIES = ("three_dimensional", "plane_strain_2d", "plane_stress_2d")
IMPLEMENTED_SOLID_DIMENSIONALITIES = ("three_dimensional",)
VALID_FLOW_DIMENSIONALITY_MODES = ("three_dimensional", "thin_3d_no_slip_z_walls", "d2q9_planar", "d3q19_quasi_2d_periodic_z")
IMPLEMENTED_FLOW_DIMENSIONALITY_MODES = ("three_dimensional", "thin_3d_no_slip_z_walls")
VALID_BOUNDARY_AXES = ("x",)
VALID_BOUNDARY_SIDES = ("min", "max")

_TUPLE3_FIELDS = ("target_u_lbm", "initial_solid_velocity_norm", "gravity", "box_min", "box_max")


def _as_float_tuple(values, name):
    if len(values) != 3:
        raise ValueError(f"{name} must contain exactly three values")
    return tuple(float(v) for v in values)


def _as_float_pair(values, name):
    if len(values) != 2:
        raise ValueError(f"{name} must contain exactly two values")
    return tuple(float(v) for v in values)


def _one_of(value, allowed, name):
    if value not in allowed:
        raise ValueError(f"{name} must be one of {allowed}")


@dataclass(frozen=True)
class FSIDriverConfig:
    coupling_mode: str = "penalty"
    geometry_type: str = "box"
    geometry_config_path: Optional[str] = None
    n_grid: int = 32
    n_particles: int = 4096
    n_lbm_steps: int = 20
    mpm_substeps_per_lbm_step: int = 10
    mpm_dt: float = 4.0e-4
    target_u_lbm: Tuple[float, float, float] = (0.02, 0.0, 0.0)
    initial_solid_velocity_norm: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    gravity: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    lbm_boundary_condition_mode: str = "default_periodic"
    lbm_open_boundary_semantics: str = "equilibrium_all_population_reset"
    open_boundary_limiter_enabled: bool = False
    open_boundary_rho_min: float = 0.8
    open_boundary_rho_max: float = 1.2
    open_boundary_u_max: float = 0.1
    open_boundary_noneq_cap: float = 0.05
    open_boundary_population_floor: Optional[float] = None
    velocity_inlet_axis: str = "x"
    velocity_inlet_side: str = "min"
    pressure_outlet_side: str = "max"
    physical_duct_length_m: float = 1.0
    target_inlet_velocity_mps: Optional[float] = None
    official_fsi_dt_s: Optional[float] = None
    target_u_lbm_for_dimensional_mapping: Optional[float] = None
    lbm_substeps_per_fsi_step: int = 1
    lbm_dt_phys_override_s: Optional[float] = None
    fluid_density_kg_m3: float = 1.225
    fluid_kinematic_viscosity_m2_s: float = 1.5e-5
    target_reynolds_number: Optional[float] = None
    lbm_viscosity_semantics: str = "legacy_external"
    lbm_min_tau_margin: float = 1.0e-4
    lbm_tau_stability_policy: str = "report_only"
    fsi_exchange_mode: str = "one_lbm_step_per_fsi_step"
    lbm_restart_path: Optional[str] = None
    lbm_restart_required: bool = False
    lbm_restart_scope: str = "rho_velocity_populations"
    box_min: Tuple[float, float, float] = (0.25, 0.35, 0.25)
    box_max: Tuple[float, float, float] = (0.55, 0.65, 0.55)
    dynamic_solid_threshold: float = 0.5
    beta_lbm: float = 1.0e-3
    penalty_force_cap_lbm: float = 1.0e-4
    mb_reaction_scale: float = 1.0
    mb_force_cap_norm: float = 1.0e-4
    reaction_transfer_mode: str = "engineering"
    solid_model: str = "finite_deformation_mpm"
    solid_dimensionality: str = "three_dimensional"
    flow_dimensionality_mode: str = "three_dimensional"
    link_area_policy: str = "inverse_length"
    link_area_scale_min: float = 0.25
    link_area_scale_max: float = 2.0
    boundary_motion_mode: str = "static"
    boundary_motion_config_path: Optional[str] = None
    boundary_motion_report_enabled: bool = False
    wall_velocity_application_mode: str = "disabled"
    wall_velocity_application_config_path: Optional[str] = None
    wall_velocity_application_report_enabled: bool = False
    geometry_motion_mode: str = "static"
    geometry_motion_config_path: Optional[str] = None
    geometry_motion_report_enabled: bool = False
    geometry_motion_application_mode: str = "disabled"
    geometry_motion_application_config_path: Optional[str] = None
    geometry_motion_application_report_enabled: bool = False
    mpm_planar_constraint_mode: str = "disabled"
    mpm_planar_constraint_axis: str = "z"
    mpm_velocity_damping: float = 0.0
    mpm_damping_application: str = "disabled"
    fluent_like_monitor_enabled: bool = False
    fluent_like_monitor_physical_point_m: Tuple[float, float] = (0.0505, 0.0095)
    fluent_like_monitor_nearest_count: int = 8
    fluent_like_monitor_radius_m: Optional[float] = None
    output_interval: int = 10
    write_vtk: bool = False
    write_particles: bool = False
    quality_check_enabled: bool = False
    quality_check_strict: bool = False
    quality_report_path: Optional[str] = None
    lattice_shape: Optional[Tuple[int, int, int]] = None
    lbm_population_storage: str = "f32"
    lbm_collision: str = "mrt"
    mpm_transfer_order: str = "atomic"

    def __post_init__(self):
        for name in _TUPLE3_FIELDS:
            object.__setattr__(self, name, _as_float_tuple(getattr(self, name), name))
        object.__setattr__(
            self,
            "fluent_like_monitor_physical_point_m",
            _as_float_pair(self.fluent_like_monitor_physical_point_m, "fluent_like_monitor_physical_point_m"),
        )
        if self.lattice_shape is not None:
            object.__setattr__(self, "lattice_shape", tuple(int(v) for v in self.lattice_shape))
        self._validate_modes()
        self._validate_motion()
        self._validate_mpm_controls()
        self._validate_lbm_boundary()
        self._validate_scalars()
        self._validate_fsi_exchange_mapping()
        self._validate_lbm_viscosity_mapping()
        self._validate_coupling_scalars()

    # -- validation (fsi_config.py:174-332) ---------------------------------
    def _validate_modes(self):
        _one_of(self.coupling_mode, VALID_COUPLING_MODES, "coupling_mode")
        _one_of(self.reaction_transfer_mode, VALID_REACTION_TRANSFER_MODES, "reaction_transfer_mode")
        if self.reaction_transfer_mode == "interface_traction_conservative":
            raise ValueError("reaction_transfer_mode='interface_traction_conservative' is planned but not implemented")
        if self.reaction_transfer_mode == "link_area_experimental" and self.coupling_mode != "moving_boundary":
            raise ValueError("reaction_transfer_mode='link_area_experimental' requires coupling_mode='moving_boundary'")
        _one_of(self.link_area_policy, VALID_LINK_AREA_POLICIES, "link_area_policy")

    def _validate_motion(self):
        _one_of(self.boundary_motion_mode, VALID_BOUNDARY_MOTION_MODES, "boundary_motion_mode")
        if self.boundary_motion_mode == "static" and self.boundary_motion_config_path is not None:
            raise ValueError("boundary_motion_config_path must be None when boundary_motion_mode='static'")
        if self.boundary_motion_mode == "prescribed_kinematic" and not self.boundary_motion_config_path:
            raise ValueError("boundary_motion_config_path is required when boundary_motion_mode='prescribed_kinematic'")
        _one_of(self.wall_velocity_application_mode, VALID_WALL_VELOCITY_APPLICATION_MODES,
                "wall_velocity_application_mode")
        if self.wall_velocity_application_mode == "disabled" and self.wall_velocity_application_config_path is not None:
            raise ValueError("wall_velocity_application_config_path must be None when wall_velocity_application_mode='disabled'")
        if self.wall_velocity_application_mode == "solid_vel_experimental":
            if self.boundary_motion_mode != "prescribed_kinematic":
                raise ValueError("wall_velocity_application_mode='solid_vel_experimental' requires boundary_motion_mode='prescribed_kinematic'")
            if not self.wall_velocity_application_config_path:
                raise ValueError("wall_velocity_application_config_path is required when wall_velocity_application_mode='solid_vel_experimental'")
            if not _path_exists(self.wall_velocity_application_config_path):
                raise ValueError("wall_velocity_application_config_path must exist when wall_velocity_application_mode='solid_vel_experimental'")
        _one_of(self.geometry_motion_mode, VALID_GEOMETRY_MOTION_MODES, "geometry_motion_mode")
        _one_of(self.geometry_motion_application_mode, VALID_GEOMETRY_MOTION_APPLICATION_MODES,
                "geometry_motion_application_mode")
        if self.geometry_motion_mode == "static":
            if self.geometry_motion_config_path is not None:
                raise ValueError("geometry_motion_config_path must be None when geometry_motion_mode='static'")
            if self.geometry_motion_application_mode != "disabled":
                raise ValueError("geometry_motion_application_mode must be disabled when geometry_motion_mode='static'")
        if self.geometry_motion_mode == "prescribed_kinematic":
            if not self.geometry_motion_config_path:
                raise ValueError("geometry_motion_config_path is required when geometry_motion_mode='prescribed_kinematic'")
            if not _path_exists(self.geometry_motion_config_path):
                raise ValueError("geometry_motion_config_path must exist when geometry_motion_mode='prescribed_kinematic'")
        if self.geometry_motion_application_mode == "disabled" and self.geometry_motion_application_config_path is not None:
            raise ValueError("geometry_motion_application_config_path must be None when geometry_motion_application_mode='disabled'")
        if self.geometry_motion_application_mode == "diagnostic_only":
            if self.geometry_motion_mode != "prescribed_kinematic":
                raise ValueError("geometry_motion_application_mode='diagnostic_only' requires geometry_motion_mode='prescribed_kinematic'")
            if not self.geometry_motion_application_config_path:
                raise ValueError("geometry_motion_application_config_path is required when geometry_motion_application_mode='diagnostic_only'")
            if not _path_exists(self.geometry_motion_application_config_path):
                raise ValueError("geometry_motion_application_config_path must exist when geometry_motion_application_mode='diagnostic_only'")

    def _validate_mpm_controls(self):
        _one_of(self.mpm_planar_constraint_mode, VALID_MPM_PLANAR_CONSTRAINT_MODES, "mpm_planar_constraint_mode")
        _one_of(self.mpm_planar_constraint_axis, VALID_MPM_PLANAR_CONSTRAINT_AXES, "mpm_planar_constraint_axis")
        if float(self.mpm_velocity_damping) < 0.0:
            raise ValueError("mpm_velocity_damping must be non-negative")
        _one_of(self.mpm_damping_application, VALID_MPM_DAMPING_APPLICATIONS, "mpm_damping_application")
        if self.mpm_damping_application == "disabled" and float(self.mpm_velocity_damping) != 0.0:
            raise ValueError("mpm_velocity_damping must be 0 when mpm_damping_application is disabled")
        if self.mpm_damping_application == "particle_velocity_post_g2p" and float(self.mpm_velocity_damping) <= 0.0:
            raise ValueError("mpm_velocity_damping must be positive when damping is enabled")

    def _validate_lbm_boundary(self):
        _one_of(self.geometry_type, VALID_GEOMETRY_TYPES, "geometry_type")
        _one_of(self.lbm_boundary_condition_mode, VALID_LBM_BOUNDARY_CONDITION_MODES, "lbm_boundary_condition_mode")
        _one_of(self.lbm_open_boundary_semantics, VALID_LBM_OPEN_BOUNDARY_SEMANTICS, "lbm_open_boundary_semantics")
        if self.lbm_open_boundary_semantics not in IMPLEMENTED_LBM_OPEN_BOUNDARY_SEMANTICS:
            raise ValueError(f"lbm_open_boundary_semantics={self.lbm_open_boundary_semantics!r} is not implemented")
        if self.open_boundary_rho_min <= 0.0:
            raise ValueError("open_boundary_rho_min must be positive")
        if self.open_boundary_rho_max <= self.open_boundary_rho_min:
            raise ValueError("open_boundary_rho_min must be less than open_boundary_rho_max")
        if self.open_boundary_u_max <= 0.0:
            raise ValueError("open_boundary_u_max must be positive")
        if self.open_boundary_noneq_cap <= 0.0:
            raise ValueError("open_boundary_noneq_cap must be positive")
        _one_of(self.lbm_viscosity_semantics, VALID_LBM_VISCOSITY_SEMANTICS, "lbm_viscosity_semantics")
        _one_of(self.lbm_tau_stability_policy, VALID_LBM_TAU_STABILITY_POLICIES, "lbm_tau_stability_policy")
        if float(self.lbm_min_tau_margin) <= 0.0:
            raise ValueError("lbm_min_tau_margin must be positive")
        _one_of(self.solid_model, VALID_SOLID_MODELS, "solid_model")
        if self.solid_model not in IMPLEMENTED_SOLID_MODELS:
            raise ValueError(f"solid_model={self.solid_model!r} is not implemented")
        _one_of(self.solid_dimensionality, VALID_SOLID_DIMENSIONALITIES, "solid_dimensionality")
        if self.solid_dimensionality not in IMPLEMENTED_SOLID_DIMENSIONALITIES:
            raise ValueError(f"solid_dimensionality={self.solid_dimensionality!r} is not implemented")
        _one_of(self.flow_dimensionality_mode, VALID_FLOW_DIMENSIONALITY_MODES, "flow_dimensionality_mode")
        if self.flow_dimensionality_mode not in IMPLEMENTED_FLOW_DIMENSIONALITY_MODES:
            raise ValueError(f"flow_dimensionality_mode={self.flow_dimensionality_mode!r} is not implemented")
        _one_of(self.velocity_inlet_axis, VALID_BOUNDARY_AXES, "velocity_inlet_axis")
        _one_of(self.velocity_inlet_side, VALID_BOUNDARY_SIDES, "velocity_inlet_side")
        _one_of(self.pressure_outlet_side, VALID_BOUNDARY_SIDES, "pressure_outlet_side")
        _one_of(self.fsi_exchange_mode, VALID_FSI_EXCHANGE_MODES, "fsi_exchange_mode")
        _one_of(self.lbm_restart_scope, VALID_LBM_RESTART_SCOPES, "lbm_restart_scope")
        if self.lbm_restart_required and not self.lbm_restart_path:
            raise ValueError("lbm_restart_path is required when lbm_restart_required is true")
        if self.lbm_boundary_condition_mode == "duct_velocity_inlet_pressure_outlet":
            if self.velocity_inlet_axis != "x":
                raise ValueError("duct_velocity_inlet_pressure_outlet currently supports only x-axis inlet/outlet")
            if self.velocity_inlet_side == self.pressure_outlet_side:
                raise ValueError("velocity_inlet_side and pressure_outlet_side must differ")

    def _validate_scalars(self):
        positive_ints = ("n_grid", "n_particles", "n_lbm_steps", "mpm_substeps_per_lbm_step")
        for name in positive_ints:
            if getattr(self, name) <= 0:
                raise ValueError(f"{name} must be positive")
        if self.mpm_dt <= 0.0:
            raise ValueError("mpm_dt must be positive")
        if self.physical_duct_length_m <= 0.0:
            raise ValueError("physical_duct_length_m must be positive")
        if self.lbm_substeps_per_fsi_step <= 0:
            raise ValueError("lbm_substeps_per_fsi_step must be positive")
        for name in ("target_inlet_velocity_mps", "official_fsi_dt_s", "target_u_lbm_for_dimensional_mapping",
                     "lbm_dt_phys_override_s"):
            _validate_optional_positive(getattr(self, name), name)
        if self.fluid_density_kg_m3 <= 0.0:
            raise ValueError("fluid_density_kg_m3 must be positive")
        if self.fluid_kinematic_viscosity_m2_s <= 0.0:
            raise ValueError("fluid_kinematic_viscosity_m2_s must be positive")
        _validate_optional_positive(self.target_reynolds_number, "target_reynolds_number")
        _validate_optional_positive(self.fluent_like_monitor_radius_m, "fluent_like_monitor_radius_m")
        if self.fluent_like_monitor_nearest_count <= 0:
            raise ValueError("fluent_like_monitor_nearest_count must be positive")
        _one_of(self.lbm_population_storage, ("f32", "f16", "bf16"), "lbm_population_storage")
        _one_of(self.lbm_collision, ("mrt", "regularized"), "lbm_collision")
        _one_of(self.mpm_transfer_order, ("atomic", "particle"), "mpm_transfer_order")
        if self.lattice_shape is not None:
            if len(self.lattice_shape) != 3 or min(self.lattice_shape) <= 0:
                raise ValueError("lattice_shape must hold three positive integers")
            if self.lattice_shape[0] != self.n_grid:
                raise ValueError("lattice_shape[0] must equal n_grid")

    def _validate_coupling_scalars(self):
        if self.dynamic_solid_threshold < 0.0:
            raise ValueError("dynamic_solid_threshold must be non-negative")
        for name in ("beta_lbm", "penalty_force_cap_lbm", "mb_reaction_scale", "mb_force_cap_norm",
                     "link_area_scale_min", "link_area_scale_max"):
            if getattr(self, name) <= 0.0:
                raise ValueError(f"{name} must be positive")
        if self.link_area_scale_min > self.link_area_scale_max:
            raise ValueError("link_area_scale_min must be <= link_area_scale_max")
        if self.output_interval <= 0:
            raise ValueError("output_interval must be positive")

    def _validate_fsi_exchange_mapping(self) -> None:
        """fsi_config.py:334-365."""
        if self.fsi_exchange_mode == "one_lbm_step_per_fsi_step":
            if self.lbm_substeps_per_fsi_step != 1:
                raise ValueError("lbm_substeps_per_fsi_step must be 1 when fsi_exchange_mode='one_lbm_step_per_fsi_step'")
            if self.lbm_dt_phys_override_s is not None:
                raise ValueError("lbm_dt_phys_override_s requires fsi_exchange_mode='lbm_subcycled_per_fsi_step'")
            return
        if self.coupling_mode != "moving_boundary":
            raise ValueError("lbm_subcycled_per_fsi_step currently requires coupling_mode='moving_boundary'")
        required = ("target_inlet_velocity_mps", "official_fsi_dt_s", "target_u_lbm_for_dimensional_mapping",
                    "lbm_dt_phys_override_s")
        missing = [name for name in required if getattr(self, name) is None]
        if missing:
            raise ValueError("subcycled FSI exchange requires: " + ", ".join(missing))
        expected_official_dt = float(self.lbm_substeps_per_fsi_step) * float(self.lbm_dt_phys_override_s)
        if not math.isclose(float(self.official_fsi_dt_s), expected_official_dt, rel_tol=1.0e-12, abs_tol=1.0e-15):
            raise ValueError("official_fsi_dt_s must equal lbm_substeps_per_fsi_step * lbm_dt_phys_override_s")
        if not math.isclose(float(self.target_u_lbm[0]), float(self.target_u_lbm_for_dimensional_mapping),
                            rel_tol=1.0e-12, abs_tol=1.0e-15):
            raise ValueError("target_u_lbm[0] must equal target_u_lbm_for_dimensional_mapping")
        dx_phys_m = float(self.physical_duct_length_m) / float(self.n_grid)
        mapped = float(self.target_u_lbm_for_dimensional_mapping) * dx_phys_m / float(self.lbm_dt_phys_override_s)
        if not math.isclose(mapped, float(self.target_inlet_velocity_mps), rel_tol=1.0e-12, abs_tol=1.0e-9):
            raise ValueError("dimensional mapping does not recover target_inlet_velocity_mps")

    def _validate_lbm_viscosity_mapping(self) -> None:
        """fsi_config.py:367-374."""
        mapping = self.lbm_viscosity_mapping_report()
        if not math.isfinite(float(mapping["tau"])) or float(mapping["tau"]) <= 0.5:
            raise ValueError("LBM viscosity mapping must produce finite tau > 0.5")
        if not math.isfinite(float(mapping["lbm_niu"])) or float(mapping["lbm_niu"]) <= 0.0:
            raise ValueError("LBM viscosity mapping must produce finite positive lbm_niu")
        if self.lbm_tau_stability_policy == "strict" and not bool(mapping["tau_margin_pass"]):
            raise ValueError("LBM tau stability margin failed strict policy")

    def lbm_viscosity_mapping_report(self) -> dict:
        """fsi_config.py:376-445: legacy niu = 0.1 (tau = 0.5333) or physical nu -> lattice nu."""
        common = {
            "lbm_viscosity_semantics": self.lbm_viscosity_semantics,
            "fluid_density_kg_m3": float(self.fluid_density_kg_m3),
            "fluid_density_used_for_lbm_rho0": False,
            "fluid_kinematic_viscosity_m2_s": float(self.fluid_kinematic_viscosity_m2_s),
            "target_reynolds_number": None if self.target_reynolds_number is None else float(self.target_reynolds_number),
            "lbm_min_tau_margin": float(self.lbm_min_tau_margin),
            "lbm_tau_stability_policy": self.lbm_tau_stability_policy,
            "mach_proxy": self._mach_proxy(),
            "physical_viscosity_validation_claim": False,
        }
        reynolds = self._reynolds_from_config()
        if self.lbm_viscosity_semantics == "legacy_external":
            legacy_niu = 0.1
            tau = tau_from_legacy_external_solver_parameter(legacy_niu)
            report = dict(common)
            report.update({
                "physical_mapping_used": False,
                "legacy_lbm_niu": legacy_niu,
                "nu_lbm": None,
                "lbm_niu": legacy_niu,
                "tau": tau,
                "tau_minus_half": float(tau) - 0.5,
                "tau_margin_pass": float(tau) - 0.5 >= float(self.lbm_min_tau_margin),
                "lbm_relaxation_semantics": LEGACY_EXTERNAL_SOLVER_RELAXATION_PARAMETER,
                "reynolds_from_config": reynolds,
                "target_reynolds_match": self._target_reynolds_match(reynolds),
                "physical_reynolds_direct_simulation_feasible_with_current_lbm": None,
            })
            return report
        if self.lbm_viscosity_semantics != "physical_nu_mapping":
            raise ValueError(f"unsupported lbm_viscosity_semantics: {self.lbm_viscosity_semantics}")
        dx_phys_m = float(self.physical_duct_length_m) / float(self.n_grid)
        dt_phys_s = float(self.lbm_dt_phys_override_s if self.lbm_dt_phys_override_s is not None
                          else self.mpm_substeps_per_lbm_step * self.mpm_dt)
        nu_lbm = float(self.fluid_kinematic_viscosity_m2_s) * dt_phys_s / (dx_phys_m * dx_phys_m)
        tau = tau_from_lattice_kinematic_viscosity(nu_lbm)
        tau_margin_pass = float(tau) - 0.5 >= float(self.lbm_min_tau_margin)
        report = dict(common)
        report.update({
            "physical_mapping_used": True,
            "dx_phys_m": dx_phys_m,
            "dt_phys_s": dt_phys_s,
            "nu_lbm": nu_lbm,
            "lbm_niu": nu_lbm,
            "tau": tau,
            "tau_minus_half": float(tau) - 0.5,
            "tau_margin_pass": tau_margin_pass,
            "lbm_relaxation_semantics": STANDARD_LATTICE_KINEMATIC_VISCOSITY,
            "reynolds_from_config": reynolds,
            "target_reynolds_match": self._target_reynolds_match(reynolds),
            "physical_reynolds_direct_simulation_feasible_with_current_lbm": bool(
                tau_margin_pass and common["mach_proxy"] <= 0.2),
        })
        return report

    def _mach_proxy(self) -> float:
        return float(math.sqrt(sum(float(v) * float(v) for v in self.target_u_lbm)) / math.sqrt(1.0 / 3.0))

    def _reynolds_from_config(self) -> Optional[float]:
        height = self._duct_height_for_reynolds()
        if self.target_inlet_velocity_mps is None or height is None:
            return None
        return float(self.target_inlet_velocity_mps) * height / float(self.fluid_kinematic_viscosity_m2_s)

    def _target_reynolds_match(self, reynolds: Optional[float]) -> Optional[bool]:
        if self.target_reynolds_number is None or reynolds is None:
            return None
        return math.isclose(reynolds, float(self.target_reynolds_number), rel_tol=1.0e-9, abs_tol=1.0e-9)

    def _duct_height_for_reynolds(self) -> Optional[float]:
        if self.geometry_config_path:
            path = Path(self.geometry_config_path)
            if path.is_file():
                with path.open("r", encoding="utf-8") as f:
                    data = json.load(f)
                dimensional = data.get("dimensional_reference") or {}
                if "duct_height_m" in dimensional:
                    return float(dimensional["duct_height_m"])
        return None

    @classmethod
    def from_json(cls, path):
        with open(path, "r", encoding="utf-8") as f:
            data = json.load(f)
        return cls(**data)

    def to_dict(self):
        data = asdict(self)
        for key in _TUPLE3_FIELDS + ("fluent_like_monitor_physical_point_m",):
            data[key] = list(data[key])
        if data.get("lattice_shape") is not None:
            data["lattice_shape"] = list(data["lattice_shape"])
        return data

    def make_unified_sim_config(self):
        mapping = self.lbm_viscosity_mapping_report()
        return UnifiedSimConfig(
            n_grid=self.n_grid,
            mpm_dt=self.mpm_dt,
            mpm_substeps_per_lbm_step=self.mpm_substeps_per_lbm_step,
            lbm_dt_phys_override_s=self.lbm_dt_phys_override_s,
            lbm_niu=float(mapping["lbm_niu"]),
            lbm_rho0=1.0,
            lbm_relaxation_semantics=str(mapping["lbm_relaxation_semantics"]),
            lbm_open_boundary_semantics=self.lbm_open_boundary_semantics,
            open_boundary_limiter_enabled=bool(self.open_boundary_limiter_enabled),
            open_boundary_rho_min=float(self.open_boundary_rho_min),
            open_boundary_rho_max=float(self.open_boundary_rho_max),
            open_boundary_u_max=float(self.open_boundary_u_max),
            open_boundary_noneq_cap=float(self.open_boundary_noneq_cap),
            open_boundary_population_floor=self.open_boundary_population_floor,
            lattice_shape=self.lattice_shape,
            lbm_population_storage=self.lbm_population_storage,
            lbm_collision=self.lbm_collision,
        )

    def make_mpm_control_overrides(self):
        return {
            "mpm_damping_application": self.mpm_damping_application,
            "mpm_planar_constraint_axis": self.mpm_planar_constraint_axis,
            "mpm_planar_constraint_mode": self.mpm_planar_constraint_mode,
            "mpm_velocity_damping": float(self.mpm_velocity_damping),
        }


def _validate_optional_positive(value, name: str) -> None:
    if value is not None and float(value) <= 0.0:
        raise ValueError(f"{name} must be positive when provided")


def _path_exists(path) -> bool:
    return Path(path).is_file()
