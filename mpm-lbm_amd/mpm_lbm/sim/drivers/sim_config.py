"""UnifiedSimConfig (reference: sim/drivers/sim_config.py:9-99).

Extension: `lattice_shape` = (nx, ny, nz) for non-cubic lattices (the
reference is cubic only, nx = ny = nz = n_grid).  Cells stay cubic with
dx_norm = domain_length / nx; the domain becomes [0,1] x [0,ny/nx] x [0,nz/nx].
"""
from dataclasses import dataclass
from typing import Optional, Tuple

from ..lbm.config import LBMConfig
from ..lbm.relaxation_semantics import LEGACY_EXTERNAL_SOLVER_RELAXATION_PARAMETER
from ..mpm.config import MPMConfig


@dataclass(frozen=True)
class UnifiedSimConfig:
    n_grid: int = 32
    domain_length: float = 1.0
    mpm_dt: float = 4.0e-4
    mpm_substeps_per_lbm_step: int = 10
    lbm_niu: float = 0.1
    lbm_rho0: float = 1.0
    lbm_relaxation_semantics: str = LEGACY_EXTERNAL_SOLVER_RELAXATION_PARAMETER
    lbm_open_boundary_semantics: str = "equilibrium_all_population_reset"
    open_boundary_limiter_enabled: bool = False
    open_boundary_rho_min: float = 0.8
    open_boundary_rho_max: float = 1.2
    open_boundary_u_max: float = 0.1
    open_boundary_noneq_cap: float = 0.05
    open_boundary_population_floor: Optional[float] = None
    lbm_dt_phys_override_s: Optional[float] = None
    lattice_shape: Optional[Tuple[int, int, int]] = None
    lbm_population_storage: str = "f32"
    lbm_collision: str = "mrt"

    def __post_init__(self):
        checks = (
            (self.n_grid > 0, "n_grid must be positive"),
            (self.domain_length > 0.0, "domain_length must be positive"),
            (self.mpm_dt > 0.0, "mpm_dt must be positive"),
            (self.mpm_substeps_per_lbm_step > 0, "mpm_substeps_per_lbm_step must be positive"),
            (self.lbm_niu > 0.0, "lbm_niu must be positive"),
            (self.lbm_rho0 > 0.0, "lbm_rho0 must be positive"),
            (self.open_boundary_rho_min > 0.0, "open_boundary_rho_min must be positive"),
            (self.open_boundary_rho_max > self.open_boundary_rho_min,
             "open_boundary_rho_min must be less than open_boundary_rho_max"),
            (self.open_boundary_u_max > 0.0, "open_boundary_u_max must be positive"),
            (self.open_boundary_noneq_cap > 0.0, "open_boundary_noneq_cap must be positive"),
            (self.lbm_dt_phys_override_s is None or self.lbm_dt_phys_override_s > 0.0,
             "lbm_dt_phys_override_s must be positive when provided"),
        )
        for ok, message in checks:
            if not ok:
                raise ValueError(message)
        if self.lattice_shape is not None:
            shape = tuple(int(v) for v in self.lattice_shape)
            if len(shape) != 3 or min(shape) <= 0:
                raise ValueError("lattice_shape must hold three positive integers")
            if shape[0] != self.n_grid:
                raise ValueError("lattice_shape[0] must equal n_grid")
            object.__setattr__(self, "lattice_shape", shape)

    @property
    def nx(self) -> int:
        return self.n_grid

    @property
    def ny(self) -> int:
        return self.n_grid if self.lattice_shape is None else self.lattice_shape[1]

    @property
    def nz(self) -> int:
        return self.n_grid if self.lattice_shape is None else self.lattice_shape[2]

    @property
    def dx_norm(self) -> float:
        return self.domain_length / float(self.n_grid)

    @property
    def domain_extent(self) -> Tuple[float, float, float]:
        return (self.nx * self.dx_norm, self.ny * self.dx_norm, self.nz * self.dx_norm)

    @property
    def lbm_dt_phys(self) -> float:
        if self.lbm_dt_phys_override_s is not None:
            return float(self.lbm_dt_phys_override_s)
        return self.mpm_substeps_per_lbm_step * self.mpm_dt

    def make_lbm_config(self, **overrides) -> LBMConfig:
        values = {
            "nx": self.nx,
            "ny": self.ny,
            "nz": self.nz,
            "niu": self.lbm_niu,
            "rho0": self.lbm_rho0,
            "relaxation_semantics": self.lbm_relaxation_semantics,
            "open_boundary_semantics": self.lbm_open_boundary_semantics,
            "open_boundary_limiter_enabled": self.open_boundary_limiter_enabled,
            "open_boundary_rho_min": self.open_boundary_rho_min,
            "open_boundary_rho_max": self.open_boundary_rho_max,
            "open_boundary_u_max": self.open_boundary_u_max,
            "open_boundary_noneq_cap": self.open_boundary_noneq_cap,
            "open_boundary_population_floor": self.open_boundary_population_floor,
            "population_storage": self.lbm_population_storage,
            "collision": self.lbm_collision,
        }
        values.update(overrides)
        return LBMConfig(**values)

    def make_mpm_config(self, **overrides) -> MPMConfig:
        values = {"n_grid": self.n_grid, "dx": self.dx_norm, "dt": self.mpm_dt,
                  "grid_shape": (self.nx, self.ny, self.nz)}
        values.update(overrides)
        return MPMConfig(**values)
