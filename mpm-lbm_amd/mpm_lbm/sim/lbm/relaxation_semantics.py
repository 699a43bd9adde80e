"""LBM relaxation-parameter conventions (reference: sim/lbm/relaxation_semantics.py:9-19)."""

LEGACY_EXTERNAL_SOLVER_RELAXATION_PARAMETER = "legacy_external_solver_relaxation_parameter"
STANDARD_LATTICE_KINEMATIC_VISCOSITY = "standard_lattice_kinematic_viscosity"
DEFAULT_TAU_CONVENTION = LEGACY_EXTERNAL_SOLVER_RELAXATION_PARAMETER


def tau_from_legacy_external_solver_parameter(niu: float) -> float:
    """tau = niu / 3 + 1/2, the legacy `niu` convention of the upstream solver."""
    return float(niu) / 3.0 + 0.5


def tau_from_lattice_kinematic_viscosity(nu_lbm: float) -> float:
    """tau = 3 nu + 1/2 for lattice kinematic viscosity nu (c_s^2 = 1/3)."""
    return 3.0 * float(nu_lbm) + 0.5
