"""CPU oracle of the reference FSI substep: a ctypes wrapper over
liboracle_fsi.so (fsi_oracle.cpp) plus the driver-level step order.

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and bench.py's
CPU-baseline leg.  Product code never imports it.

Arrays use the reference's Taichi layouts: lattice fields (nx,ny,nz[,19|3])
C-order, particle fields (P[,3[,3]]).  `dtype` selects the FP32 (parity) or
FP64 (precision reference) instantiation.  Parameters are derived from an
FSIDriverConfig with the reference's own formulas, restated here
(relaxation_semantics.py:14-19, fluid.py:295-304, solid.py:48-53,
projection.py:9-11, penalty.py:31, moving_boundary.py:30,
sim_config.py:63-71) so the product's parameter plumbing is checked too.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "liboracle_fsi.so"
_LIB = None
Q = 19
E = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
              [1, 1, 0], [-1, -1, 0], [1, -1, 0], [-1, 1, 0], [1, 0, 1], [-1, 0, -1], [1, 0, -1],
              [-1, 0, 1], [0, 1, 1], [0, -1, -1], [0, 1, -1], [0, -1, 1]], dtype=np.int64)


def lib():
    global _LIB
    if _LIB is None:
        if not LIB_PATH.is_file():
            raise RuntimeError(f"{LIB_PATH} not built (run __graft_entry__.build())")
        _LIB = ctypes.CDLL(str(LIB_PATH))
    return _LIB


def _structs(T):
    class LbmArgs(ctypes.Structure):
        _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                    ("s_v", T), ("s_other", T), ("rho0", T), ("fext", T * 3),
                    ("bc_x_left", ctypes.c_int), ("bc_x_right", ctypes.c_int),
                    ("rho_bc_x_left", T), ("rho_bc_x_right", T),
                    ("vel_bc_x_left", T * 3), ("vel_bc_x_right", T * 3)]

    class MpmArgs(ctypes.Structure):
        _fields_ = [("gnx", ctypes.c_int), ("gny", ctypes.c_int), ("gnz", ctypes.c_int),
                    ("n_particles", ctypes.c_int), ("bound", ctypes.c_int),
                    ("damping_enabled", ctypes.c_int), ("lock_z_enabled", ctypes.c_int),
                    ("dx", T), ("inv_dx", T), ("dt", T), ("mu", T), ("la", T), ("gravity", T * 3),
                    ("velocity_damping", T), ("domain_max", T * 3)]

    class CplArgs(ctypes.Structure):
        _fields_ = [(n, T) for n in ("inv_dx_norm", "cell_volume_norm", "vel_scale", "eps_mass", "beta_lbm",
                                     "penalty_phi_min", "force_cap_lbm", "force_density_scale",
                                     "reaction_scale", "mb_phi_min", "mb_force_cap_norm", "area_scale")]

    class MbStats(ctypes.Structure):
        _fields_ = [("link_count", ctypes.c_longlong), ("max_correction", T),
                    ("net_fluid_impulse", T * 3), ("net_solid_force", T * 3),
                    ("link_count_by_dir", ctypes.c_longlong * Q), ("fluid_impulse_by_dir", T * (3 * Q)),
                    ("solid_force_by_dir", T * (3 * Q)), ("corr_abs_sum_by_dir", T * Q),
                    ("corr_abs_max_by_dir", T * Q)]

    class ProjStats(ctypes.Structure):
        _fields_ = [("projected_mass", T), ("projected_volume_raw", T), ("projected_volume_clamped", T),
                    ("max_phi_raw", T), ("active_cell_count", ctypes.c_longlong)]

    return LbmArgs, MpmArgs, CplArgs, MbStats, ProjStats


def _ob_args(T):
    class ObArgs(ctypes.Structure):
        _fields_ = [("mode", ctypes.c_int), ("limiter", ctypes.c_int), ("floor_enabled", ctypes.c_int),
                    ("rho_min", T), ("rho_max", T), ("u_max", T), ("noneq_cap", T), ("floor", T)]
    return ObArgs


class ObCounts(ctypes.Structure):
    """rho clips, velocity clips, non-equilibrium clips, population floors,
    reconstructed populations of one Boundary_condition call."""
    _fields_ = [("c", ctypes.c_longlong * 5)]


OB_COUNT_NAMES = ("rho_clip_count", "velocity_clip_count", "noneq_clip_count", "population_floor_count",
                  "reconstructed_population_count")
# lbm_open_boundary_semantics -> the oracle's mode (fsi_oracle.cpp OB_*)
OB_MODES = {"equilibrium_all_population_reset": 0, "regularized_velocity_pressure": 1,
            "regularized_velocity_pressure_limited": 2, "convective_pressure_outlet_experimental": 3}
_OB_ARGS = {np.float32: _ob_args(ctypes.c_float), np.float64: _ob_args(ctypes.c_double)}


def ob_args(config, dtype=np.float32):
    """The open-boundary options of an FSIDriverConfig or LBMConfig (the
    reference's LBMFluid3D.__init__, fluid.py:49-81)."""
    sem = (getattr(config, "lbm_open_boundary_semantics", None) or getattr(config, "open_boundary_semantics", None)
           or "equilibrium_all_population_reset")
    ob = _OB_ARGS[dtype]()
    ob.mode = OB_MODES[sem]
    ob.limiter = int(bool(getattr(config, "open_boundary_limiter_enabled", False)))
    floor = getattr(config, "open_boundary_population_floor", None)
    ob.floor_enabled = int(floor is not None)
    ob.rho_min = getattr(config, "open_boundary_rho_min", 0.8)  # fsi_config.py:86-89 defaults
    ob.rho_max = getattr(config, "open_boundary_rho_max", 1.2)
    ob.u_max = getattr(config, "open_boundary_u_max", 0.1)
    ob.noneq_cap = getattr(config, "open_boundary_noneq_cap", 0.05)
    ob.floor = float(floor) if floor is not None else -1.0e30
    return ob


_STRUCTS = {np.float32: _structs(ctypes.c_float), np.float64: _structs(ctypes.c_double)}


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


# ---------------------------------------------------------------------------
# Parameters, restated from the reference
# ---------------------------------------------------------------------------
def lbm_tau(relaxation_semantics: str, niu: float) -> float:
    if relaxation_semantics == "standard_lattice_kinematic_viscosity":
        return 3.0 * float(niu) + 0.5  # relaxation_semantics.py:18-19
    return float(niu) / 3.0 + 0.5  # relaxation_semantics.py:14-15


def driver_parameters(config):
    """Everything the kernels need, from an FSIDriverConfig (duck-typed)."""
    n = config.n_grid
    shape = tuple(config.lattice_shape) if getattr(config, "lattice_shape", None) else (n, n, n)
    dx_norm = 1.0 / float(n)  # sim_config.py:63-65 (domain_length 1)
    dt_lbm = (float(config.lbm_dt_phys_override_s) if config.lbm_dt_phys_override_s is not None
              else config.mpm_substeps_per_lbm_step * config.mpm_dt)  # sim_config.py:67-71
    if config.lbm_viscosity_semantics == "legacy_external":
        niu, sem = 0.1, "legacy"
    else:
        dx_phys = config.physical_duct_length_m / float(n)
        niu, sem = config.fluid_kinematic_viscosity_m2_s * dt_lbm / (dx_phys * dx_phys), "standard_lattice_kinematic_viscosity"
    tau = lbm_tau(sem, niu)
    s_v = 1.0 / tau
    s_o = 8.0 * (2.0 - s_v) / (8.0 - s_v)  # fluid.py:300-301
    duct = config.lbm_boundary_condition_mode == "duct_velocity_inlet_pressure_outlet"
    bc = dict(bc_x_left=0, bc_x_right=0, rho_l=1.0, rho_r=1.0, vel_l=(0.0, 0.0, 0.0), vel_r=(0.0, 0.0, 0.0))
    if duct:
        tgt = tuple(float(v) for v in config.target_u_lbm)
        if config.velocity_inlet_side == "min":
            bc.update(bc_x_left=2, bc_x_right=1, vel_l=tgt)
        else:
            bc.update(bc_x_left=1, bc_x_right=2, vel_r=tgt)
    return dict(shape=shape, dx_norm=dx_norm, dt_lbm=dt_lbm, tau=tau, s_v=s_v, s_o=s_o, bc=bc)


class OracleFSI:
    """Reference-structured FSI state and step paths (fsi_driver.py:425-575)."""

    def __init__(self, config, solid_flags, particles, mpm_material, dtype=np.float32):
        self.config = config
        self.dtype = dtype
        T = dtype
        LbmArgs, MpmArgs, CplArgs, MbStats, ProjStats = _STRUCTS[dtype]
        self._MbStats, self._ProjStats = MbStats, ProjStats
        self.sfx = "f32" if dtype == np.float32 else "f64"
        prm = driver_parameters(config)
        self.prm = prm
        nx, ny, nz = prm["shape"]
        self.shape = (nx, ny, nz)
        N = nx * ny * nz
        la = LbmArgs()
        la.nx, la.ny, la.nz = nx, ny, nz
        la.s_v, la.s_other, la.rho0 = prm["s_v"], prm["s_o"], 1.0
        la.fext[:] = [0.0, 0.0, 0.0]
        la.bc_x_left, la.bc_x_right = prm["bc"]["bc_x_left"], prm["bc"]["bc_x_right"]
        la.rho_bc_x_left, la.rho_bc_x_right = prm["bc"]["rho_l"], prm["bc"]["rho_r"]
        la.vel_bc_x_left[:] = list(prm["bc"]["vel_l"])
        la.vel_bc_x_right[:] = list(prm["bc"]["vel_r"])
        if dtype == np.float32:  # the reference stores these as f32 fields
            for name in ("s_v", "s_other"):
                setattr(la, name, float(np.float32(getattr(la, name))))
        self.la = la
        self.ob = ob_args(config, dtype)
        self.ob_step = ObCounts()
        self.ob_run = np.zeros(5, np.int64)
        # lattice fields
        self.F = np.zeros(N * Q, T)
        self.f = np.zeros(N * Q, T)
        self.rho = np.zeros(N, T)
        self.v = np.zeros(N * 3, T)
        self.solid = np.array(solid_flags, dtype=np.int8).reshape(-1)  # a copy
        self.static_solid = self.solid.copy()
        self.old_solid = self.solid.copy()
        self.reinit = np.zeros(N, np.int8)
        self.phi = np.zeros(N, T)
        self.smass = np.zeros(N, T)
        self.svel = np.zeros(N * 3, T)
        self.cell_force = np.zeros(N * 3, T)
        self.hydro = np.zeros(N * 3, T)
        self.accum = np.zeros(N * 3, T)
        self.mb_sample_count = 0
        self.accum_norm_max = 0.0
        self.mean_norm_max = 0.0
        self.mbstats = MbStats()
        getattr(lib(), f"orc_lbm_init_{self.sfx}")(ctypes.byref(la), _p(self.F), _p(self.f), _p(self.rho), _p(self.v))
        # MPM
        P = int(particles["x"].shape[0])
        self.P = P
        E_mod, nu, p_rho = mpm_material
        mu = E_mod / (2.0 * (1.0 + nu))
        lam = E_mod * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))  # solid.py:48-53
        ma = MpmArgs()
        ma.gnx, ma.gny, ma.gnz = nx, ny, nz
        ma.n_particles, ma.bound = P, 3
        ma.damping_enabled, ma.lock_z_enabled = 0, 0
        ma.dx, ma.inv_dx, ma.dt, ma.mu, ma.la = prm["dx_norm"], 1.0 / prm["dx_norm"], config.mpm_dt, mu, lam
        ma.gravity[:] = list(config.gravity)
        ma.velocity_damping = 0.0
        ma.domain_max[:] = [nx * prm["dx_norm"], ny * prm["dx_norm"], nz * prm["dx_norm"]]
        self.ma = ma
        self.x = np.array(particles["x"], T).reshape(-1)  # a copy: the caller's cloud stays as given
        self.fixed_x = self.x.copy()
        self.pv = np.zeros(P * 3, T)
        v0 = np.asarray(config.initial_solid_velocity_norm, T)
        self.fixed_mask = np.ascontiguousarray(np.asarray(particles.get("fixed_base_mask", np.zeros(P)), np.int8))
        for d in range(3):
            self.pv[d::3] = np.where(self.fixed_mask != 0, T(0), v0[d])
        self.C = np.zeros(P * 9, T)
        self.Fp = np.tile(np.eye(3, dtype=T).reshape(-1), P)
        self.Jp = np.ones(P, T)
        self.mass = np.array(particles["mass"], T)
        self.vol0 = np.array(particles["vol0"], T)
        G = nx * ny * nz
        self.grid_v = np.zeros(G * 3, T)
        self.grid_m = np.zeros(G, T)
        self.grid_f = np.zeros(G * 3, T)
        # coupling
        ca = CplArgs()
        ca.inv_dx_norm = 1.0 / prm["dx_norm"]
        ca.cell_volume_norm = prm["dx_norm"] ** 3
        ca.vel_scale = prm["dt_lbm"] / prm["dx_norm"]
        ca.eps_mass = 1.0e-12
        ca.beta_lbm, ca.penalty_phi_min, ca.force_cap_lbm = config.beta_lbm, 1.0e-6, config.penalty_force_cap_lbm
        ca.force_density_scale = prm["dx_norm"] / prm["dt_lbm"] ** 2
        ca.reaction_scale = config.mb_reaction_scale if config.coupling_mode == "moving_boundary" else 1.0
        ca.mb_phi_min, ca.mb_force_cap_norm = 1.0e-6, config.mb_force_cap_norm
        ca.area_scale = 1.0
        self.ca = ca
        self.proj_stats = ProjStats()
        self.reaction_stats = np.zeros(9, T)
        self.penalty_stats = np.zeros(9, T)
        self.total_mpm_substeps = 0
        self.total_lbm_substeps = 0

    def _call(self, name, *args):
        getattr(lib(), f"orc_{name}_{self.sfx}")(*args)

    # ---- LBM passes (fluid.py)
    def colission(self):
        name = "colission_regularized" if getattr(self.config, "lbm_collision", "mrt") == "regularized" else "colission"
        self._call(name, ctypes.byref(self.la), _p(self.F), _p(self.f), _p(self.rho), _p(self.v),
                   _p(self.solid), _p(self.cell_force))

    def boundary_condition(self):
        """Boundary_condition (fluid.py:538-570): the legacy reset or an open-boundary reconstruction."""
        if self.ob.mode == 0:
            self._call("bc_legacy", ctypes.byref(self.la), _p(self.F), _p(self.v), _p(self.solid))
            return
        self._call("bc_open", ctypes.byref(self.la), ctypes.byref(self.ob), _p(self.F), _p(self.rho), _p(self.v),
                   _p(self.solid), ctypes.byref(self.ob_step))
        self.ob_run += np.asarray(self.ob_step.c[:], np.int64)

    def lbm_step(self):
        """step(): colission, streaming1, Boundary_condition, streaming3 (fluid.py:2099-2103)."""
        self.colission()
        self._call("streaming1", ctypes.byref(self.la), _p(self.f), _p(self.F), _p(self.solid))
        self.boundary_condition()
        self._call("streaming3", ctypes.byref(self.la), _p(self.F), _p(self.f), _p(self.rho), _p(self.v),
                   _p(self.solid), _p(self.cell_force))

    def lbm_step_moving_bounceback(self):
        """step_moving_bounceback (fluid.py:2105-2111)."""
        self.colission()
        self._call("streaming_mb", ctypes.byref(self.la), _p(self.f), _p(self.F), _p(self.solid), _p(self.rho),
                   _p(self.svel), _p(self.hydro), ctypes.byref(self.mbstats))
        self.boundary_condition()
        self._call("streaming3", ctypes.byref(self.la), _p(self.F), _p(self.f), _p(self.rho), _p(self.v),
                   _p(self.solid), _p(self.cell_force))

    def update_dynamic_solid(self, threshold):
        self._call("dynamic_solid", ctypes.byref(self.la), (ctypes.c_float if self.dtype == np.float32 else ctypes.c_double)(threshold), _p(self.phi), _p(self.static_solid),
                   _p(self.solid), _p(self.old_solid), _p(self.reinit), _p(self.svel), _p(self.F), _p(self.f),
                   _p(self.rho), _p(self.v))

    def clear_accumulator(self):
        """fluid.py:461-466."""
        self.accum[:] = 0
        self.mb_sample_count = 0
        self.accum_norm_max = 0.0
        self.mean_norm_max = 0.0

    def accumulate(self):
        """fluid.py:469-473."""
        self.mb_sample_count += 1
        self.accum += self.hydro
        norms = np.sqrt((self.accum.reshape(-1, 3).astype(self.dtype) ** 2).sum(axis=1))
        self.accum_norm_max = max(self.accum_norm_max, float(norms.max(initial=0.0)))

    def finalize_accumulator(self):
        """fluid.py:476-486."""
        self.mean_norm_max = 0.0
        if self.mb_sample_count > 0:
            inv = self.dtype(1.0) / self.dtype(self.mb_sample_count)
            mean = self.accum * inv
            self.hydro[:] = mean
            self.mean_norm_max = float(np.sqrt((mean.reshape(-1, 3) ** 2).sum(axis=1)).max(initial=0.0))

    # ---- MPM + coupling
    def project(self):
        self._call("project", ctypes.byref(self.la), ctypes.byref(self.ma), ctypes.byref(self.ca), _p(self.x),
                   _p(self.pv), _p(self.Jp), _p(self.mass), _p(self.vol0), _p(self.phi), _p(self.smass), _p(self.svel),
                   ctypes.byref(self.proj_stats))

    def penalty_force(self):
        self._call("penalty_force", ctypes.byref(self.la), ctypes.byref(self.ca), _p(self.solid), _p(self.phi),
                   _p(self.rho), _p(self.v), _p(self.svel), _p(self.cell_force), _p(self.hydro), _p(self.penalty_stats))

    def p2g(self):
        self._call("p2g", ctypes.byref(self.ma), _p(self.x), _p(self.pv), _p(self.C), _p(self.Fp), _p(self.Jp),
                   _p(self.mass), _p(self.vol0), _p(self.grid_v), _p(self.grid_m), _p(self.grid_f))

    def reaction(self, mode):
        self._call("reaction", ctypes.byref(self.la), ctypes.byref(self.ma), ctypes.byref(self.ca), mode, _p(self.x),
                   _p(self.Jp), _p(self.vol0), _p(self.hydro), _p(self.phi), _p(self.grid_f), _p(self.reaction_stats))

    def grid_update(self):
        self._call("grid_update", ctypes.byref(self.ma), _p(self.grid_v), _p(self.grid_m), _p(self.grid_f))

    def g2p(self):
        self._call("g2p", ctypes.byref(self.ma), _p(self.x), _p(self.pv), _p(self.C), _p(self.Fp), _p(self.Jp),
                   _p(self.fixed_mask), _p(self.fixed_x), _p(self.grid_v))

    def mpm_substeps(self, reaction_mode):
        for _ in range(self.config.mpm_substeps_per_lbm_step):
            self.p2g()  # includes clear_grid
            if reaction_mode:
                self.reaction(reaction_mode)
            self.grid_update()
            self.g2p()
            self.total_mpm_substeps += 1

    def update_area_scale(self):
        """LinkAreaMovingBoundaryCoupler3D.update_area_scale_from_lbm
        (link_area.py:84-109 with link_area_accounting.py:42-60, :86-110): the
        area-weighted x solid force of the last substep's links over their
        plain x solid force, clipped; stored as the reaction's float scale."""
        c = self.config
        length = np.linalg.norm(E.astype(np.float64), axis=1)
        moving = length > 0
        w = np.zeros(Q)
        w[moving] = {"uniform": np.ones(Q), "inverse_length": 1.0 / np.where(moving, length, 1.0),
                     "length": length}[c.link_area_policy][moving]
        by_dir = np.asarray(self.mbstats.solid_force_by_dir[:], dtype=self.dtype).reshape(Q, 3)
        weighted_x = np.sum(by_dir * w[:, None], axis=0, dtype=np.float64)[0]
        net_x = float(self.mbstats.net_solid_force[0])
        raw = abs(float(weighted_x)) / (abs(net_x) + 1.0e-12)
        scale = float(np.clip(raw, c.link_area_scale_min, c.link_area_scale_max))
        if not np.isfinite(scale):
            scale = c.link_area_scale_min
        self.ca.area_scale = float(self.dtype(scale))
        self.raw_area_scale = raw
        return scale

    def _mb_reaction_substeps(self):
        if getattr(self.config, "reaction_transfer_mode", "engineering") == "link_area_experimental":
            self.update_area_scale()
        self.mpm_substeps(2)

    # ---- driver step paths
    def step_once(self):
        c = self.config
        if c.fsi_exchange_mode == "lbm_subcycled_per_fsi_step":
            self.project()
            self.clear_accumulator()
            for _ in range(c.lbm_substeps_per_fsi_step):
                self.update_dynamic_solid(c.dynamic_solid_threshold)
                self.lbm_step_moving_bounceback()
                self.accumulate()
                self.total_lbm_substeps += 1
            self.finalize_accumulator()
            self._mb_reaction_substeps()
            return
        if c.coupling_mode == "none":
            self.project()
            self.lbm_step()
            self.mpm_substeps(0)
        elif c.coupling_mode == "penalty":
            self.project()
            self.penalty_force()
            self.lbm_step()
            self.mpm_substeps(1)
        else:
            self.project()
            self.update_dynamic_solid(c.dynamic_solid_threshold)
            self.lbm_step_moving_bounceback()
            self._mb_reaction_substeps()
        self.total_lbm_substeps += 1

    # ---- views in reference shapes
    def fields(self):
        nx, ny, nz = self.shape
        sh = (nx, ny, nz)
        return {
            "F": self.F.reshape(sh + (Q,)), "rho": self.rho.reshape(sh), "v": self.v.reshape(sh + (3,)),
            "solid": self.solid.reshape(sh), "solid_phi": self.phi.reshape(sh), "solid_mass": self.smass.reshape(sh),
            "solid_vel": self.svel.reshape(sh + (3,)), "cell_force": self.cell_force.reshape(sh + (3,)),
            "hydro_force": self.hydro.reshape(sh + (3,)),
            "x": self.x.reshape(-1, 3), "pv": self.pv.reshape(-1, 3), "C": self.C.reshape(-1, 3, 3),
            "Fp": self.Fp.reshape(-1, 3, 3), "Jp": self.Jp.copy(),
            "grid_v": self.grid_v.reshape(sh + (3,)), "grid_m": self.grid_m.reshape(sh),
            "grid_f": self.grid_f.reshape(sh + (3,)),
        }


def inv_M(dtype=np.float64):
    out = np.zeros(Q * Q, dtype)
    getattr(lib(), "orc_inv_M_" + ("f32" if dtype == np.float32 else "f64"))(_p(out))
    return out.reshape(Q, Q)


def polar(Fm, dtype=np.float64):
    F = np.ascontiguousarray(np.asarray(Fm, dtype).reshape(9))
    R = np.zeros(9, dtype)
    getattr(lib(), "orc_polar_" + ("f32" if dtype == np.float32 else "f64"))(_p(F), _p(R))
    return R.reshape(3, 3)
