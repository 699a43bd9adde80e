"""LBMFluid3D: D3Q19 MRT lattice Boltzmann fluid on gfx950.

Drop-in for the reference's LBMFluid3D (sim/lbm/fluid.py:31-2111) on the FSI
path: same constructor, field names and methods used by FSIDriver3D and the
couplers.  The reference runs colission / streaming / Boundary_condition /
streaming3 as four passes with an f -> F copy; here one fused kernel
(`fsi_lbm_substep`) collides each fluid cell's populations and pushes them
into the other of two SoA buffers, so a substep reads and writes 19 floats
per fluid cell once.  Moments rho/v are materialised lazily (a `moments`
pass runs when a caller reads them), and the moving-wall hydro force of a
substep is gathered by the wall cells during the next substep.

Only the legacy `equilibrium_all_population_reset` open boundary with x-axis
boundary conditions is implemented (the FSIDriver3D default); other
open-boundary semantics, y/z boundary conditions and sparse storage raise.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from ... import _native
from ..._device import FieldView, ScalarView, bits_to_float, default_device, read_struct, stream_handle, struct_buffer
from .config import LBMConfig
from .relaxation_semantics import (
    LEGACY_EXTERNAL_SOLVER_RELAXATION_PARAMETER,
    STANDARD_LATTICE_KINEMATIC_VISCOSITY,
    tau_from_lattice_kinematic_viscosity,
    tau_from_legacy_external_solver_parameter,
)

EQUILIBRIUM_ALL_POPULATION_RESET = "equilibrium_all_population_reset"
# lbm_open_boundary_semantics with a GPU form (fluid.py:14-19, :538-570)
OPEN_BC_CODES = {
    EQUILIBRIUM_ALL_POPULATION_RESET: _native.OPEN_BC_EQUILIBRIUM,
    "regularized_velocity_pressure": _native.OPEN_BC_REGULARIZED,
    "regularized_velocity_pressure_limited": _native.OPEN_BC_REGULARIZED_LIMITED,
    "convective_pressure_outlet_experimental": _native.OPEN_BC_CONVECTIVE,
}
OB_COUNTERS = ("rho_clip", "velocity_clip", "noneq_clip", "population_floor", "reconstructed_population")
UNKNOWN_X_MIN_POPULATIONS = (1, 7, 9, 11, 13)
UNKNOWN_X_MAX_POPULATIONS = (2, 8, 10, 12, 14)

# D3Q19 tables in the reference's order (fluid.py:359-373, :252)
E = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
              [1, 1, 0], [-1, -1, 0], [1, -1, 0], [-1, 1, 0], [1, 0, 1], [-1, 0, -1], [1, 0, -1],
              [-1, 0, 1], [0, 1, 1], [0, -1, -1], [0, 1, -1], [0, -1, 1]], dtype=np.int32)
W = np.array([1.0 / 3.0] + [1.0 / 18.0] * 6 + [1.0 / 36.0] * 12)
LR = [0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17]


def mrt_rates(tau: float):
    """s_v = 1/tau, s_other = 8(2 - s_v)/(8 - s_v) (fluid.py:300-301)."""
    s_v = 1.0 / tau
    return s_v, 8.0 * (2.0 - s_v) / (8.0 - s_v)


class _PopulationView(FieldView):
    """F / f (post-stream populations, (nx,ny,nz,19)) decoded from the storage format."""

    def __init__(self, lbm, space_shape):
        super().__init__(lambda: lbm._F[lbm.cur], space_shape, 19, before_write=lbm._sync_hydro,
                         after_write=lbm._populations_written)
        self._lbm = lbm

    def _decoded(self):
        lbm = self._lbm
        raw = lbm._F[lbm.cur]
        if lbm.population_storage == "f32":
            return raw
        half = torch.float16 if lbm.population_storage == "f16" else torch.bfloat16
        return raw.view(half).to(torch.float32) + lbm._w_col

    def to_numpy(self):
        t = self._decoded().reshape((19,) + self.space_shape)
        return np.array(t.permute(1, 2, 3, 0).contiguous().cpu().numpy(), copy=True)

    def from_numpy(self, arr):
        a = np.asarray(arr, dtype=np.float32)
        if a.shape != self.shape:
            raise ValueError(f"expected array of shape {self.shape}, got {a.shape}")
        lbm = self._lbm
        lbm._sync_hydro()
        # rho and v are fields of their own in the reference: writing F leaves
        # them as they were, so settle any lazy moments of the old F first
        lbm._sync_moments()
        src = torch.from_numpy(np.ascontiguousarray(a)).to(lbm.device).reshape(-1, 19).t()
        if lbm.population_storage == "f32":
            lbm._F[lbm.cur].copy_(src)
        else:
            half = torch.float16 if lbm.population_storage == "f16" else torch.bfloat16
            lbm._F[lbm.cur].copy_((src - lbm._w_col).to(half).view(torch.int16))
        lbm._populations_written()


class LBMFluid3D:
    def __init__(self, config: LBMConfig, native: _native.Native | None = None, device=None):
        if config.sparse_storage:
            raise NotImplementedError(
                "Step 2 only supports dense storage. Enable sparse storage after FSI coupling is validated."
            )
        if config.open_boundary_semantics not in OPEN_BC_CODES:
            raise NotImplementedError(
                f"open_boundary_semantics={config.open_boundary_semantics!r} is not implemented on the GPU path; "
                f"implemented: {', '.join(OPEN_BC_CODES)}"
            )
        for side in ("bc_y_left", "bc_y_right", "bc_z_left", "bc_z_right"):
            if getattr(config, side) != 0:
                raise NotImplementedError(f"{side}={getattr(config, side)}: only x-axis boundary conditions are implemented")
        for side in ("bc_x_left", "bc_x_right"):
            if getattr(config, side) not in (0, 1, 2):
                raise ValueError(f"{side} must be 0 (periodic), 1 (pressure) or 2 (velocity)")
        self.config = config
        self.native = native if native is not None else _native.load_hip()
        self.device = torch.device(device) if device is not None else default_device()
        self.nx, self.ny, self.nz = int(config.nx), int(config.ny), int(config.nz)
        if min(self.nx, self.ny, self.nz) < 3:
            raise ValueError("LBMFluid3D needs at least 3 cells along every axis")
        self.N = self.nx * self.ny * self.nz
        self.fx, self.fy, self.fz = (float(v) for v in config.force)
        self.niu = config.niu
        self.rho0 = config.rho0
        self.relaxation_semantics = config.relaxation_semantics
        self.open_boundary_semantics = config.open_boundary_semantics
        self._open_bc = OPEN_BC_CODES[config.open_boundary_semantics]
        self.open_boundary_limiter_enabled = bool(config.open_boundary_limiter_enabled)
        self.open_boundary_population_floor_enabled = config.open_boundary_population_floor is not None
        self.sparse_storage = False
        self.bc_x_left, self.bc_x_right = int(config.bc_x_left), int(config.bc_x_right)
        self.e = E.copy()
        self.w = W.copy()
        self.LR = list(LR)

        N, dev = self.N, self.device
        f32 = dict(dtype=torch.float32, device=dev)
        self.population_storage = getattr(config, "population_storage", "f32")
        self.collision = getattr(config, "collision", "mrt")
        self._storage_code = {"f32": _native.STORE_F32, "f16": _native.STORE_F16_DEV,
                              "bf16": _native.STORE_BF16_DEV}[self.population_storage]
        pop_dtype = torch.float32 if self.population_storage == "f32" else torch.int16
        self._F = [torch.zeros((19, N), dtype=pop_dtype, device=dev), torch.zeros((19, N), dtype=pop_dtype, device=dev)]
        self._w_col = torch.tensor(W, dtype=torch.float32, device=dev)[:, None]
        self._rho = torch.full((N,), float(self.rho0), **f32)
        self._vel = torch.zeros((3, N), **f32)
        self._solid = torch.zeros((N,), dtype=torch.int8, device=dev)
        self._static_solid = torch.zeros((N,), dtype=torch.int8, device=dev)
        self._old_solid = torch.zeros((N,), dtype=torch.int8, device=dev)
        self._reinit = torch.zeros((N,), dtype=torch.int8, device=dev)
        self._cellinfo = torch.zeros((N,), dtype=torch.int32, device=dev)
        self._phi = torch.zeros((N,), **f32)
        self._smass = torch.zeros((N,), **f32)
        self._svel = torch.zeros((3, N), **f32)
        self._cell_force = torch.zeros((3, N), **f32)
        self._hydro = torch.zeros((3, N), **f32)
        self._link = None  # moving-wall link buffers, allocated on first use
        self._link_corr = None
        self._accum = None  # (N, 4) fp64: accumulated x, y, z and running max of the norm
        self._reduce_ws = None  # finalize records, 8 doubles per 256-cell workgroup
        self._counters = struct_buffer(_native.LbmCounters, dev)
        self._stats_ws = struct_buffer(_native.LbmStats, dev, _native.REDUCE_BLOCKS)
        self._stats_out = struct_buffer(_native.LbmStats, dev)

        self.cur = 0
        self._moments_fresh = True
        self._cellinfo_valid = False
        self._cell_force_active = False
        self._phi_version = 0
        self._rebuilt_phi_version = None
        self._pending = None  # link_c parity of a substep whose wall forces are not gathered yet
        self._pending_accumulate = False
        self._hydro_from_links = False
        self.mb_sample_count = 0
        self._last_link_parity = None  # link_c parity of the last moving-wall substep, while its links are current
        self._link_stats_ws = None
        # The moving-wall subcycle runs as one resident launch with the lattice
        # on chip (fsi_lbm_subcycle_resident) where that entry point covers the
        # lattice; False keeps one launch per substep.  Both give the same bits.
        self.resident_subcycle = True
        self._res_declined = False
        self._res_halo = None
        self._res_sync = None
        self.tau_f = None
        self.s_v = None
        self.s_other = None
        self._params = None
        self._state = _native.LbmState()
        self._refresh_state()
        self._make_views()

    # ------------------------------------------------------------------ setup
    def _make_views(self):
        sp = (self.nx, self.ny, self.nz)
        self.F = _PopulationView(self, sp)
        self.f = self.F
        self.rho = FieldView(lambda: self._rho, sp, before_read=self._sync_moments, after_write=self._moments_written)
        self.v = FieldView(lambda: self._vel, sp, 3, before_read=self._sync_moments, after_write=self._moments_written)
        self.solid = FieldView(lambda: self._solid, sp, after_write=self._flags_written, np_dtype=np.int8)
        self.static_solid = FieldView(lambda: self._static_solid, sp, after_write=self._flags_written, np_dtype=np.int8)
        self.old_solid = FieldView(lambda: self._old_solid, sp, np_dtype=np.int8)
        self.reinit_flag = FieldView(lambda: self._reinit, sp, np_dtype=np.int8)
        self.solid_phi = FieldView(lambda: self._phi, sp, after_write=self._phi_written)
        self.solid_mass = FieldView(lambda: self._smass, sp)
        self.solid_vel = FieldView(lambda: self._svel, sp, 3, after_write=self._phi_written)
        # rho / v of the last step include that step's cell force (streaming3 runs
        # at the end of step(), fluid.py:2099-2103): settle them before it changes
        self.cell_force = FieldView(lambda: self._cell_force, sp, 3, before_write=self._sync_moments,
                                    after_write=self._cell_force_written)
        self.hydro_force = FieldView(lambda: self._hydro, sp, 3, before_read=self._sync_hydro)
        self.hydro_force_accum = FieldView(self._accum_f32, sp, 3)
        self.bb_link_count = ScalarView(lambda: self.get_moving_boundary_stats()["bb_link_count"])
        self.bb_max_correction = ScalarView(lambda: self.get_moving_boundary_stats()["bb_max_correction"])
        self.mb_subcycle_force_sample_count = ScalarView(lambda: self.mb_sample_count)

    def _accum_f32(self):
        if self._accum is None:
            return torch.zeros((3, self.N), dtype=torch.float32, device=self.device)
        return self._accum[:, :3].t().contiguous().to(torch.float32)

    def _ensure_link_buffers(self):
        if self._link is None:
            dev = self.device
            self._link = [torch.zeros((18, self.N), dtype=torch.float32, device=dev) for _ in range(2)]
            self._link_corr = torch.zeros((18, self.N), dtype=torch.float32, device=dev)
            self._accum = torch.zeros((self.N, 4), dtype=torch.float64, device=dev)
            self._reduce_ws = torch.zeros((8 * ((self.N + 255) // 256),), dtype=torch.float64, device=dev)
            self._refresh_state()

    def _refresh_state(self):
        p = _native.ptr
        s = self._state
        s.F[0], s.F[1] = p(self._F[0]), p(self._F[1])
        s.rho, s.vel = p(self._rho), p(self._vel)
        s.solid, s.static_solid, s.old_solid = p(self._solid), p(self._static_solid), p(self._old_solid)
        s.cellinfo = p(self._cellinfo)
        s.solid_phi, s.solid_mass, s.solid_vel = p(self._phi), p(self._smass), p(self._svel)
        s.cell_force, s.hydro_force = p(self._cell_force), p(self._hydro)
        if self._link is not None:
            s.link_c[0], s.link_c[1] = p(self._link[0]), p(self._link[1])
            s.link_corr, s.hydro_accum, s.reduce_ws = p(self._link_corr), p(self._accum), p(self._reduce_ws)
        s.counters = p(self._counters)

    def _build_params(self):
        c = self.config
        prm = _native.LbmParams()
        prm.nx, prm.ny, prm.nz = self.nx, self.ny, self.nz
        prm.bc_x_left, prm.bc_x_right = self.bc_x_left, self.bc_x_right
        prm.s_v, prm.s_other, prm.rho0 = self.s_v, self.s_other, float(self.rho0)
        prm.force[:] = [self.fx, self.fy, self.fz]
        prm.rho_bc_x_left, prm.rho_bc_x_right = float(c.rho_bc_x_left), float(c.rho_bc_x_right)
        prm.vel_bc_x_left[:] = [float(v) for v in self.bc_vel_x_left]
        prm.vel_bc_x_right[:] = [float(v) for v in c.vel_bc_x_right]
        prm.dynamic_solid_threshold = 0.5
        prm.storage = self._storage_code
        prm.collision = _native.COLLIDE_REGULARIZED if self.collision == "regularized" else _native.COLLIDE_MRT
        prm.open_bc = self._open_bc
        prm.ob_limiter = int(self.open_boundary_limiter_enabled)
        prm.ob_floor_enabled = int(self.open_boundary_population_floor_enabled)
        prm.ob_rho_min, prm.ob_rho_max = float(c.open_boundary_rho_min), float(c.open_boundary_rho_max)
        prm.ob_u_max, prm.ob_noneq_cap = float(c.open_boundary_u_max), float(c.open_boundary_noneq_cap)
        prm.ob_floor = float(c.open_boundary_population_floor) if self.open_boundary_population_floor_enabled else -1.0e30
        self._params = prm

    def _stream(self):
        return stream_handle(self.device)

    def _call(self, name, *args):
        self.native.call(name, *args)

    # ------------------------------------------------------ reference surface
    def init_simulation(self):
        """Relaxation rates and initial equilibrium (fluid.py:285-317)."""
        if self.relaxation_semantics == STANDARD_LATTICE_KINEMATIC_VISCOSITY:
            self.tau_f = tau_from_lattice_kinematic_viscosity(self.niu)
        elif self.relaxation_semantics == LEGACY_EXTERNAL_SOLVER_RELAXATION_PARAMETER:
            self.tau_f = tau_from_legacy_external_solver_parameter(self.niu)
        else:
            raise ValueError(f"unsupported LBM relaxation semantics: {self.relaxation_semantics}")
        self.s_v, self.s_other = mrt_rates(self.tau_f)
        self.bc_vel_x_left = tuple(float(v) for v in self.config.vel_bc_x_left)
        self._build_params()
        self._call("fsi_lbm_init", ctypes.byref(self._params), ctypes.byref(self._state), self._stream())
        self._moments_fresh = True
        self._cellinfo_valid = False

    def init_geo(self, filename):
        """Read 0/1 solid flags written in Fortran order (fluid.py:347-353)."""
        data = np.loadtxt(filename)
        data[data > 0] = 1
        flags = np.reshape(data, (self.nx, self.ny, self.nz), order="F").astype(np.int8)
        self.set_solid_flags(flags)

    def set_solid_flags(self, flags):
        flags = np.asarray(flags, dtype=np.int8).reshape(self.nx, self.ny, self.nz)
        self.solid.from_numpy(flags)
        self.copy_solid_to_static()

    def copy_solid_to_static(self):
        self._static_solid.copy_(self._solid)
        self._old_solid.copy_(self._solid)
        self._cellinfo_valid = False

    def set_bc_vel_x0(self, vel):
        self.bc_x_left = 2
        self.bc_vel_x_left = tuple(float(v) for v in vel)
        if self._params is not None:
            self._build_params()

    def set_uniform_cell_force(self, fx: float, fy: float, fz: float):
        self._sync_moments()
        self._cell_force[0].fill_(float(fx))
        self._cell_force[1].fill_(float(fy))
        self._cell_force[2].fill_(float(fz))
        self._cell_force_written()

    def step(self):
        """colission + streaming1 + Boundary_condition + streaming3 (fluid.py:2099-2103)."""
        self._ensure_ready()
        self._launch_substep(_native.SUB_STORED_MOMENTS if self._moments_fresh else 0)

    def step_moving_bounceback(self):
        """Moving-wall variant (fluid.py:2105-2111); the link forces of this
        substep reach hydro_force when read or when accumulated."""
        self._ensure_ready()
        self._ensure_link_buffers()
        flags = _native.SUB_MOVING_WALL | _native.SUB_DIAG
        if self._moments_fresh:
            flags |= _native.SUB_STORED_MOMENTS
        parity = 0
        if self._pending is not None:
            parity = 1 - self._pending
            if self._pending_accumulate:
                flags |= _native.SUB_GATHER_PREV
        self._launch_substep(flags, parity)
        self._pending = parity
        self._pending_accumulate = False
        self._hydro_from_links = True
        self._last_link_parity = parity

    def update_dynamic_solid(self, threshold: float):
        """update_dynamic_solid + reinitialize_new_fluid_cells (fluid.py:1848-1877).
        Idempotent until solid_phi / solid_vel change, as in the reference."""
        threshold = float(threshold)
        key = (self._phi_version, threshold)
        if self._rebuilt_phi_version == key and self._cellinfo_valid:
            return
        self._sync_moments()
        # the last substep's link forces are fixed by the flags they were made under
        self._sync_hydro()
        self._params.dynamic_solid_threshold = threshold
        self._last_link_parity = None  # the flags no longer describe the last substep's links
        self._call("fsi_lbm_rebuild", ctypes.byref(self._params), ctypes.byref(self._state), 1, self.cur, self._stream())
        self._cellinfo_valid = True
        self._rebuilt_phi_version = key

    def reinitialize_new_fluid_cells(self):
        """Performed by update_dynamic_solid in the same pass."""

    def clear_moving_boundary_force_accumulator(self):
        """fluid.py:461-466."""
        self._ensure_link_buffers()
        self._call("fsi_lbm_mb_begin", ctypes.byref(self._params), ctypes.byref(self._state), self._stream())
        self.mb_sample_count = 0
        self._pending_accumulate = False

    def accumulate_moving_boundary_force_sample(self):
        """fluid.py:469-473 (the wall cells add the sample during the next substep or the finalize)."""
        if self._pending is None:
            raise RuntimeError("accumulate_moving_boundary_force_sample needs a preceding step_moving_bounceback")
        self._pending_accumulate = True
        self.mb_sample_count += 1

    def finalize_moving_boundary_force_accumulator(self):
        """hydro_force = accumulated mean (fluid.py:476-486)."""
        if self._pending is None or not self._pending_accumulate:
            raise RuntimeError("finalize_moving_boundary_force_accumulator needs an accumulated substep")
        self._call("fsi_lbm_mb_finalize", ctypes.byref(self._params), ctypes.byref(self._state), self.cur,
                   self._pending, self.mb_sample_count, 1, self._stream())
        self._pending = None
        self._pending_accumulate = False
        self._hydro_from_links = False

    def run_moving_boundary_subcycle(self, n_substeps: int, threshold: float, accumulate: bool = True):
        """The FSIDriver3D subcycle in one call (fsi_driver.py:533-550):
        clear accumulator, rebuild flags, n fused substeps, finalize, moments."""
        n = int(n_substeps)
        if n <= 0:
            raise ValueError("n_substeps must be positive")
        self._ensure_ready()
        self._ensure_link_buffers()
        if accumulate:
            self.clear_moving_boundary_force_accumulator()
        self.update_dynamic_solid(threshold)
        base = _native.SUB_MOVING_WALL | (_native.SUB_CELL_FORCE if self._cell_force_active else 0)
        parity = 0
        launches = n
        if self._run_resident(n, base, accumulate):
            self.cur = (self.cur + n) % 2
            self._moments_fresh = False
            parity = n % 2
            launches = 0
        for t in range(launches):
            flags = base
            if t == 0 and self._moments_fresh:
                flags |= _native.SUB_STORED_MOMENTS
            if t > 0 and accumulate:
                flags |= _native.SUB_GATHER_PREV
            if t == n - 1:
                flags |= _native.SUB_DIAG
            self._call("fsi_lbm_substep", ctypes.byref(self._params), ctypes.byref(self._state), self.cur, parity,
                       flags, self._stream())
            self.cur = 1 - self.cur
            self._open_bc_after_substep(flags)
            self._moments_fresh = False
            parity = 1 - parity
        last = 1 - parity
        self._last_link_parity = last
        if accumulate:
            self.mb_sample_count = n
            self._call("fsi_lbm_mb_finalize", ctypes.byref(self._params), ctypes.byref(self._state), self.cur, last,
                       n, 1, self._stream())
            self._pending = None
            self._hydro_from_links = False
        else:
            self._pending = last
            self._pending_accumulate = False
            self._hydro_from_links = True
        self._sync_moments()

    def _resident_shape_ok(self):
        return (self.resident_subcycle and not self._res_declined and self._storage_code == _native.STORE_F32
                and self._open_bc == _native.OPEN_BC_EQUILIBRIUM and self.nz == 64 and self.nx % 4 == 0
                and self.ny % 4 == 0)

    def _run_resident(self, n, base, accumulate):
        """The subcycle's n substeps as one fsi_lbm_subcycle_resident launch;
        False (nothing run) where it does not cover the lattice or device."""
        if not self._resident_shape_ok():
            return False
        if self._res_sync is None:
            self._res_halo = torch.empty((_native.res_halo_floats(self.N),), dtype=torch.float32, device=self.device)
            self._res_sync = torch.zeros((_native.res_sync_words(self.nx, self.ny),), dtype=torch.int32,
                                         device=self.device)
        flags = base | _native.SUB_DIAG
        if self._moments_fresh:
            flags |= _native.SUB_STORED_MOMENTS
        if accumulate:
            flags |= _native.SUB_GATHER_PREV
        ok = self.native.try_call("fsi_lbm_subcycle_resident", ctypes.byref(self._params), ctypes.byref(self._state),
                                  self.cur, int(n), flags, _native.ptr(self._res_halo), _native.ptr(self._res_sync),
                                  self._stream())
        if not ok:
            self._res_declined = True
        return ok

    def check_resident_waits(self):
        """Raise if a resident subcycle went on without a neighbour block
        (a wait that timed out: its results are invalid); the sync words are
        cleared so the next call starts clean."""
        if self._res_sync is None:
            return
        nblk = _native.res_blocks(self.nx, self.ny)
        timeouts = int(self._res_sync[64 * nblk].item())
        if timeouts:
            self._res_sync.zero_()
            raise RuntimeError(f"resident LBM subcycle: {timeouts} neighbour waits timed out; results are invalid")

    # --------------------------------------------------------- open boundaries
    def Boundary_condition(self):
        """Boundary_condition (fluid.py:538-570) on the current populations in
        place, with the stored rho, v as the neighbours' moments; clears the
        step counters first (fluid.py:539).  Substeps apply it themselves."""
        self._ensure_ready()
        # the neighbours' moments are those streaming3 gave the last step (fluid.py:1027-1030)
        self._sync_moments()
        self._call("fsi_lbm_open_bc", ctypes.byref(self._params), ctypes.byref(self._state), self.cur,
                   _native.SUB_STORED_MOMENTS, self._stream())

    def _ob_counters(self):
        k = read_struct(self._counters, _native.LbmCounters)
        return [int(v) for v in k.ob_step], [int(v) for v in k.ob_run]

    def _write_ob_counters(self, step=None, run=None):
        k = read_struct(self._counters, _native.LbmCounters)
        if step is not None:
            k.ob_step[:] = [int(v) for v in step]
        if run is not None:
            k.ob_run[:] = [int(v) for v in run]
        raw = np.frombuffer(bytes(k), dtype=np.uint8).copy()
        self._counters.copy_(torch.from_numpy(raw).to(self._counters.device))

    def clear_open_boundary_limiter_step_counters(self):
        """fluid.py:572-587 (the limiter counters)."""
        self._write_ob_counters(step=[0] * 5)

    def clear_open_boundary_limiter_run_counters(self):
        """fluid.py:590-637 (the limiter counters)."""
        self._write_ob_counters(step=[0] * 5, run=[0] * 5)

    def set_open_boundary_limiter_run_counters(self, rho_clip_count, velocity_clip_count, noneq_clip_count,
                                               population_floor_count, reconstructed_population_count):
        """fluid.py:639-693 (restart of the run counters)."""
        self._write_ob_counters(step=[0] * 5, run=[rho_clip_count, velocity_clip_count, noneq_clip_count,
                                                  population_floor_count, reconstructed_population_count])

    def get_open_boundary_limiter_stats(self):
        """The limiter part of fluid.py:804-925 (the flux-controller and
        mass-neutral entries belong to open-boundary modes without a GPU form)."""
        step, run = self._ob_counters()
        out = {}
        for name, a, b in zip(OB_COUNTERS, step, run):
            out[f"{name}_count_step"] = a
            out[f"{name}_count_run"] = b
        out["limiter_activation_count_step"] = sum(step[:4])
        out["limiter_activation_count"] = sum(run[:4])
        out["limiter_activation_denominator"] = run[4]
        out["limiter_activation_fraction"] = float(sum(run[:4]) / run[4] if run[4] else 0.0)
        return out

    # -------------------------------------------------------- lazy state sync
    def _ensure_ready(self):
        if self._params is None:
            raise RuntimeError("call init_simulation() first")
        if not self._cellinfo_valid:
            self._last_link_parity = None
            self._call("fsi_lbm_rebuild", ctypes.byref(self._params), ctypes.byref(self._state), 0, self.cur,
                       self._stream())
            self._cellinfo_valid = True

    def _launch_substep(self, flags, parity=0):
        self._last_link_parity = None
        if self._cell_force_active:
            flags |= _native.SUB_CELL_FORCE
        self._call("fsi_lbm_substep", ctypes.byref(self._params), ctypes.byref(self._state), self.cur, parity, flags,
                   self._stream())
        self.cur = 1 - self.cur
        self._open_bc_after_substep(flags)
        self._moments_fresh = False

    def _open_bc_after_substep(self, flags):
        """The open modes' Boundary_condition on the substep's output F[cur]
        (the legacy reset runs inside the substep)."""
        if self._open_bc != _native.OPEN_BC_EQUILIBRIUM:
            self._call("fsi_lbm_open_bc", ctypes.byref(self._params), ctypes.byref(self._state), self.cur,
                       flags & (_native.SUB_STORED_MOMENTS | _native.SUB_CELL_FORCE), self._stream())

    def _sync_moments(self):
        if self._moments_fresh or self._params is None:
            return
        if not self._cellinfo_valid:
            self._ensure_ready()
        flags = _native.SUB_CELL_FORCE if self._cell_force_active else 0
        self._call("fsi_lbm_moments", ctypes.byref(self._params), ctypes.byref(self._state), self.cur, flags,
                   self._stream())
        self._moments_fresh = True

    def _sync_hydro(self):
        if self._hydro_from_links and self._pending is not None:
            self._call("fsi_lbm_mb_finalize", ctypes.byref(self._params), ctypes.byref(self._state), self.cur,
                       self._pending, 1, 0, self._stream())
            self._hydro_from_links = False

    def drop_pending_links(self):
        """Forget the wall links of the last substep without gathering them
        (state loaded from a file replaces what they would give)."""
        self._pending = None
        self._pending_accumulate = False
        self._hydro_from_links = False

    def _populations_written(self):
        other = self._F[1 - self.cur]
        other.copy_(self._F[self.cur])

    def _moments_written(self):
        self._moments_fresh = True

    def _flags_written(self):
        self._cellinfo_valid = False

    def _phi_written(self):
        self._phi_version += 1

    def _cell_force_written(self):
        self._cell_force_active = True

    def mark_projection_updated(self):
        self._phi_version += 1

    # ------------------------------------------------------------ diagnostics
    def reduce_stats(self) -> dict:
        """All LBM reductions for one diagnostics row, as a dict of floats."""
        self._sync_moments()
        self._sync_hydro()
        self._call("fsi_lbm_reduce_stats", ctypes.byref(self._params), ctypes.byref(self._state),
                   _native.ptr(self._stats_ws), _native.ptr(self._stats_out), self._stream())
        st = read_struct(self._stats_out, _native.LbmStats)
        return {name: float(getattr(st, name)) for name in _native.LBM_STATS_FIELDS}

    def get_stats(self):
        """fluid.py:1931-1968."""
        r = self.reduce_stats()
        if r["fluid_count"] == 0:
            return {"max_v": 0.0, "rho_min": float("nan"), "rho_max": float("nan"), "mass_total": 0.0,
                    "force_norm_max": 0.0}
        return {"max_v": r["max_v"], "rho_min": r["rho_min"], "rho_max": r["rho_max"], "mass_total": r["mass_total"],
                "force_norm_max": r["force_norm_max"]}

    def _read_counters(self):
        return read_struct(self._counters, _native.LbmCounters)

    def get_moving_boundary_stats(self):
        """fluid.py:2067-2079 (last substep's link statistics)."""
        self._sync_hydro()
        k = self._read_counters()
        solid_force = tuple(float(v) for v in k.bb_net_solid_force)
        return {
            "bb_link_count": int(k.bb_link_count),
            "bb_max_correction": bits_to_float(k.bb_max_correction_bits),
            "bb_net_fluid_impulse": tuple(-v for v in solid_force),
            "bb_net_solid_force": solid_force,
        }

    def get_moving_boundary_directional_stats(self):
        """fluid.py:2080-2090: per-direction link statistics of the last
        moving-wall substep (zeros before one, or once a flag rebuild has
        followed it); the fluid impulse of a link is minus its solid force."""
        out = np.zeros((19, _native.LINK_STATS_VALUES), dtype=np.float64)
        if self._last_link_parity is not None and self._link is not None:
            if self._link_stats_ws is None:
                self._link_stats_ws = torch.zeros((_native.LINK_STATS_WS_DOUBLES + out.size,), dtype=torch.float64,
                                                  device=self.device)
            ws = self._link_stats_ws
            res = ws[_native.LINK_STATS_WS_DOUBLES:]
            self._call("fsi_lbm_link_stats", ctypes.byref(self._params), ctypes.byref(self._state), self.cur,
                       self._last_link_parity, _native.ptr(ws), _native.ptr(res), self._stream())
            out = res.cpu().numpy().reshape(19, _native.LINK_STATS_VALUES)
        # fp64 sums (the reference's tables are fp32 atomics): the net force of
        # get_moving_boundary_stats is summed in fp64 link by link too, so
        # the two stay consistent where the directions' forces cancel
        solid = out[:, 1:4].copy()
        return {
            "link_count_by_dir": out[:, 0].astype(np.int32),
            "fluid_impulse_by_dir": -solid,
            "solid_force_by_dir": solid,
            "correction_abs_sum_by_dir": out[:, 4].copy(),
            "correction_abs_max_by_dir": out[:, 5].copy(),
        }

    def get_moving_boundary_accumulation_stats(self):
        """fluid.py:2093-2097."""
        k = self._read_counters()
        return {
            "mb_subcycle_force_sample_count": int(k.mb_sample_count),
            "mb_subcycle_force_accum_norm_max": bits_to_float(k.accum_norm_max_bits),
            "mb_subcycle_force_mean_norm_max": bits_to_float(k.mean_norm_max_bits),
        }

    def get_max_v(self):
        return self.get_stats()["max_v"]
