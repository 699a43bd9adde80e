"""MPMSolid3D: MLS-MPM / APIC solid with fixed-corotated elasticity on gfx950.

Drop-in for the reference's MPMSolid3D (sim/mpm/solid.py:10-419): the same
constructor checks, particle / grid fields and methods, with p2g, grid_update
and g2p as hand-written HIP kernels (`fsi_mpm_*`).  Particle state lives in
SoA device buffers; the grid spans `config.grid_shape` nodes (n_grid^3 in the
reference).
"""
from __future__ import annotations

import ctypes
import math
import os

import numpy as np
import torch

from ... import _native
from ..._device import BinBuffers, FieldView, default_device, read_struct, stream_handle, struct_buffer
from .config import MPMConfig


class MPMSolid3D:
    def __init__(self, config: MPMConfig, n_particles: int, native: _native.Native | None = None, device=None,
                 domain_extent=(1.0, 1.0, 1.0)):
        if n_particles <= 0:
            raise ValueError("n_particles must be positive")
        valid_planar_modes = ("disabled", "lock_z")
        valid_damping_modes = ("disabled", "particle_velocity_post_g2p")
        if config.mpm_planar_constraint_mode not in valid_planar_modes:
            raise ValueError(f"mpm_planar_constraint_mode must be one of {valid_planar_modes}")
        if config.mpm_planar_constraint_axis != "z":
            raise ValueError("mpm_planar_constraint_axis currently supports only z")
        if config.mpm_velocity_damping < 0.0:
            raise ValueError("mpm_velocity_damping must be non-negative")
        if config.mpm_damping_application not in valid_damping_modes:
            raise ValueError(f"mpm_damping_application must be one of {valid_damping_modes}")
        if config.mpm_damping_application == "disabled" and config.mpm_velocity_damping != 0.0:
            raise ValueError("mpm_velocity_damping must be 0 when damping is disabled")
        self.config = config
        self.native = native if native is not None else _native.load_hip()
        self.device = torch.device(device) if device is not None else default_device()
        self.dim = 3
        self.n_grid = config.n_grid
        self.grid_shape = tuple(config.grid_shape)
        self.dx = config.dx
        self.inv_dx = 1.0 / config.dx
        self.dt = config.dt
        self.n_particles = int(n_particles)
        self.bound = config.bound
        self.use_apic = config.use_apic
        # order of the grid-node sums of substep_coupled: "atomic" (LDS tiles,
        # float atomics; fastest) or "particle" (binned, in particle order;
        # deterministic)
        self.transfer_order = "atomic"
        # subcycle_coupled in the "atomic" order runs as one resident launch
        # (fsi_mpm_subcycle_coupled) when the device holds it and there are at
        # least resident_min_substeps substeps: with fewer, the launch's fixed
        # costs (its clear of the state's grid by the particle workgroups
        # alone, two barriers) outweigh what it saves (measured: 1 substep per
        # FSI step is slower resident, 10 faster)
        self.resident_subcycle = True
        self.resident_min_substeps = 4
        self._res_declined = False
        self._res_scratch = None
        self._res_sync = None
        self.gravity_x, self.gravity_y, self.gravity_z = config.gravity
        self.box_min_x, self.box_min_y, self.box_min_z = config.box_min
        self.box_max_x, self.box_max_y, self.box_max_z = config.box_max
        self.mpm_planar_lock_z_enabled = 1 if config.mpm_planar_constraint_mode == "lock_z" else 0
        self.mpm_velocity_damping = float(config.mpm_velocity_damping)
        self.mpm_damping_enabled = 1 if config.mpm_damping_application == "particle_velocity_post_g2p" else 0
        self.young_modulus = config.young_modulus
        self.poisson_ratio = config.poisson_ratio
        # Lame parameters (solid.py:48-53)
        self.mu = config.young_modulus / (2.0 * (1.0 + config.poisson_ratio))
        self.la = config.young_modulus * config.poisson_ratio / (
            (1.0 + config.poisson_ratio) * (1.0 - 2.0 * config.poisson_ratio))
        box_volume = ((self.box_max_x - self.box_min_x) * (self.box_max_y - self.box_min_y)
                      * (self.box_max_z - self.box_min_z))
        self.p_vol0 = box_volume / float(n_particles)
        self.p_mass0 = self.p_vol0 * config.p_rho
        self.particles_per_axis = max(1, math.ceil(n_particles ** (1.0 / 3.0)))
        self.domain_extent = tuple(float(v) for v in domain_extent)

        P, dev = self.n_particles, self.device
        f32 = dict(dtype=torch.float32, device=dev)
        self._x = torch.zeros((3, P), **f32)
        self._v = torch.zeros((3, P), **f32)
        self._C = torch.zeros((9, P), **f32)
        self._F = torch.zeros((9, P), **f32)
        self._F[0] = 1.0
        self._F[4] = 1.0
        self._F[8] = 1.0
        self._Jp = torch.ones((P,), **f32)
        self._mass = torch.full((P,), float(self.p_mass0), **f32)
        self._vol0 = torch.full((P,), float(self.p_vol0), **f32)
        self._fixed_mask = torch.zeros((P,), dtype=torch.int8, device=dev)
        self._fixed_x = torch.zeros((3, P), **f32)
        G = int(np.prod(self.grid_shape))
        self._grid_v = torch.zeros((3, G), **f32)
        self._grid_m = torch.zeros((G,), **f32)
        self._grid_f = torch.zeros((3, G), **f32)
        self._stats_ws = struct_buffer(_native.MpmStats, dev, _native.REDUCE_BLOCKS)
        self._stats_out = struct_buffer(_native.MpmStats, dev)

        prm = _native.MpmParams()
        prm.gnx, prm.gny, prm.gnz = self.grid_shape
        prm.n_particles = P
        prm.bound = int(self.bound)
        prm.damping_enabled = self.mpm_damping_enabled
        prm.lock_z_enabled = self.mpm_planar_lock_z_enabled
        prm.dx, prm.inv_dx, prm.dt = float(self.dx), float(self.inv_dx), float(self.dt)
        prm.mu, prm.la = float(self.mu), float(self.la)
        prm.gravity[:] = [float(g) for g in config.gravity]
        prm.velocity_damping = self.mpm_velocity_damping
        prm.domain_max[:] = list(self.domain_extent)
        self._params = prm
        self._state = _native.MpmState()
        p = _native.ptr
        s = self._state
        s.x, s.v, s.C, s.F, s.Jp = p(self._x), p(self._v), p(self._C), p(self._F), p(self._Jp)
        s.mass, s.vol0, s.fixed_mask, s.fixed_x = p(self._mass), p(self._vol0), p(self._fixed_mask), p(self._fixed_x)
        s.grid_v, s.grid_m, s.grid_f = p(self._grid_v), p(self._grid_m), p(self._grid_f)

        sp = (P,)
        gs = self.grid_shape
        self.x = FieldView(lambda: self._x, sp, 3)
        self.v = FieldView(lambda: self._v, sp, 3)
        self.C = FieldView(lambda: self._C, sp, (3, 3))
        self.F = FieldView(lambda: self._F, sp, (3, 3))
        self.Jp = FieldView(lambda: self._Jp, sp)
        self.mass = FieldView(lambda: self._mass, sp)
        self.vol0 = FieldView(lambda: self._vol0, sp)
        self.fixed_mask = FieldView(lambda: self._fixed_mask, sp, np_dtype=np.int8)
        self.fixed_x = FieldView(lambda: self._fixed_x, sp, 3)
        self.grid_v = FieldView(lambda: self._grid_v, gs, 3)
        self.grid_m = FieldView(lambda: self._grid_m, gs)
        self.grid_f_ext = FieldView(lambda: self._grid_f, gs, 3)

    def _stream(self):
        return stream_handle(self.device)

    def _call(self, name):
        self.native.call(name, ctypes.byref(self._params), ctypes.byref(self._state), self._stream())

    # ------------------------------------------------------------- set-up
    def init_box(self):
        """Regular lattice inside [box_min, box_max] (solid.py:127-159), in f32 as the reference."""
        ppa = self.particles_per_axis
        p = np.arange(self.n_particles)
        idx = np.stack([p % ppa, (p // ppa) % ppa, (p // (ppa * ppa)) % ppa], axis=1).astype(np.float32)
        denom = np.float32(max(1, ppa - 1) + 1)
        local = (idx + np.float32(0.5)) / denom
        lo = np.array([self.box_min_x, self.box_min_y, self.box_min_z], dtype=np.float32)
        span = np.array([self.box_max_x - self.box_min_x, self.box_max_y - self.box_min_y,
                         self.box_max_z - self.box_min_z], dtype=np.float32)
        x = lo + span * local
        self.x.from_numpy(x)
        self.fixed_x.from_numpy(x)
        self._v.zero_()
        self.reset_deformation_state()
        self._mass.fill_(float(self.p_mass0))
        self._vol0.fill_(float(self.p_vol0))
        self._fixed_mask.zero_()

    def reset_deformation_state(self):
        self._C.zero_()
        self._F.zero_()
        self._F[0] = 1.0
        self._F[4] = 1.0
        self._F[8] = 1.0
        self._Jp.fill_(1.0)

    def init_from_numpy(self, x_np, vol0_np, mass_np, v_np=None, fixed_mask_np=None):
        """solid.py:168-210 (same checks and messages)."""
        x_arr = np.asarray(x_np, dtype=np.float32)
        vol_arr = np.asarray(vol0_np, dtype=np.float32)
        mass_arr = np.asarray(mass_np, dtype=np.float32)
        P = self.n_particles
        if x_arr.shape != (P, 3):
            raise ValueError(f"x_np must have shape ({P}, 3)")
        if vol_arr.shape != (P,):
            raise ValueError(f"vol0_np must have shape ({P},)")
        if mass_arr.shape != (P,):
            raise ValueError(f"mass_np must have shape ({P},)")
        if v_np is None:
            v_arr = np.zeros_like(x_arr)
        else:
            v_arr = np.asarray(v_np, dtype=np.float32)
            if v_arr.shape != (P, 3):
                raise ValueError(f"v_np must have shape ({P}, 3)")
        if fixed_mask_np is None:
            fixed_arr = np.zeros((P,), dtype=np.int8)
        else:
            fixed_arr = np.asarray(fixed_mask_np, dtype=np.int8)
            if fixed_arr.shape != (P,):
                raise ValueError(f"fixed_mask_np must have shape ({P},)")
        if not np.all(np.isfinite(x_arr)):
            raise ValueError("x_np must be finite")
        if not np.all(np.isfinite(v_arr)):
            raise ValueError("v_np must be finite")
        if not np.all((fixed_arr == 0) | (fixed_arr == 1)):
            raise ValueError("fixed_mask_np must contain only 0/1 values")
        if not np.all(np.isfinite(vol_arr)) or np.any(vol_arr <= 0.0):
            raise ValueError("vol0_np must be finite and positive")
        if not np.all(np.isfinite(mass_arr)) or np.any(mass_arr <= 0.0):
            raise ValueError("mass_np must be finite and positive")
        self.x.from_numpy(x_arr)
        self.v.from_numpy(v_arr)
        self.vol0.from_numpy(vol_arr)
        self.mass.from_numpy(mass_arr)
        self.fixed_x.from_numpy(x_arr)
        self.fixed_mask.from_numpy(fixed_arr)
        self.reset_deformation_state()
        self.apply_fixed_particle_constraints()

    def set_uniform_velocity(self, vx: float, vy: float, vz: float):
        """solid.py:324-330: fixed particles get zero velocity."""
        moving = (self._fixed_mask == 0).to(torch.float32)
        self._v[0] = float(vx) * moving
        self._v[1] = float(vy) * moving
        self._v[2] = float(vz) * moving

    def apply_fixed_particle_constraints(self):
        """solid.py:332-339."""
        fixed = self._fixed_mask != 0
        if not bool(fixed.any()):
            return
        keep = (~fixed).to(torch.float32)
        for d in range(3):
            self._x[d] = torch.where(fixed, self._fixed_x[d], self._x[d])
            self._v[d] *= keep
        self._C *= keep
        ident = torch.tensor([1, 0, 0, 0, 1, 0, 0, 0, 1], dtype=torch.float32, device=self.device)[:, None]
        self._F.copy_(torch.where(fixed[None, :], ident.expand_as(self._F), self._F))
        self._Jp.copy_(torch.where(fixed, torch.ones_like(self._Jp), self._Jp))

    # -------------------------------------------------------------- substep
    def clear_grid(self):
        self._call("fsi_mpm_clear_grid")

    def p2g(self):
        self._call("fsi_mpm_p2g")

    def grid_update(self):
        self._call("fsi_mpm_grid_update")

    def g2p(self):
        self._call("fsi_mpm_g2p")

    def substep(self):
        self.clear_grid()
        self.p2g()
        self.grid_update()
        self.g2p()

    def substep_coupled(self, lbm=None, coupler=None, mode: int = 0):
        """One substep with the coupling reaction of `coupler` (mode 1: penalty,
        2: moving boundary; 0: none) in one fused device sequence
        (fsi_mpm_substep_coupled): clear_grid, p2g + reaction, grid_update, g2p."""
        if getattr(self, "_coupled_ws", None) is None:
            self._coupled_ws = torch.zeros(_native.coupled_workspace_doubles(self.n_particles), dtype=torch.float64,
                                           device=self.device)
        if self.transfer_order == "particle":
            self._substep_ordered(lbm, coupler, mode)
            return
        if mode:
            stream = coupler._bind(lbm)
            lbm._sync_hydro()
            self.native.call("fsi_mpm_substep_coupled", ctypes.byref(lbm._params), ctypes.byref(lbm._state),
                             ctypes.byref(self._params), ctypes.byref(self._state), ctypes.byref(coupler.params),
                             _native.ptr(coupler._counters_buf), int(mode), _native.ptr(self._coupled_ws), stream)
        else:
            self.native.call("fsi_mpm_substep_coupled", None, None, ctypes.byref(self._params),
                             ctypes.byref(self._state), None, None, 0, None, self._stream())

    def subcycle_coupled(self, n: int, lbm=None, coupler=None, mode: int = 0):
        """n calls of substep_coupled (the MPM substeps of one FSI step, the
        lattice unchanged in between).  In the "atomic" order they run as one
        launch that keeps the particles on chip (fsi_mpm_subcycle_coupled)
        when the device holds every particle workgroup at once and n is at
        least resident_min_substeps; otherwise, or when resident_subcycle is
        off, as n substeps."""
        if (self.resident_subcycle and not self._res_declined and self.transfer_order == "atomic"
                and n >= max(1, self.resident_min_substeps)):
            if self._run_resident(int(n), lbm, coupler, mode):
                return
            self._res_declined = True
        for _ in range(int(n)):
            self.substep_coupled(lbm, coupler, mode)

    def _run_resident(self, n, lbm, coupler, mode) -> bool:
        if getattr(self, "_coupled_ws", None) is None:
            self._coupled_ws = torch.zeros(_native.coupled_workspace_doubles(self.n_particles), dtype=torch.float64,
                                           device=self.device)
        if self._res_sync is None:
            G = int(np.prod(self.grid_shape))
            self._res_scratch = torch.zeros(_native.mpm_res_scratch_floats(G), dtype=torch.float32, device=self.device)
            self._res_sync = torch.zeros(_native.MPM_RES_SYNC_WORDS, dtype=torch.int32, device=self.device)
        scratch, sync = _native.ptr(self._res_scratch), _native.ptr(self._res_sync)
        if mode:
            stream = coupler._bind(lbm)
            lbm._sync_hydro()
            return self.native.try_call("fsi_mpm_subcycle_coupled", ctypes.byref(lbm._params),
                                        ctypes.byref(lbm._state), ctypes.byref(self._params),
                                        ctypes.byref(self._state), ctypes.byref(coupler.params),
                                        _native.ptr(coupler._counters_buf), int(mode), n,
                                        _native.ptr(self._coupled_ws), scratch, sync, stream)
        return self.native.try_call("fsi_mpm_subcycle_coupled", None, None, ctypes.byref(self._params),
                                    ctypes.byref(self._state), None, None, 0, n, None, scratch, sync, self._stream())

    def check_resident_waits(self):
        """Raise if a resident MPM subcycle went on past a barrier without
        every workgroup (a wait that timed out: its results are invalid); the
        sync words and the grid scratch are cleared so the next call starts
        clean."""
        if self._res_sync is None:
            return
        timeouts = int(self._res_sync[_native.MPM_RES_TIMEOUT_WORD].item())
        if timeouts:
            self._res_sync.zero_()
            self._res_scratch.zero_()
            raise RuntimeError(f"resident MPM subcycle: {timeouts} barrier waits timed out; results are invalid")

    def _substep_ordered(self, lbm, coupler, mode):
        """substep_coupled with every grid-node sum in particle order
        (fsi_mpm_substep_ordered): deterministic, and in the order of the
        reference's kernels run sequentially."""
        if getattr(self, "_bins", None) is None:
            self._bins = BinBuffers(self.n_particles, self.grid_shape, self.device)
            self._ordered_terms = torch.zeros(_native.ORDERED_TERMS_FLOATS * self.n_particles, dtype=torch.float32,
                                              device=self.device)
        if mode:
            stream = coupler._bind(lbm)
            lbm._sync_hydro()
            self.native.call("fsi_mpm_substep_ordered", ctypes.byref(lbm._params), ctypes.byref(lbm._state),
                             ctypes.byref(self._params), ctypes.byref(self._state), ctypes.byref(coupler.params),
                             _native.ptr(coupler._counters_buf), int(mode), ctypes.byref(self._bins.struct),
                             _native.ptr(self._ordered_terms), _native.ptr(self._coupled_ws), stream)
        else:
            self.native.call("fsi_mpm_substep_ordered", None, None, ctypes.byref(self._params),
                             ctypes.byref(self._state), None, None, 0, ctypes.byref(self._bins.struct),
                             _native.ptr(self._ordered_terms), None, self._stream())

    # ----------------------------------------------------------- diagnostics
    def reduce_stats(self) -> dict:
        self.native.call("fsi_mpm_reduce_stats", ctypes.byref(self._params), ctypes.byref(self._state),
                         _native.ptr(self._stats_ws), _native.ptr(self._stats_out), self._stream())
        st = read_struct(self._stats_out, _native.MpmStats)
        return {name: float(getattr(st, name)) for name in _native.MPM_STATS_FIELDS}

    def get_stats(self):
        """solid.py:359-375."""
        r = self.reduce_stats()
        min_x = (r["min_x0"], r["min_x1"], r["min_x2"])
        max_x = (r["max_x0"], r["max_x1"], r["max_x2"])
        return {
            "min_x": tuple(float(np.float32(v)) for v in min_x),
            "max_x": tuple(float(np.float32(v)) for v in max_x),
            "min_y": float(np.float32(min_x[1])),
            "max_y": float(np.float32(max_x[1])),
            "max_speed": r["max_speed"],
            "min_J": r["min_J"],
            "max_J": r["max_J"],
            "total_mass": r["total_mass"],
            "mass_weighted_velocity": tuple(
                (r[k] / r["total_mass"]) if r["total_mass"] > 0 else 0.0 for k in ("mv_x", "mv_y", "mv_z")),
        }

    def get_fixed_particle_stats(self):
        """solid.py:377-401."""
        mask = self.fixed_mask.to_numpy().astype(bool)
        count = int(np.count_nonzero(mask))
        if count == 0:
            return {"fixed_base_constraint_applied": False, "fixed_base_particle_count": 0,
                    "fixed_base_max_displacement_norm": 0.0, "fixed_base_max_velocity_norm": 0.0}
        x, fx, v = self.x.to_numpy(), self.fixed_x.to_numpy(), self.v.to_numpy()
        return {
            "fixed_base_constraint_applied": True,
            "fixed_base_particle_count": count,
            "fixed_base_max_displacement_norm": float(np.max(np.linalg.norm(x[mask] - fx[mask], axis=1))),
            "fixed_base_max_velocity_norm": float(np.max(np.linalg.norm(v[mask], axis=1))),
        }

    def export_particles(self, out_dir: str, prefix: str = "particles"):
        os.makedirs(out_dir, exist_ok=True)
        np.save(os.path.join(out_dir, f"{prefix}_x.npy"), self.x.to_numpy())
        np.save(os.path.join(out_dir, f"{prefix}_v.npy"), self.v.to_numpy())
        np.save(os.path.join(out_dir, f"{prefix}_F.npy"), self.F.to_numpy())
        np.save(os.path.join(out_dir, f"{prefix}_J.npy"), self.Jp.to_numpy())
