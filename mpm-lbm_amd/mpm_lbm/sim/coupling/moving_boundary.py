"""MovingBoundaryFSICoupler3D (reference: sim/coupling/moving_boundary.py:5-130).

Per particle: gather hydro_force and phi over 27 cells; if the sampled phi
exceeds phi_min, F_p = hydro * (dx / dt^2) * vol0 * max(J, 0) * scale,
capped at force_cap_norm, and w * F_p is scattered onto grid_f_ext.
"""
from __future__ import annotations

from ..._device import bits_to_float
from ... import _native
from ._common import CouplingBase


class MovingBoundaryFSICoupler3D(CouplingBase):
    def __init__(self, sim_config, reaction_scale: float = 1.0, force_cap_norm: float = 1.0e-2,
                 phi_min: float = 1.0e-6, native=None):
        if reaction_scale <= 0.0:
            raise ValueError("reaction_scale must be positive")
        if force_cap_norm <= 0.0:
            raise ValueError("force_cap_norm must be positive")
        if phi_min < 0.0:
            raise ValueError("phi_min must be non-negative")
        super().__init__(sim_config, native)
        self.reaction_scale = float(reaction_scale)
        self.force_cap_norm = float(force_cap_norm)
        self.phi_min = float(phi_min)
        self.force_density_scale_lbm_to_norm = self.dx_norm / (self.lbm_dt_phys ** 2)
        self.params.reaction_scale = self.reaction_scale
        self.params.mb_phi_min = self.phi_min
        self.params.mb_force_cap_norm = self.force_cap_norm

    def clear_reaction_diagnostics(self):
        """The reaction entry point clears its counters itself."""

    def add_moving_boundary_reaction_to_mpm_grid(self, solid, lbm):
        stream = self._bind(lbm)
        lbm._sync_hydro()
        self._call("fsi_mpm_reaction", *self._refs(lbm._params, lbm._state, solid._params, solid._state, self.params),
                   _native.ptr(self._counters_buf), 2, stream)

    def get_stats(self):
        k = self._counters()
        return {
            "n_grid": self.n_grid,
            "dx_norm": self.dx_norm,
            "inv_dx_norm": self.inv_dx_norm,
            "lbm_dt_phys": self.lbm_dt_phys,
            "reaction_scale": self.reaction_scale,
            "force_cap_norm": self.force_cap_norm,
            "phi_min": self.phi_min,
            "force_density_scale_lbm_to_norm": self.force_density_scale_lbm_to_norm,
            "active_reaction_particle_count": int(k.active_reaction_particle_count),
            "max_particle_reaction_norm": bits_to_float(k.max_particle_reaction_bits),
            "max_grid_reaction_norm": bits_to_float(k.max_grid_reaction_bits),
            "net_particle_reaction_force": tuple(float(v) for v in k.net_particle_reaction),
            "net_grid_reaction_force": tuple(float(v) for v in k.net_grid_reaction),
        }
