"""PenaltyFSICoupler3D (reference: sim/coupling/penalty.py:5-152).

build_penalty_force: on fluid cells with phi > phi_min,
F = beta phi rho (u_s - u) capped at force_cap_lbm, hydro = -F.
add_lbm_reaction_to_mpm_grid: per particle, gather hydro over 27 cells and
scatter w * hydro * (dx / dt^2) * vol0 * max(J, 0) * reaction_scale onto
grid_f_ext.
"""
from __future__ import annotations

from ..._device import bits_to_float
from ... import _native
from ._common import CouplingBase


class PenaltyFSICoupler3D(CouplingBase):
    def __init__(self, sim_config, beta_lbm: float = 1.0e-3, phi_min: float = 1.0e-6, force_cap_lbm: float = 1.0e-4,
                 reaction_scale: float = 1.0, native=None):
        if beta_lbm <= 0.0:
            raise ValueError("beta_lbm must be positive")
        if phi_min < 0.0:
            raise ValueError("phi_min must be non-negative")
        if force_cap_lbm <= 0.0:
            raise ValueError("force_cap_lbm must be positive")
        if reaction_scale <= 0.0:
            raise ValueError("reaction_scale must be positive")
        super().__init__(sim_config, native)
        self.beta_lbm = float(beta_lbm)
        self.phi_min = float(phi_min)
        self.force_cap_lbm = float(force_cap_lbm)
        self.reaction_scale = float(reaction_scale)
        self.force_density_scale_lbm_to_norm = self.dx_norm / (self.lbm_dt_phys ** 2)
        self.params.beta_lbm = self.beta_lbm
        self.params.penalty_phi_min = self.phi_min
        self.params.force_cap_lbm = self.force_cap_lbm
        self.params.reaction_scale = self.reaction_scale
        self._react_counters = None

    def clear_force_fields(self, lbm):
        # v of the last step carries its cell force (streaming3 at the end of
        # step(), fluid.py:2099-2103) and build_penalty_force reads that v
        lbm._sync_moments()
        lbm._cell_force.zero_()
        lbm._hydro.zero_()
        lbm._hydro_from_links = False

    def build_penalty_force(self, lbm):
        stream = self._bind(lbm)
        lbm._sync_moments()
        self._call("fsi_penalty_force", *self._refs(lbm._params, lbm._state, self.params),
                   _native.ptr(self._counters_buf), stream)
        lbm._cell_force_active = True
        lbm._hydro_from_links = False

    def add_lbm_reaction_to_mpm_grid(self, solid, lbm):
        stream = self._bind(lbm)
        self._call("fsi_mpm_reaction", *self._refs(lbm._params, lbm._state, solid._params, solid._state, self.params),
                   _native.ptr(self._counters_buf), 1, stream)

    def get_stats(self):
        k = self._counters()
        return {
            "n_grid": self.n_grid,
            "dx_norm": self.dx_norm,
            "inv_dx_norm": self.inv_dx_norm,
            "lbm_dt_phys": self.lbm_dt_phys,
            "beta_lbm": self.beta_lbm,
            "phi_min": self.phi_min,
            "force_cap_lbm": self.force_cap_lbm,
            "reaction_scale": self.reaction_scale,
            "force_density_scale_lbm_to_norm": self.force_density_scale_lbm_to_norm,
            "active_force_cell_count": int(k.active_force_cell_count),
            "max_cell_force_norm": bits_to_float(k.max_cell_force_bits),
            "max_hydro_force_norm": bits_to_float(k.max_hydro_force_bits),
            "net_cell_force": tuple(float(v) for v in k.net_cell_force),
            "net_hydro_force": tuple(float(v) for v in k.net_hydro_force),
            "max_reaction_grid_force_norm": bits_to_float(k.max_grid_reaction_bits),
            "net_reaction_grid_force": tuple(float(v) for v in k.net_grid_reaction),
        }
