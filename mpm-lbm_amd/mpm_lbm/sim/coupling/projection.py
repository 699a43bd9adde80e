"""MPMToLBMProjector3D (reference: sim/coupling/projection.py:5-111).

One `fsi_project` call clears the projection fields, scatters particle
volume fraction, mass and mass-weighted lattice velocity with quadratic
B-spline weights (27 nodes, inside the lattice only), then normalises
velocity by mass (mass > 1e-12) and clamps phi to [0, 1].
"""
from __future__ import annotations

import ctypes

from ..._device import BinBuffers, bits_to_float
from ... import _native
from ._common import CouplingBase


class MPMToLBMProjector3D(CouplingBase):
    def __init__(self, sim_config, native=None):
        super().__init__(sim_config, native)
        self.cell_volume_norm = sim_config.dx_norm ** 3
        self.vel_scale_norm_to_lbm = sim_config.lbm_dt_phys / sim_config.dx_norm
        self.eps_mass = 1.0e-12
        self.transfer_order = "atomic"  # or "particle": binned, sums in particle order
        self._bins = None

    def project(self, solid, lbm, clear=True):
        if not clear:
            raise NotImplementedError("project(clear=False) is not supported")
        stream = self._bind(lbm)
        refs = self._refs(lbm._params, lbm._state, solid._params, solid._state, self.params)
        if self.transfer_order == "particle":
            # each cell's sums in particle order (deterministic)
            shape = (lbm.nx, lbm.ny, lbm.nz)
            if self._bins is None or self._bins.shape != shape or self._bins.key.numel() < solid.n_particles:
                self._bins = BinBuffers(solid.n_particles, shape, lbm.device)
            self._call("fsi_project_ordered", *refs, _native.ptr(self._counters_buf), ctypes.byref(self._bins.struct),
                       stream)
        else:
            self._call("fsi_project", *refs, _native.ptr(self._counters_buf), stream)
        lbm.mark_projection_updated()

    def get_stats(self):
        k = self._counters()
        return {
            "projected_mass": float(k.projected_mass),
            "projected_volume_raw": float(k.projected_volume_raw),
            "projected_volume_clamped": float(k.projected_volume_clamped),
            "max_phi_raw": bits_to_float(k.max_phi_raw_bits),
            "active_cell_count": int(k.active_cell_count),
            "vel_scale_norm_to_lbm": float(self.vel_scale_norm_to_lbm),
            "cell_volume_norm": float(self.cell_volume_norm),
        }
