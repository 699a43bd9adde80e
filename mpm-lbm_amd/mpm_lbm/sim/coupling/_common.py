"""Shared plumbing of the coupling operators (coupling constants + device counters)."""
from __future__ import annotations

import ctypes

from ... import _native
from ..._device import read_struct, stream_handle, struct_buffer


class CouplingBase:
    def __init__(self, sim_config, native=None):
        self.n_grid = sim_config.n_grid
        self.dx_norm = sim_config.dx_norm
        self.inv_dx_norm = 1.0 / sim_config.dx_norm
        self.lbm_dt_phys = sim_config.lbm_dt_phys
        self._native = native
        self._counters_buf = None
        self.params = _native.CouplingParams()
        self.params.inv_dx_norm = self.inv_dx_norm
        self.params.cell_volume_norm = sim_config.dx_norm ** 3
        self.params.vel_scale_norm_to_lbm = sim_config.lbm_dt_phys / sim_config.dx_norm
        self.params.eps_mass = 1.0e-12
        self.params.force_density_scale = sim_config.dx_norm / (sim_config.lbm_dt_phys ** 2)

    def _bind(self, lbm):
        if self._native is None:
            self._native = lbm.native
        if self._counters_buf is None or self._counters_buf.device != lbm.device:
            self._counters_buf = struct_buffer(_native.CouplingCounters, lbm.device)
        return stream_handle(lbm.device)

    def _counters(self):
        if self._counters_buf is None:
            return _native.CouplingCounters()
        return read_struct(self._counters_buf, _native.CouplingCounters)

    def _call(self, name, *args):
        self._native.call(name, *args)

    @staticmethod
    def _refs(*objs):
        return [ctypes.byref(o) for o in objs]
