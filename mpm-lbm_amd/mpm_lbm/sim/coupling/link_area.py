"""LinkAreaMovingBoundaryCoupler3D (reference: sim/coupling/link_area.py:1-221).

The experimental link-area reaction transfer: the engineering
moving-boundary reaction with one more global factor, the area scale.  Once
per FSI step, before the MPM substeps, the scale is the ratio of the
area-weighted x solid force of the last substep's links to their plain x
solid force, clipped to [area_scale_min, area_scale_max] (link_area.py:84-109).
The reaction kernels read it from the coupling parameters (area_scale).
"""
from __future__ import annotations

import numpy as np

from .link_area_accounting import VALID_AREA_POLICIES, LinkAreaMomentumAccounting3D
from .moving_boundary import MovingBoundaryFSICoupler3D


class LinkAreaMovingBoundaryCoupler3D(MovingBoundaryFSICoupler3D):
    def __init__(self, sim_config, area_policy: str = "inverse_length", reaction_scale: float = 1.0,
                 force_cap_norm: float = 1.0e-5, phi_min: float = 1.0e-6, area_scale_min: float = 0.25,
                 area_scale_max: float = 2.0, native=None):
        if area_policy not in VALID_AREA_POLICIES:
            raise ValueError(f"area_policy must be one of {VALID_AREA_POLICIES}")
        if area_scale_min <= 0.0:
            raise ValueError("area_scale_min must be positive")
        if area_scale_max <= 0.0:
            raise ValueError("area_scale_max must be positive")
        if area_scale_min > area_scale_max:
            raise ValueError("area_scale_min must be <= area_scale_max")
        super().__init__(sim_config, reaction_scale=reaction_scale, force_cap_norm=force_cap_norm, phi_min=phi_min,
                         native=native)
        self.area_policy = area_policy
        self.area_scale_min = float(area_scale_min)
        self.area_scale_max = float(area_scale_max)
        self.set_area_summary(1.0, 1.0, 0.0, 0.0, 0.0, 0.0)

    def set_area_summary(self, raw_area_scale, area_scale, area_proxy_total, wx, wy, wz):
        """link_area.py:118-133 (the scale kept as float32, as the reference's field)."""
        self.raw_area_scale = float(np.float32(raw_area_scale))
        self.params.area_scale = float(np.float32(area_scale))
        self.area_proxy_total = float(np.float32(area_proxy_total))
        self.area_weighted_hydro_sum = tuple(float(np.float32(v)) for v in (wx, wy, wz))

    def set_area_scale(self, value):
        self.params.area_scale = float(np.float32(value))

    @property
    def area_scale(self):
        return float(self.params.area_scale)

    def update_area_scale_from_lbm(self, lbm):
        """link_area.py:84-109."""
        summary = LinkAreaMomentumAccounting3D.summarize_link_accounting(lbm, policy=self.area_policy)
        numerator = abs(float(summary["area_weighted_solid_force_x"]))
        denominator = abs(float(summary["bb_net_solid_force_x"])) + 1.0e-12
        raw = numerator / denominator
        scale = float(np.clip(raw, self.area_scale_min, self.area_scale_max))
        if not np.isfinite(scale):
            scale = self.area_scale_min
        self.set_area_summary(raw, scale, summary["area_proxy_total"], summary["area_weighted_solid_force_x"],
                              summary["area_weighted_solid_force_y"], summary["area_weighted_solid_force_z"])
        out = dict(summary)
        out.update({"area_policy": self.area_policy, "raw_area_scale": float(raw), "area_scale": scale})
        return out

    def add_link_area_reaction_to_mpm_grid(self, solid, lbm):
        """link_area.py:139-194: the engineering reaction times area_scale."""
        self.add_moving_boundary_reaction_to_mpm_grid(solid, lbm)

    def get_stats(self):
        """link_area.py:196-221."""
        out = super().get_stats()
        out.update({
            "area_policy": self.area_policy,
            "area_scale_min": self.area_scale_min,
            "area_scale_max": self.area_scale_max,
            "raw_area_scale": self.raw_area_scale,
            "area_scale": self.area_scale,
            "area_proxy_total": self.area_proxy_total,
            "area_weighted_hydro_sum": self.area_weighted_hydro_sum,
        })
        return out
