"""Deterministic particle sampling (reference: sim/geometry/sampler.py:74-175).

Candidates are cell centres of a `resolution`^3 lattice over the domain; the
inside ones are thinned to n_particles by evenly spaced indices
(np.linspace(0, n-1, count) truncated to int).  Box and duct_flap_proxy
geometries are supported; the others belong to campaign code outside this
package and raise NotImplementedError.
"""
import math

import numpy as np

from .config import GeometryConfig
from .duct_flap_proxy import duct_flap_proxy_component_masks, duct_flap_proxy_sampling_stats, inside_duct_flap_proxy


class GeometrySampler3D:
    def __init__(self, config: GeometryConfig):
        self.config = config
        self.domain_min = np.asarray(config.domain_min, dtype=np.float64)
        self.domain_max = np.asarray(config.domain_max, dtype=np.float64)
        self.domain_span = self.domain_max - self.domain_min
        self.domain_volume = float(np.prod(self.domain_span))
        if config.geometry_type not in ("box", "duct_flap_proxy"):
            raise NotImplementedError(f"geometry_type={config.geometry_type!r} sampling is not part of this package")

    def inside(self, points: np.ndarray) -> np.ndarray:
        pts = np.asarray(points, dtype=np.float64)
        if self.config.geometry_type == "box":
            lo = np.asarray(self.config.box_min)
            hi = np.asarray(self.config.box_max)
            return np.all((pts >= lo) & (pts <= hi), axis=1)
        return inside_duct_flap_proxy(pts, self.config)

    def component_masks(self, points: np.ndarray) -> dict:
        if self.config.geometry_type == "duct_flap_proxy":
            return duct_flap_proxy_component_masks(points, self.config)
        return {self.config.geometry_type: self.inside(points)}

    def _candidate_points(self, resolution: int) -> np.ndarray:
        coords = self.domain_min + (np.arange(resolution, dtype=np.float64) + 0.5)[:, None] * (
            self.domain_span / float(resolution))
        x, y, z = np.meshgrid(coords[:, 0], coords[:, 1], coords[:, 2], indexing="ij")
        return np.column_stack((x.ravel(), y.ravel(), z.ravel()))

    def _find_inside_candidates(self):
        min_fraction = 0.015 if self.config.geometry_type != "box" else 0.025
        resolution = max(int(self.config.particles_per_axis_hint),
                         int(math.ceil((self.config.n_particles / min_fraction) ** (1.0 / 3.0))), 16)
        inside = None
        for _ in range(8):
            candidates = self._candidate_points(resolution)
            inside = self.inside(candidates)
            if int(np.count_nonzero(inside)) >= self.config.n_particles:
                return candidates, inside, resolution
            resolution = int(math.ceil(resolution * 1.35))
        raise RuntimeError(
            f"geometry_type={self.config.geometry_type} produced only {int(np.count_nonzero(inside))} "
            f"inside candidates for {self.config.n_particles} particles")

    def sample_particles(self) -> dict:
        candidates, inside, resolution = self._find_inside_candidates()
        points = candidates[inside]
        n = self.config.n_particles
        if len(points) < n:
            raise RuntimeError(f"not enough inside points: have {len(points)}, need {n}")
        selected = points.copy() if len(points) == n else points[np.linspace(0, len(points) - 1, n, dtype=np.int64)].copy()
        accepted = int(np.count_nonzero(inside))
        geometry_volume = self.domain_volume * float(accepted) / float(resolution ** 3)
        vol0 = np.full(n, geometry_volume / n, dtype=np.float32)
        mass = vol0 * np.float32(self.config.p_rho)
        stats = {
            "geometry_type": self.config.geometry_type,
            "candidate_resolution": int(resolution),
            "candidate_count": int(len(candidates)),
            "accepted_count": int(len(points)),
            "particle_count": int(len(selected)),
            "geometry_volume": float(geometry_volume),
            "deterministic": bool(self.config.deterministic),
        }
        cloud = {"x": selected.astype(np.float32), "vol0": vol0, "mass": mass.astype(np.float32),
                 "geometry_volume": float(geometry_volume), "sampling_stats": stats}
        if self.config.geometry_type == "duct_flap_proxy":
            masks = self.component_masks(selected)
            stats.update({f"{name}_particle_count": int(np.count_nonzero(m)) for name, m in masks.items()})
            stats.update(duct_flap_proxy_sampling_stats(self.config))
            cloud["fixed_base_mask"] = masks["fixed_base"].astype(np.int8)
            cloud["free_tip_proxy_mask"] = masks["free_tip_proxy"].astype(np.int8)
        return cloud
