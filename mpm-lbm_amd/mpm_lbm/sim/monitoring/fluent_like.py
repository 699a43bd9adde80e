"""Nearest-particle displacement monitor (reference: sim/monitoring/fluent_like.py:6-51)."""
from __future__ import annotations

import numpy as np


def _points(values, name: str) -> np.ndarray:
    arr = np.asarray(values, dtype=np.float64)
    if arr.ndim != 2 or arr.shape[1] != 3:
        raise ValueError(f"{name} must have shape (n, 3)")
    if len(arr) == 0:
        raise ValueError(f"{name} must not be empty")
    if not np.all(np.isfinite(arr)):
        raise ValueError(f"{name} must be finite")
    return arr


def official_point_like_displacement(initial_positions, current_positions, target_point_norm, nearest_count: int,
                                     scale_m: float, radius_norm: float | None = None) -> dict:
    """Mean displacement of the particles initially nearest the target point
    (all within radius_norm when given, else the nearest_count closest; stable order)."""
    initial = _points(initial_positions, "initial_positions")
    current = _points(current_positions, "current_positions")
    target = np.asarray(target_point_norm, dtype=np.float64)
    if target.shape != (3,) or not np.all(np.isfinite(target)):
        raise ValueError("target_point_norm must be a finite 3-vector")
    if initial.shape != current.shape:
        raise ValueError("initial_positions and current_positions must have the same shape")
    if nearest_count <= 0:
        raise ValueError("nearest_count must be positive")
    if scale_m <= 0.0:
        raise ValueError("scale_m must be positive")
    if radius_norm is not None and radius_norm <= 0.0:
        raise ValueError("radius_norm must be positive when provided")
    dist = np.linalg.norm(initial - target[None, :], axis=1)
    order = np.argsort(dist, kind="stable")
    selected = order[: min(int(nearest_count), len(order))]
    if radius_norm is not None:
        within = order[dist[order] <= float(radius_norm)]
        if len(within):
            selected = within
    disp = np.mean(current[selected] - initial[selected], axis=0) * float(scale_m)
    return {
        "selected_particle_count": int(len(selected)),
        "selected_particle_indices": [int(i) for i in selected.tolist()],
        "target_point_norm": [float(v) for v in target.tolist()],
        "nearest_distance_norm_min": float(np.min(dist[selected])),
        "nearest_distance_norm_max": float(np.max(dist[selected])),
        "official_point_like_total_displacement_m": float(np.linalg.norm(disp)),
        "official_point_like_x_displacement_m": float(disp[0]),
        "official_point_like_y_displacement_m": float(disp[1]),
        "official_point_like_z_displacement_m": float(disp[2]),
        "monitor_is_direct_fluent_equivalent": False,
        "monitor_scope": "nearest-particle proxy for Fluent structural point surface displacement",
    }
