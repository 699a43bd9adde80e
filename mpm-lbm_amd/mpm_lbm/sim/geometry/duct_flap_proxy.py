"""Procedural duct + flap geometry (reference: sim/geometry/duct_flap_proxy.py).

Cell flags use cell centres (i + 0.5) / n and inclusive bounds in float64
(duct_flap_proxy.py:190-232, :235-238); the fixed base is the lowest
max(thickness, 0.2 height) of each flap and the free tip the highest
(:57-86).  Extension: `duct_flap_proxy_static_geometry` accepts a
(nx, ny, nz) shape for non-cubic lattices with cubic cells of size 1/nx.
"""
from __future__ import annotations

import math
from typing import Any

import numpy as np

DEFAULT_DUCT = {"x": [0.0, 1.0], "y": [0.3, 0.7], "z": [0.45, 0.55]}
DEFAULT_FLAP = {
    "anchor_x": 0.505,
    "anchor_y": 0.3,
    "height_over_duct_height": 0.25,
    "thickness_over_duct_height": 0.075,
    "normalized_height": 0.10,
    "normalized_thickness": 0.03,
    "z": [0.45, 0.55],
    "fixed_base": True,
    "mirrored_pair": False,
}
DEFAULT_MATERIAL_REFERENCE = {
    "density": 1600.0,
    "youngs_modulus": 1.0e6,
    "poisson_ratio": 0.47,
    "used_for_exact_structural_model": False,
}


def _duct_dict(config) -> dict:
    return dict(DEFAULT_DUCT if getattr(config, "duct", None) is None else config.duct)


def _flap_dict(config) -> dict:
    return dict(DEFAULT_FLAP if getattr(config, "flap", None) is None else config.flap)


def _material_dict(config) -> dict:
    value = getattr(config, "material_reference", None)
    return dict(DEFAULT_MATERIAL_REFERENCE if value is None else value)


def _check_interval(values, name: str) -> None:
    if len(values) != 2:
        raise ValueError(f"{name} must contain exactly two values")
    lo, hi = float(values[0]), float(values[1])
    if not math.isfinite(lo) or not math.isfinite(hi) or lo >= hi:
        raise ValueError(f"{name} must be finite and increasing")


def _check_number(value, name: str, position: bool) -> None:
    number = float(value)
    if not math.isfinite(number):
        raise ValueError(f"{name} must be finite")
    if not position and number <= 0.0:
        raise ValueError(f"{name} must be positive")


def validate_duct_flap_proxy_config(config) -> None:
    duct, flap, material = _duct_dict(config), _flap_dict(config), _material_dict(config)
    for key in ("x", "y", "z"):
        _check_interval(duct[key], f"duct.{key}")
    _check_interval(flap["z"], "flap.z")
    for key in ("anchor_x", "anchor_y", "height_over_duct_height", "thickness_over_duct_height"):
        _check_number(flap[key], f"flap.{key}", position=key.startswith("anchor_"))
    for key in ("normalized_height", "normalized_thickness"):
        _check_number(flap[key], f"flap.{key}", position=False)
    for key in ("density", "youngs_modulus", "poisson_ratio"):
        _check_number(material[key], f"material_reference.{key}", position=False)
    if bool(material.get("used_for_exact_structural_model", True)):
        raise ValueError("duct_flap_proxy material_reference must not claim exact structural-model equivalence")


def _inside(points: np.ndarray, lo, hi) -> np.ndarray:
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    return np.all((points >= lo) & (points <= hi), axis=1)


def flap_bounds(config) -> list[tuple[np.ndarray, np.ndarray, str]]:
    flap = _flap_dict(config)
    half = 0.5 * float(flap["normalized_thickness"])
    ax, ay, h = float(flap["anchor_x"]), float(flap["anchor_y"]), float(flap["normalized_height"])
    z0, z1 = float(flap["z"][0]), float(flap["z"][1])
    bounds = [(np.array([ax - half, ay, z0]), np.array([ax + half, ay + h, z1]), "y_min")]
    if bool(flap.get("mirrored_pair", False)):
        top = float(_duct_dict(config)["y"][1])
        bounds.append((np.array([ax - half, top - h, z0]), np.array([ax + half, top, z1]), "y_max"))
    return bounds


def duct_flap_proxy_component_masks(points: np.ndarray, config) -> dict[str, np.ndarray]:
    pts = np.asarray(points, dtype=np.float64)
    if pts.ndim != 2 or pts.shape[1] != 3:
        raise ValueError("points must have shape (n, 3)")
    if not np.all(np.isfinite(pts)):
        raise ValueError("points must be finite")
    duct, flap = _duct_dict(config), _flap_dict(config)
    lo = [duct["x"][0], duct["y"][0], duct["z"][0]]
    hi = [duct["x"][1], duct["y"][1], duct["z"][1]]
    out = {
        "duct_context": _inside(pts, lo, hi),
        "flap": np.zeros(len(pts), dtype=bool),
        "fixed_base": np.zeros(len(pts), dtype=bool),
        "free_tip_proxy": np.zeros(len(pts), dtype=bool),
    }
    base_thickness = max(float(flap["normalized_thickness"]), 0.2 * float(flap["normalized_height"]))
    for fmin, fmax, side in flap_bounds(config):
        inside = _inside(pts, fmin, fmax)
        out["flap"] |= inside
        if side == "y_max":
            base_lo = [fmin[0], max(fmin[1], fmax[1] - base_thickness), fmin[2]]
            base = _inside(pts, base_lo, fmax)
            tip_hi = [fmax[0], min(fmax[1], fmin[1] + base_thickness), fmax[2]]
            tip = _inside(pts, fmin, tip_hi)
        else:
            base_hi = [fmax[0], min(fmax[1], fmin[1] + base_thickness), fmax[2]]
            base = _inside(pts, fmin, base_hi)
            tip_lo = [fmin[0], max(fmin[1], fmax[1] - base_thickness), fmin[2]]
            tip = _inside(pts, tip_lo, fmax)
        out["fixed_base"] |= base & inside
        out["free_tip_proxy"] |= tip & inside
    return out


def inside_duct_flap_proxy(points: np.ndarray, config) -> np.ndarray:
    return duct_flap_proxy_component_masks(points, config)["flap"].copy()


def geometry_ratios(config) -> dict[str, float]:
    duct, flap = _duct_dict(config), _flap_dict(config)
    height = float(duct["y"][1]) - float(duct["y"][0])
    return {
        "height_over_duct_height": float(flap["normalized_height"]) / height,
        "thickness_over_duct_height": float(flap["normalized_thickness"]) / height,
    }


def duct_flap_proxy_sampling_stats(config) -> dict[str, Any]:
    duct, flap, material = _duct_dict(config), _flap_dict(config), _material_dict(config)
    ratios = geometry_ratios(config)
    return {
        "duct_flap_proxy_enabled": True,
        "duct_x_min": float(duct["x"][0]), "duct_x_max": float(duct["x"][1]),
        "duct_y_min": float(duct["y"][0]), "duct_y_max": float(duct["y"][1]),
        "duct_z_min": float(duct["z"][0]), "duct_z_max": float(duct["z"][1]),
        "flap_anchor_x": float(flap["anchor_x"]), "flap_anchor_y": float(flap["anchor_y"]),
        "flap_normalized_height": float(flap["normalized_height"]),
        "flap_normalized_thickness": float(flap["normalized_thickness"]),
        "flap_count": len(flap_bounds(config)),
        "flap_height_over_duct_height": float(ratios["height_over_duct_height"]),
        "flap_thickness_over_duct_height": float(ratios["thickness_over_duct_height"]),
        "material_density": float(material["density"]),
        "material_youngs_modulus": float(material["youngs_modulus"]),
        "material_poisson_ratio": float(material["poisson_ratio"]),
        "used_for_mpm_config": bool(material.get("used_for_mpm_config", False)),
        "used_for_exact_structural_model": bool(material["used_for_exact_structural_model"]),
    }


def cell_centres(shape) -> np.ndarray:
    """Cell centres (i + 0.5) / nx of a lattice with cubic cells (duct_flap_proxy.py:193-195)."""
    nx = float(shape[0])
    axes = [(np.arange(n, dtype=np.float64) + 0.5) / nx for n in shape]
    return np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1).reshape(-1, 3)


def duct_flap_proxy_static_geometry(n_grid, config, include_flap: bool = False) -> tuple[np.ndarray, dict[str, Any]]:
    """int8 solid flags (1 = solid) of shape (nx, ny, nz); n_grid may be an int or a 3-tuple."""
    shape = (int(n_grid),) * 3 if np.isscalar(n_grid) else tuple(int(v) for v in n_grid)
    if min(shape) <= 0:
        raise ValueError("n_grid must be positive")
    points = cell_centres(shape)
    duct = _duct_dict(config)
    duct_mask = _inside(points, [duct["x"][0], duct["y"][0], duct["z"][0]],
                        [duct["x"][1], duct["y"][1], duct["z"][1]]).reshape(shape)
    solid = np.ones(shape, dtype=np.int8)
    solid[duct_mask] = 0
    flap_cells = 0
    if include_flap:
        mask = np.zeros(len(points), dtype=bool)
        for fmin, fmax, _side in flap_bounds(config):
            mask |= _inside(points, fmin, fmax)
        mask = mask.reshape(shape)
        solid[mask] = 1
        flap_cells = int(np.count_nonzero(mask))
    fluid = solid == 0
    solid_count = int(np.count_nonzero(solid))
    name = f"geo_duct_flap_proxy_{shape[0]}.dat" if len(set(shape)) == 1 else "geo_duct_flap_proxy_{}x{}x{}.dat".format(*shape)
    report = {
        "all_fluid_geometry_used": bool(solid_count == 0),
        "duct_wall_cell_count": int(solid_count - flap_cells),
        "flap_static_cell_count": flap_cells,
        "flap_count": len(flap_bounds(config)),
        "fluid_cell_count": int(np.count_nonzero(fluid)),
        "geometry_name": name,
        "inlet_fluid_cell_count": int(np.count_nonzero(fluid[0])),
        "pressure_outlet_fluid_cell_count": int(np.count_nonzero(fluid[shape[0] - 1])),
        "solid_cell_count": solid_count,
        "static_geometry_scope": "duct walls only; flap is represented by MPM particles unless include_flap is true",
        "top_bottom_wall_cells_present": bool(solid_count > 0),
    }
    return solid, report
