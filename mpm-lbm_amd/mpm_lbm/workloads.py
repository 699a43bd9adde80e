"""Benchmark workloads: duct-with-solid FSI cases and their particle clouds.

Geometry: a duct spanning the lattice in x with one-cell solid walls at the
y and z extremes (duct_flap_proxy geometry), a velocity inlet at x = 0 and a
rho = 1 outlet at x = nx-1 (legacy equilibrium boundary conditions).  Solids
are boxes of whole lattice cells filled with 8 particles per cell at the
offsets (0.25, 0.75) of each axis plus a seeded U(-0.25, 0.25) dx jitter, with
vol0 = dx^3 / 8 and mass = rho_s vol0.  Flaps have their bottom cell layer
pinned (fixed_mask) and their top cell layer marked as the free tip.

Time steps are chosen for stability of the explicit MPM (dt below the
elastic CFL limit dx / sqrt(E / rho_s)), and the lattice viscosity is set
through `physical_nu_mapping` so that tau = 3 nu_lbm + 1/2.
"""
from __future__ import annotations

import json
import os
from dataclasses import dataclass, field
from typing import Optional, Tuple

import numpy as np

from .sim.drivers.fsi_config import FSIDriverConfig


@dataclass
class Workload:
    name: str
    config: FSIDriverConfig
    cloud: dict
    geometry_json: str
    material: Tuple[float, float, float]  # E, nu, rho_s
    notes: dict = field(default_factory=dict)

    @property
    def shape(self):
        n = self.config.n_grid
        return tuple(self.config.lattice_shape) if self.config.lattice_shape else (n, n, n)

    @property
    def n_cells(self):
        nx, ny, nz = self.shape
        return nx * ny * nz


def solid_box_cloud(shape, cell_lo, cell_hi, rho_s, seed, fixed_bottom=False, tip_top=False):
    """8 jittered particles per lattice cell in the cell box [cell_lo, cell_hi)."""
    nx = shape[0]
    dx = 1.0 / nx
    rng = np.random.default_rng(seed)
    cells = np.stack(np.meshgrid(*[np.arange(lo, hi) for lo, hi in zip(cell_lo, cell_hi)], indexing="ij"),
                     axis=-1).reshape(-1, 3)
    sub = np.stack(np.meshgrid([0.25, 0.75], [0.25, 0.75], [0.25, 0.75], indexing="ij"), axis=-1).reshape(-1, 3)
    base = (cells[:, None, :] + sub[None, :, :]).reshape(-1, 3)
    jitter = rng.uniform(-0.25, 0.25, size=base.shape)
    x = ((base + jitter) * dx).astype(np.float32)
    cell_of = np.repeat(cells, 8, axis=0)
    P = len(x)
    vol0 = np.full(P, dx ** 3 / 8.0, dtype=np.float32)
    cloud = {"x": x, "vol0": vol0, "mass": (vol0 * np.float32(rho_s)).astype(np.float32)}
    if fixed_bottom:
        cloud["fixed_base_mask"] = (cell_of[:, 1] == cell_lo[1]).astype(np.int8)
    if tip_top:
        cloud["free_tip_proxy_mask"] = (cell_of[:, 1] == cell_hi[1] - 1).astype(np.int8)
    return cloud


def duct_geometry_dict(shape, n_particles, material):
    nx, ny, nz = shape
    h = 1.0 / nx
    E_mod, nu, rho_s = material
    return {
        "geometry_type": "duct_flap_proxy",
        "n_particles": int(n_particles),
        "duct": {"x": [0.0, 1.0], "y": [h, (ny - 1) * h], "z": [h, (nz - 1) * h]},
        "material_reference": {"density": float(rho_s), "youngs_modulus": float(E_mod), "poisson_ratio": float(nu),
                               "used_for_mpm_config": True, "used_for_exact_structural_model": False},
    }


def _write_geometry(out_dir, name, data):
    os.makedirs(out_dir, exist_ok=True)
    path = os.path.join(out_dir, f"{name}_geometry.json")
    with open(path, "w", encoding="utf-8") as f:
        json.dump(data, f, indent=2)
    return path


def _viscosity_for_tau(tau, n, dt):
    nu_lbm = (tau - 0.5) / 3.0
    dx_phys = 1.0 / n
    return nu_lbm * dx_phys * dx_phys / dt


def penalty_block(n=32, steps=100, u_in=0.04, seed=1, out_dir="/tmp/fsi_workloads", block=(8, 8, 4),
                  E_mod=5.0e4, nu=0.3, rho_s=2.0, tau=0.6, mpm_dt=1.0e-4, beta=1.0e-3, cap=1.0e-4) -> Workload:
    """Plain penalty-mode FSI: free elastic block at the duct centre, one LBM step per FSI step."""
    shape = (n, n, n)
    lo = tuple((n - b) // 2 for b in block)
    hi = tuple(l + b for l, b in zip(lo, block))
    material = (E_mod, nu, rho_s)
    cloud = solid_box_cloud(shape, lo, hi, rho_s, seed)
    P = len(cloud["x"])
    geo = _write_geometry(out_dir, f"penalty_{n}", duct_geometry_dict(shape, P, material))
    cfg = FSIDriverConfig(
        coupling_mode="penalty", geometry_type="duct_flap_proxy", geometry_config_path=geo, n_grid=n, n_particles=P,
        n_lbm_steps=steps, mpm_substeps_per_lbm_step=1, mpm_dt=mpm_dt, target_u_lbm=(u_in, 0.0, 0.0),
        lbm_boundary_condition_mode="duct_velocity_inlet_pressure_outlet", lbm_viscosity_semantics="physical_nu_mapping",
        fluid_kinematic_viscosity_m2_s=_viscosity_for_tau(tau, n, mpm_dt), beta_lbm=beta, penalty_force_cap_lbm=cap,
        output_interval=max(1, steps // 10))
    return Workload(f"penalty_{n}^3", cfg, cloud, geo, material, {"tau": tau, "seed": seed})


def moving_boundary_flap(n=64, steps=50, substeps=120, u_in=0.05, seed=2, out_dir="/tmp/fsi_workloads",
                         flap=(4, 32, 24), E_mod=1.0e5, nu=0.35, rho_s=2.0, tau=0.6, mpm_dt=2.0e-5, mpm_substeps=10,
                         shape: Optional[Tuple[int, int, int]] = None, mb_cap=1.0e-4,
                         geometry_name: Optional[str] = None) -> Workload:
    """Subcycled moving_boundary FSI: a flap anchored on the y = 0 wall, `substeps` LBM substeps per FSI step.
    The LBM step is dt_lbm = mpm_substeps * mpm_dt / substeps so both solvers advance one FSI step per exchange."""
    shape = tuple(shape) if shape else (n, n, n)
    nx, ny, nz = shape
    fx, fy, fz = flap
    lo = (nx // 2 - fx // 2, 1, nz // 2 - fz // 2)
    hi = (lo[0] + fx, lo[1] + fy, lo[2] + fz)
    material = (E_mod, nu, rho_s)
    cloud = solid_box_cloud(shape, lo, hi, rho_s, seed, fixed_bottom=True, tip_top=True)
    P = len(cloud["x"])
    geo = _write_geometry(out_dir, geometry_name or "mb_{}x{}x{}_p{}".format(*shape, P),
                          duct_geometry_dict(shape, P, material))
    dt_lbm = mpm_substeps * mpm_dt / substeps
    dx_phys = 1.0 / nx
    cfg = FSIDriverConfig(
        coupling_mode="moving_boundary", geometry_type="duct_flap_proxy", geometry_config_path=geo, n_grid=nx,
        n_particles=P, n_lbm_steps=steps, mpm_substeps_per_lbm_step=mpm_substeps, mpm_dt=mpm_dt,
        target_u_lbm=(u_in, 0.0, 0.0), lbm_boundary_condition_mode="duct_velocity_inlet_pressure_outlet",
        lbm_viscosity_semantics="physical_nu_mapping", fluid_kinematic_viscosity_m2_s=_viscosity_for_tau(tau, nx, dt_lbm),
        fsi_exchange_mode="lbm_subcycled_per_fsi_step", lbm_substeps_per_fsi_step=substeps,
        lbm_dt_phys_override_s=dt_lbm, official_fsi_dt_s=substeps * dt_lbm,
        target_u_lbm_for_dimensional_mapping=u_in, target_inlet_velocity_mps=u_in * dx_phys / dt_lbm,
        mb_force_cap_norm=mb_cap, output_interval=max(1, steps // 10),
        lattice_shape=None if len(set(shape)) == 1 else shape)
    return Workload("moving_boundary_{}x{}x{}".format(*shape), cfg, cloud, geo, material,
                    {"tau": tau, "seed": seed, "substeps": substeps, "dt_lbm": dt_lbm})


def moving_boundary_unsubcycled(n=32, steps=10, **kw) -> Workload:
    """moving_boundary with one LBM step per FSI step (no force accumulator)."""
    w = moving_boundary_flap(n=n, steps=steps, substeps=1, **kw)
    c = w.config
    cfg = FSIDriverConfig(**{**c.to_dict(), "fsi_exchange_mode": "one_lbm_step_per_fsi_step",
                             "lbm_substeps_per_fsi_step": 1, "lbm_dt_phys_override_s": None,
                             "official_fsi_dt_s": None, "target_u_lbm_for_dimensional_mapping": None,
                             "target_inlet_velocity_mps": None, "mpm_substeps_per_lbm_step": 1,
                             "lattice_shape": c.lattice_shape,
                             "fluid_kinematic_viscosity_m2_s": _viscosity_for_tau(w.notes["tau"], n, c.mpm_dt)})
    return Workload(w.name + "_unsubcycled", cfg, w.cloud, w.geometry_json, w.material, w.notes)


def flap_case(mode, n, substeps=120, u_in=0.05, E_mod=1.0e5, thickness=4, seed=0, out_dir="/tmp/fsi_workloads",
              rho_s=2.0, nu=0.35, tau=0.6, mpm_substeps=10, storage="f32", cfl=0.3) -> Workload:
    """A flap case of either coupling mode, with the MPM step at `cfl` times the
    elastic limit dx / sqrt(E / rho_s).  Moving-boundary instances subcycle
    `substeps` LBM steps per FSI step; penalty instances take one LBM step and
    one MPM substep per FSI step (the reference's only penalty schedule)."""
    dx = 1.0 / n
    dt_cfl = cfl * dx / np.sqrt(E_mod / rho_s)
    flap = (thickness, max(2, n // 2), max(2, (3 * n) // 8))
    name = f"{mode}_{n}_t{thickness}_s{seed}"
    if mode == "moving_boundary":
        w = moving_boundary_flap(n=n, steps=1, substeps=substeps, u_in=u_in, seed=seed, out_dir=out_dir, flap=flap,
                                 E_mod=E_mod, nu=nu, rho_s=rho_s, tau=tau, mpm_dt=dt_cfl, mpm_substeps=mpm_substeps,
                                 geometry_name=name)
    else:
        shape = (n, n, n)
        lo = (n // 2 - flap[0] // 2, 1, n // 2 - flap[2] // 2)
        hi = (lo[0] + flap[0], lo[1] + flap[1], lo[2] + flap[2])
        cloud = solid_box_cloud(shape, lo, hi, rho_s, seed, fixed_bottom=True, tip_top=True)
        P = len(cloud["x"])
        geo = _write_geometry(out_dir, name, duct_geometry_dict(shape, P, (E_mod, nu, rho_s)))
        cfg = FSIDriverConfig(
            coupling_mode=mode, geometry_type="duct_flap_proxy", geometry_config_path=geo, n_grid=n, n_particles=P,
            n_lbm_steps=1, mpm_substeps_per_lbm_step=1, mpm_dt=dt_cfl, target_u_lbm=(u_in, 0.0, 0.0),
            lbm_boundary_condition_mode="duct_velocity_inlet_pressure_outlet", lbm_viscosity_semantics="physical_nu_mapping",
            fluid_kinematic_viscosity_m2_s=_viscosity_for_tau(tau, n, dt_cfl))
        w = Workload(name, cfg, cloud, geo, (E_mod, nu, rho_s), {"tau": tau, "seed": seed})
    if storage != "f32":
        w = Workload(w.name, FSIDriverConfig(**{**w.config.to_dict(), "lbm_population_storage": storage}),
                     w.cloud, w.geometry_json, w.material, w.notes)
    return w


def batch_sweep(count=64, seed=4, sizes=(32, 48, 64, 96), substeps=120, out_dir="/tmp/fsi_workloads",
                storage="f32") -> list:
    """Batched parameter sweep: grid from `sizes`, u_in ~ U(0.02, 0.08),
    E ~ logU(1e4, 1e6), flap thickness in {2, 4, 6} cells, modes alternating
    moving_boundary / penalty, all from one seeded generator."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(rng.choice(sizes))
        u_in = float(rng.uniform(0.02, 0.08))
        E_mod = float(10 ** rng.uniform(4.0, 6.0))
        thickness = int(rng.choice([2, 4, 6]))
        mode = "moving_boundary" if k % 2 == 0 else "penalty"
        out.append(flap_case(mode, n, substeps=substeps, u_in=u_in, E_mod=E_mod, thickness=thickness, seed=seed * 1000 + k,
                             out_dir=out_dir, storage=storage))
    return out
