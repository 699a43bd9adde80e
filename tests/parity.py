"""Helpers shared by the parity tests: run the driver and the oracle side by side."""
from __future__ import annotations

import numpy as np

from mpm_lbm.sim.drivers.fsi_driver import FSIDriver3D
from oracle.oracle import OracleFSI


def rel_max_err(a, b, mask=None, floor=0.0):
    """max |a - b| / max(max |b|, floor) over the selected entries (field-norm relative error)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if mask is not None:
        a, b = a[mask], b[mask]
    scale = max(float(np.max(np.abs(b))) if b.size else 0.0, floor)
    if scale == 0.0:
        return float(np.max(np.abs(a))) if a.size else 0.0
    return float(np.max(np.abs(a - b))) / scale


def make_pair(workload, out_dir, native, device, dtype=np.float32):
    d = FSIDriver3D(workload.config, str(out_dir), particle_cloud=workload.cloud, native=native, device=device,
                    write_geometry_file=False)
    d.initialize()
    o = OracleFSI(workload.config, d.lbm.solid.to_numpy(), workload.cloud, workload.material, dtype=dtype)
    return d, o


def field_errors(d, o):
    of = o.fields()
    fluid = of["solid"] == 0
    return {
        "solid_equal": bool(np.array_equal(d.lbm.solid.to_numpy(), of["solid"])),
        "F": rel_max_err(d.lbm.F.to_numpy(), of["F"], fluid),
        "rho": rel_max_err(d.lbm.rho.to_numpy(), of["rho"], fluid),
        "v": rel_max_err(d.lbm.v.to_numpy(), of["v"], fluid),
        "phi": rel_max_err(d.lbm.solid_phi.to_numpy(), of["solid_phi"]),
        "solid_vel": rel_max_err(d.lbm.solid_vel.to_numpy(), of["solid_vel"]),
        "cell_force": rel_max_err(d.lbm.cell_force.to_numpy(), of["cell_force"]),
        "hydro": rel_max_err(d.lbm.hydro_force.to_numpy(), of["hydro_force"]),
        "x": rel_max_err(d.solid.x.to_numpy(), of["x"]),
        "pv": rel_max_err(d.solid.v.to_numpy(), of["pv"]),
        "Fp": rel_max_err(d.solid.F.to_numpy(), of["Fp"]),
        # APIC C is a velocity gradient: its natural scale is |v| / dx even when it vanishes
        "C": rel_max_err(d.solid.C.to_numpy(), of["C"],
                         floor=float(np.max(np.abs(of["pv"]))) * float(o.ma.inv_dx)),
    }


def run_and_compare(d, o, n_steps):
    history = []
    for _ in range(n_steps):
        d.step_once()
        o.step_once()
        history.append(field_errors(d, o))
    return history
