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


# Velocity-level particle fields.  With F stored in FP32 as I + H at strains
# ~1e-6, one rounding difference in grid momentum moves H by an ulp and the
# fixed-corotated stress by percent: any two FP32 implementations that sum the
# P2G / projection scatter in different orders (GPU atomics vs the oracle's
# sequential sums; the reference's own atomics run to run) differ here far
# above 1e-5.  Their bar is therefore stated against the FP32 method error
# itself: within max(1e-5, CONDITIONED_FRACTION * |FP32 oracle - FP64 oracle|).
CONDITIONED_FIELDS = ("pv", "C", "solid_vel")
CONDITIONED_FRACTION = 0.25


def make_pair(workload, out_dir, native, device, dtype=np.float32, with_fp64=False):
    d = FSIDriver3D(workload.config, str(out_dir), particle_cloud=workload.cloud, native=native, device=device,
                    write_geometry_file=False)
    d.initialize()
    o = OracleFSI(workload.config, d.lbm.solid.to_numpy(), workload.cloud, workload.material, dtype=dtype)
    if with_fp64:
        o.fp64 = OracleFSI(workload.config, d.lbm.solid.to_numpy(), workload.cloud, workload.material,
                           dtype=np.float64)
    return d, o


def method_errors(o):
    """|FP32 oracle - FP64 oracle| for the conditioned fields."""
    a, b = o.fields(), o.fp64.fields()
    return {
        "pv": rel_max_err(a["pv"], b["pv"]),
        "solid_vel": rel_max_err(a["solid_vel"], b["solid_vel"]),
        "C": rel_max_err(a["C"], b["C"], floor=float(np.max(np.abs(b["pv"]))) * float(o.ma.inv_dx)),
    }


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
        errs = field_errors(d, o)
        if getattr(o, "fp64", None) is not None:
            o.fp64.step_once()
            errs["method"] = method_errors(o)
        history.append(errs)
    return history


def check_history(history, label, tol=1.0e-5):
    """Assert the parity bar; velocity-level particle fields use the FP32 method error when it is known."""
    for step, errs in enumerate(history):
        assert errs["solid_equal"], f"{label}: solid flags differ at step {step}"
        method = errs.get("method", {})
        for k, v in errs.items():
            if k in ("solid_equal", "method"):
                continue
            bound = tol
            if k in CONDITIONED_FIELDS and k in method:
                bound = max(tol, CONDITIONED_FRACTION * method[k])
            assert v <= bound, f"{label} step {step}: {k} relative error {v:.3e} > {bound:.3e}"
