"""The precision-form case at scale and its error measures.

192x96x96 duct with a 6x48x36-cell flap (82,944 particles, seed 3), the
regularised collision at tau = 0.55, subcycled moving boundary (120 LBM
substeps per FSI step), in FP32, FP16-deviation or BF16-deviation storage.
`compare` states one run's error against a reference run of the same case:
density and velocity over the cells fluid in both (velocity relative to the
reference's largest |u|, also away from cells whose dynamic-solid test came
out differently), flap-tip displacement, and the drift of the total fluid
mass of each run from its own start.
"""
from __future__ import annotations

import numpy as np

from mpm_lbm import workloads
from mpm_lbm.sim.drivers.fsi_config import FSIDriverConfig

WORKLOAD = ("192x96x96 duct, 6x48x36 flap, 82,944 particles, regularised collision tau=0.55, "
            "moving boundary, 120 substeps per FSI step")


def setup(steps, out_dir="/tmp/fsi_bench_prec"):
    """(base workload, config(storage), state(rho, v, solid, x), compare(a, ref, m0_ref, m0_a))."""
    shape, E, rho_s = (192, 96, 96), 1.0e5, 2.0
    mpm_dt = 0.3 * (1.0 / shape[0]) / np.sqrt(E / rho_s)
    base = workloads.moving_boundary_flap(shape=shape, flap=(6, 48, 36), seed=3, tau=0.55, substeps=120,
                                          steps=steps, mpm_dt=mpm_dt, E_mod=E, rho_s=rho_s, out_dir=out_dir)
    tip = base.cloud["free_tip_proxy_mask"] != 0
    x0 = base.cloud["x"][tip].astype(np.float64).mean(0)

    def config(storage):
        return FSIDriverConfig(**{**base.config.to_dict(), "lbm_population_storage": storage,
                                  "lbm_collision": "regularized"})

    def state(rho, v, solid, x):
        fluid = solid == 0
        return {"rho": np.asarray(rho, np.float64), "v": np.asarray(v, np.float64), "fluid": fluid,
                "mass": float(np.asarray(rho, np.float64)[fluid].sum()),
                "tip": np.asarray(x, np.float64)[tip].mean(0) - x0}

    def compare(a, ref, m0_ref, m0_a):
        both = a["fluid"] & ref["fluid"]
        # cells whose dynamic-solid test (phi >= threshold) came out differently, dilated by 2 cells:
        # a tie at phi ~ threshold changes the bounce-back around it, whatever the precision
        near = a["fluid"] != ref["fluid"]
        for _ in range(2):
            grown = near.copy()
            for ax in range(3):
                grown |= np.roll(near, 1, ax) | np.roll(near, -1, ax)
            near = grown
        away = both & ~near
        du = np.abs(a["v"][both] - ref["v"][both])
        du_away = np.abs(a["v"][away] - ref["v"][away])
        umax = float(np.max(np.abs(ref["v"][both])))
        drho = np.abs(a["rho"][both] - ref["rho"][both])
        tipn = float(np.linalg.norm(ref["tip"]))
        return {
            "rho_max_err": float(drho.max()), "rho_mean_err": float(drho.mean()),
            "rho_max_err_away_from_flag_flips": float(np.abs(a["rho"][away] - ref["rho"][away]).max()),
            "u_max_rel_err": float(du.max() / umax), "u_mean_rel_err": float(du.mean() / umax),
            "u_max_rel_err_away_from_flag_flips": float(du_away.max() / umax),
            "tip_disp_rel_err": float(np.linalg.norm(a["tip"] - ref["tip"]) / tipn) if tipn > 0 else None,
            "mass_drift": (a["mass"] - m0_a) / m0_a, "mass_drift_ref": (ref["mass"] - m0_ref) / m0_ref,
            "flag_mismatch_cells": int(np.count_nonzero(a["fluid"] != ref["fluid"])),
        }

    return base, config, state, compare
