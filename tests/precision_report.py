"""Precision form at scale against the FP64 oracle (a GPU report, not a test):
FP32, FP16-deviation and BF16-deviation storage of the 192x96x96 case
(mpm_lbm.precision_case) after `--steps` FSI steps, each compared with the
FP64 oracle run of the same case.  Prints one JSON line.

    python tests/precision_report.py --steps 2
"""
import argparse
import json
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
sys.path[:0] = [str(ROOT / "mpm-lbm_amd"), str(ROOT)]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=2)
    args = ap.parse_args()
    import numpy as np
    import torch

    from mpm_lbm import precision_case
    from mpm_lbm.sim.drivers.fsi_driver import FSIDriver3D
    from oracle.oracle import OracleFSI

    base, config, state, compare = precision_case.setup(args.steps)
    runs, flags = {}, None
    for st in ("f32", "f16", "bf16"):
        d = FSIDriver3D(config(st), f"/tmp/fsi_prec_report/{st}", particle_cloud=base.cloud, write_geometry_file=False)
        d.initialize()
        s0 = state(d.lbm.rho.to_numpy(), d.lbm.v.to_numpy(), d.lbm.solid.to_numpy(), d.solid.x.to_numpy())
        flags = d.lbm.solid.to_numpy() if flags is None else flags
        for _ in range(args.steps):
            d.step_once()
        torch.cuda.synchronize()
        runs[st] = (s0, state(d.lbm.rho.to_numpy(), d.lbm.v.to_numpy(), d.lbm.solid.to_numpy(), d.solid.x.to_numpy()))
        print(f"[precision] {st}: {args.steps} FSI steps done", flush=True)
        del d
        torch.cuda.empty_cache()
    t0 = time.perf_counter()
    o = OracleFSI(config("f32"), flags, base.cloud, base.material, dtype=np.float64)
    f = o.fields()
    o0 = state(f["rho"], f["v"], f["solid"], f["x"])
    for k in range(args.steps):
        o.step_once()
        print(f"[precision] FP64 oracle step {k + 1} ({time.perf_counter() - t0:.0f} s)", flush=True)
    f = o.fields()
    ref = state(f["rho"], f["v"], f["solid"], f["x"])
    out = {"workload": precision_case.WORKLOAD, "fsi_steps": args.steps,
           "vs_fp64": {st: compare(r[1], ref, o0["mass"], r[0]["mass"]) for st, r in runs.items()},
           "fp64_mass_drift": (ref["mass"] - o0["mass"]) / o0["mass"]}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
