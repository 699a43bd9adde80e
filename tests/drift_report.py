"""Drift of the plain FP32 path over about 1000 LBM steps (a GPU report, not
a test).  The product runs beside the FP32 and FP64 oracles on the same
seeded inputs, and the per-field errors of tests/parity.py (max |a - b| /
max |b|) are printed at checkpoints, one JSON line each:

* penalty, 32^3 duct with an 8x8x4-cell block: `--penalty-steps` FSI steps
  of one LBM step each;
* moving boundary, the 64^3 flap: `--mb-steps` FSI steps of 120 LBM
  substeps each.

    python tests/drift_report.py --penalty-steps 1000 --mb-steps 9
"""
import argparse
import json
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
sys.path[:0] = [str(ROOT / "mpm-lbm_amd"), str(ROOT)]

FIELDS = ("F", "rho", "v", "phi", "solid_vel", "cell_force", "hydro", "x", "pv", "Fp", "C")


def _row(label, step, lbm_steps, errs):
    return {"case": label, "fsi_step": step, "lbm_steps": lbm_steps, "solid_equal": errs["solid_equal"],
            "vs_fp32_oracle": {k: errs[k] for k in FIELDS},
            "vs_fp64_oracle": {k: errs["vs_fp64"][k] for k in FIELDS},
            "fp32_oracle_vs_fp64": {k: errs["method"][k] for k in FIELDS}}


def run(label, w, steps, checkpoints, lbm_per_step, tmp):
    from tests.parity import field_errors, make_pair

    d, o = make_pair(w, Path(tmp) / label, None, None, with_fp64=True)
    t0 = time.perf_counter()
    for k in range(1, steps + 1):
        d.step_once()
        o.step_once()
        o.fp64.step_once()
        if k in checkpoints:
            row = _row(label, k, k * lbm_per_step, field_errors(d, o))
            row["elapsed_s"] = round(time.perf_counter() - t0, 1)
            print(json.dumps(row), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--penalty-steps", type=int, default=1000)
    ap.add_argument("--mb-steps", type=int, default=9)
    ap.add_argument("--tmp", default="/tmp/fsi_drift")
    args = ap.parse_args()
    from mpm_lbm import workloads

    if args.penalty_steps > 0:
        w = workloads.penalty_block(n=32, steps=args.penalty_steps, out_dir=args.tmp)
        marks = {1, 10, 100, 250, 500, 750, args.penalty_steps}
        run("penalty_32", w, args.penalty_steps, marks, 1, args.tmp)
    if args.mb_steps > 0:
        w = workloads.moving_boundary_flap(n=64, steps=args.mb_steps, substeps=120, out_dir=args.tmp)
        run("moving_boundary_64", w, args.mb_steps, set(range(1, args.mb_steps + 1)), 120, args.tmp)


if __name__ == "__main__":
    main()
