"""The precision, batched and served configurations at their own sizes, on the GPU.

* Precision form (192x96x96 duct, 6x48x36 flap, 82,944 particles, regularised
  collision at tau = 0.55): FP32, FP16- and BF16-deviation storage after one
  FSI step of 120 substeps, each against the FP64 oracle run of the case,
  within the bounds below.
* Batched form: the full 64-instance sweep ({32,48,64,96}^3, mixed modes),
  one exchange; every instance stays finite with its particles in the
  domain and its flags intact, and a seeded subset - a 96^3 moving-boundary
  instance among them - meets the single-instance parity bar against its own
  oracle run.
* Served form: one request of each lattice size {32,48,64,96,128}^3 in one
  stream, one exchange each, every request against its own oracle run.
"""
import numpy as np
import pytest

from mpm_lbm import _native, workloads
from oracle.oracle import OracleFSI
from tests.parity import check_history, driver_fields, field_errors, fields_errors

pytestmark = pytest.mark.gpu
TOL = 1.0e-5

# Bounds of each storage against FP64 after one FSI step (velocity relative to
# the largest |u|; "away" excludes cells next to a dynamic-solid flag that came
# out differently; tip = flap-tip displacement relative to its size).  FP32
# sits at its own rounding; FP16 (10-bit mantissa on deviations ~1e-2) adds
# ~1e-5 of |u| on average, BF16 (7-bit) ~1e-3.  Measured: u mean 7.8e-7 /
# 3.3e-5 / 4.9e-4, u max 1.3e-5 / 9.7e-4 / 1.2e-2, rho mean 4.0e-7 / 3.9e-6 /
# 6.2e-5, tip 5.9e-4 for all three (the flap follows the fluid's FP32 forces).
C3_BOUNDS = {
    "f32": {"u_mean_rel_err": 5e-5, "u_max_rel_err_away_from_flag_flips": 2e-2, "rho_mean_err": 1e-6,
            "tip_disp_rel_err": 2e-3},
    "f16": {"u_mean_rel_err": 3e-4, "u_max_rel_err_away_from_flag_flips": 2e-2, "rho_mean_err": 1e-5,
            "tip_disp_rel_err": 2e-3},
    "bf16": {"u_mean_rel_err": 5e-3, "u_max_rel_err_away_from_flag_flips": 5e-2, "rho_mean_err": 2e-4,
             "tip_disp_rel_err": 5e-3},
}


# |relative fluid-mass drift - the FP64 run's| after one FSI step (measured:
# FP32 3.7e-7, FP16 2.3e-6, BF16 5.4e-5)
C3_MASS_DRIFT_BOUND = {"f32": 2e-6, "f16": 1e-5, "bf16": 5e-4}


def test_gpu_c3_precision_form_vs_fp64(hip, tmp_path):
    import torch

    from mpm_lbm import precision_case
    from mpm_lbm.sim.drivers.fsi_driver import FSIDriver3D

    base, config, state, compare = precision_case.setup(1, out_dir=str(tmp_path))
    runs, flags = {}, None
    for st in ("f32", "f16", "bf16"):
        d = FSIDriver3D(config(st), str(tmp_path / st), particle_cloud=base.cloud, native=hip,
                        write_geometry_file=False)
        d.initialize()
        flags = d.lbm.solid.to_numpy() if flags is None else flags
        s0 = state(d.lbm.rho.to_numpy(), d.lbm.v.to_numpy(), d.lbm.solid.to_numpy(), d.solid.x.to_numpy())
        d.step_once()
        torch.cuda.synchronize()
        runs[st] = (s0, state(d.lbm.rho.to_numpy(), d.lbm.v.to_numpy(), d.lbm.solid.to_numpy(), d.solid.x.to_numpy()))
        del d
        torch.cuda.empty_cache()
    o = OracleFSI(config("f32"), flags, base.cloud, base.material, dtype=np.float64)
    f = o.fields()
    o0 = state(f["rho"], f["v"], f["solid"], f["x"])
    o.step_once()
    f = o.fields()
    ref = state(f["rho"], f["v"], f["solid"], f["x"])
    results = {st: compare(s1, ref, o0["mass"], s0["mass"]) for st, (s0, s1) in runs.items()}
    for st, errs in results.items():
        print("C3", st, errs, flush=True)
    for st, errs in results.items():
        for key, bound in C3_BOUNDS[st].items():
            assert errs[key] <= bound, (st, key, errs[key])
        # total fluid mass follows the FP64 run's (the duct is open: it is not conserved)
        assert abs(errs["mass_drift"] - errs["mass_drift_ref"]) <= C3_MASS_DRIFT_BOUND[st], (st, errs)


def _oracles(w, d):
    flags = d.lbm.static_solid.to_numpy()
    o = OracleFSI(w.config, flags, w.cloud, w.material, dtype=np.float32)
    o.fp64 = OracleFSI(w.config, flags, w.cloud, w.material, dtype=np.float64)
    return o


def _finite_and_sane(d):
    f = driver_fields(d)
    for k in ("F", "rho", "v", "x", "pv", "Fp", "C", "hydro_force", "solid_vel"):
        assert np.all(np.isfinite(f[k])), k
    x = f["x"]
    assert np.all(x > 0) and np.all(x < 1), "particles left the domain"
    static = d.lbm.static_solid.to_numpy() != 0
    assert np.all(f["solid"][static] != 0), "a static wall cell turned fluid"


def test_gpu_c4_full_sweep_one_exchange(hip, tmp_path):
    from mpm_lbm.batch import BatchedFSI

    ws = workloads.batch_sweep(count=64, seed=4, out_dir=str(tmp_path / "w"))
    sizes = [w.config.n_grid for w in ws]
    modes = [w.config.coupling_mode for w in ws]
    # seeded subset: the first moving-boundary instance at 96^3 and the first
    # instance of each other (size, mode) pair up to 64^3
    subset, seen = [], set()
    for k, (n, m) in enumerate(zip(sizes, modes)):
        key = (n, m)
        if key in seen or (n == 96 and m == "penalty") or (n == 64 and m == "penalty"):
            continue
        seen.add(key)
        subset.append(k)
    assert any(sizes[k] == 96 and modes[k] == "moving_boundary" for k in subset)
    # a fixed substep plan: the one exchange runs as the graph (no eager tuning exchange)
    batch = BatchedFSI(ws, out_dir=str(tmp_path / "batch"), native=hip, substep_plan="split")
    print("C4 instances", len(ws), "cells", batch.total_cells, "subset", [(k, sizes[k], modes[k]) for k in subset],
          flush=True)
    # the 64^3 moving-boundary instances run their subcycle as LBM_RESIDENT launches
    resident = [k for k in range(len(ws)) if batch._inst_host[k].resident]
    assert resident and all(sizes[k] == 64 and modes[k] == "moving_boundary" for k in resident)
    assert any(k in resident for k in subset)
    # the other moving-boundary instances in subcycle groups, on the graph's side
    # chain 1, beside the penalty instances on the main chain
    grouped = sorted(k for g in batch.subcycle_groups() for k in g)
    assert grouped == [k for k in range(len(ws)) if modes[k] == "moving_boundary" and k not in resident]
    assert any(k in grouped for k in subset)
    oracles = {k: _oracles(ws[k], batch.drivers[k]) for k in subset}
    batch.step()
    assert batch.graph_builds == 1 and 1 in {br for *_, br in batch._plan}
    batch.check_resident_waits()
    for d in batch.drivers:
        _finite_and_sane(d)
    for k in subset:
        o, d = oracles[k], batch.drivers[k]
        steps = 1 if modes[k] == "moving_boundary" else batch.n_substeps
        for _ in range(steps):
            o.step_once()
            o.fp64.step_once()
        errs = field_errors(d, o)
        print("C4 instance", k, sizes[k], modes[k], {q: v for q, v in errs.items() if q not in ("vs_fp64", "method")},
              flush=True)
        check_history([errs], f"C4 instance {k} ({sizes[k]}^3 {modes[k]})", TOL)


def test_gpu_c5_each_lattice_size_served(hip, tmp_path):
    from mpm_lbm.serve import FSIServer, Request, latency_summary

    mk = workloads.flap_case
    plan = [(32, "penalty"), (48, "penalty"), (64, "moving_boundary"), (96, "moving_boundary"),
            (128, "moving_boundary")]
    reqs = [Request(k, mk(m, n, substeps=120, thickness=2 + 2 * (k % 3), seed=50 + k, out_dir=str(tmp_path / "w")),
                    1, 0.25 * k)
            for k, (n, m) in enumerate(plan)]
    kept = {}

    def keep(r, d):
        kept[r.rid] = (driver_fields(d), d.lbm.static_solid.to_numpy())

    server = FSIServer(native=hip, out_dir=str(tmp_path / "srv"), n_substeps=120)
    server.run(reqs, sim_exchange_seconds=0.1, on_done=keep)
    assert all(r.done is not None and r.driver is None for r in reqs)
    for r in reqs:
        w = r.workload
        fields, flags = kept[r.rid]
        o = OracleFSI(w.config, flags, w.cloud, w.material, dtype=np.float32)
        o.fp64 = OracleFSI(w.config, flags, w.cloud, w.material, dtype=np.float64)
        steps = 1 if w.config.coupling_mode == "moving_boundary" else 120
        for _ in range(steps):
            o.step_once()
            o.fp64.step_once()
        errs = fields_errors(fields, o)
        print("C5 request", r.rid, w.config.n_grid, w.config.coupling_mode,
              {q: v for q, v in errs.items() if q not in ("vs_fp64", "method")}, flush=True)
        check_history([errs], f"C5 request {r.rid} ({w.config.n_grid}^3)", TOL)
    s = latency_summary(reqs)
    assert s["completed"] == 5
