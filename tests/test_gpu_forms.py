"""GPU parity of the paths the plain-schedule tests do not reach:

* the precision form on libfsi_hip.so: FP16 / BF16 deviation storage (MRT and
  regularised collision) against the FP64 oracle, to the same error bounds
  as the emulation build (the bounds and workload of
  tests/test_precision_forms.py), and the regularised collision in FP32 to the
  1e-5 parity bar;
* the substep variant for lattices larger than one wave generation (80^3:
  2,000 workgroups of 4 waves, more than the chip holds at once), which
  loads the populations with the flags and runs wall cells in place;
* coupling_mode "none" (no projection or reaction; gravity only);
* the served form: requests admitted and retired over a changing mix, each
  held to its own oracle run.
"""
import numpy as np
import pytest

from mpm_lbm import workloads
from mpm_lbm.sim.drivers.fsi_config import FSIDriverConfig
from oracle.oracle import OracleFSI
from tests.parity import check_history, driver_fields, field_errors, fields_errors, make_pair, run_and_compare
from tests.test_precision_forms import BOUNDS, _workload

pytestmark = pytest.mark.gpu
TOL = 1.0e-5


@pytest.mark.parametrize("collision", ["mrt", "regularized"])
@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_gpu_16bit_storage_error_vs_fp64(hip, tmp_path, storage, collision):
    d, o = make_pair(_workload(tmp_path, storage, collision), tmp_path / "run", hip, None, dtype=np.float64)
    history = run_and_compare(d, o, 3)
    print(storage, collision, {k: v for k, v in history[-1].items() if k in BOUNDS[storage]}, flush=True)
    last = history[-1]
    assert last["solid_equal"]
    for key, bound in BOUNDS[storage].items():
        assert last[key] <= bound, (storage, collision, key, last[key])


def test_gpu_regularized_fp32_meets_kernel_bar(hip, tmp_path):
    d, o = make_pair(_workload(tmp_path, "f32", "regularized"), tmp_path / "run", hip, None, with_fp64=True)
    history = run_and_compare(d, o, 3)
    for step, errs in enumerate(history):
        print("regularized f32", step, {k: v for k, v in errs.items() if k not in ("vs_fp64", "method")},
              flush=True)
    check_history(history, "regularized f32", TOL)


def test_gpu_multi_generation_lattice(hip, tmp_path):
    import torch

    n = 80
    w = workloads.moving_boundary_flap(n=n, steps=2, substeps=20, flap=(4, 40, 30), out_dir=str(tmp_path))
    tiles = -(-(n * n) // 256)
    simds = 4 * torch.cuda.get_device_properties(0).multi_processor_count
    assert tiles * n > simds  # more workgroups of 4 waves than one generation holds
    d, o = make_pair(w, tmp_path / "run", hip, None, with_fp64=True)
    history = run_and_compare(d, o, 2)
    for step, errs in enumerate(history):
        print("mb80", step, {k: v for k, v in errs.items() if k not in ("vs_fp64", "method")}, flush=True)
    check_history(history, "mb80", TOL)
    assert d.lbm.get_moving_boundary_stats()["bb_link_count"] == o.mbstats.link_count


def test_gpu_none_mode(hip, tmp_path):
    w = workloads.penalty_block(n=32, steps=3, block=(8, 8, 4), out_dir=str(tmp_path))
    cfg = FSIDriverConfig(coupling_mode="none", n_grid=32, n_particles=w.config.n_particles, n_lbm_steps=3,
                          mpm_substeps_per_lbm_step=4, mpm_dt=1e-4, gravity=(0.0, -9.8, 0.0))
    w2 = workloads.Workload("none", cfg, w.cloud, w.geometry_json, (400.0, 0.2, 1.0))
    d, o = make_pair(w2, tmp_path / "run", hip, None, with_fp64=True)
    history = run_and_compare(d, o, 3)
    for step, errs in enumerate(history):
        print("none", step, {k: v for k, v in errs.items() if k not in ("vs_fp64", "method")}, flush=True)
    check_history(history, "none", TOL, strict=True)


def test_gpu_served_requests_match_oracle(hip, tmp_path):
    from mpm_lbm.serve import FSIServer, Request, latency_summary

    mk = workloads.flap_case
    reqs = [
        Request(0, mk("moving_boundary", 24, substeps=8, thickness=2, seed=11, mpm_substeps=2, out_dir=str(tmp_path)),
                2, 0.0),
        Request(1, mk("penalty", 16, substeps=8, thickness=2, seed=12, out_dir=str(tmp_path)), 1, 0.5),
        Request(2, mk("moving_boundary", 20, substeps=8, thickness=4, seed=13, mpm_substeps=3, E_mod=3e4,
                      out_dir=str(tmp_path)), 2, 1.2),
        Request(3, mk("penalty", 12, substeps=8, thickness=2, seed=14, E_mod=2e5, out_dir=str(tmp_path)), 2, 3.5),
    ]
    server = FSIServer(native=hip, out_dir=str(tmp_path / "srv"), n_substeps=8)
    kept = {}

    def keep(r, d):  # copy the results out before the server releases the request's device state
        kept[r.rid] = (driver_fields(d), d.lbm.static_solid.to_numpy())

    server.run(reqs, sim_exchange_seconds=1.0, on_done=keep)
    assert all(r.done is not None and r.driver is None for r in reqs) and server.regroups >= 3
    st = server.stats()
    print("served stats", st, flush=True)
    # one slot table for the whole stream: a graph per launch plan, replayed across live-set changes
    assert st["graph_builds"] + st["graph_replays"] == st["exchanges"]
    for r in reqs:
        w = r.workload
        fields, flags = kept[r.rid]
        o = OracleFSI(w.config, flags, w.cloud, w.material, dtype=np.float32)
        o.fp64 = OracleFSI(w.config, flags, w.cloud, w.material, dtype=np.float64)
        steps = r.exchanges if w.config.coupling_mode == "moving_boundary" else r.exchanges * 8
        for _ in range(steps):
            o.step_once()
            o.fp64.step_once()
        errs = fields_errors(fields, o)
        print("served", r.rid, {k: v for k, v in errs.items() if k not in ("vs_fp64", "method")}, flush=True)
        check_history([errs], f"served request {r.rid}", TOL)
    s = latency_summary(reqs)
    assert s["completed"] == 4 and s["p50_s"] <= s["p99_s"] <= s["max_s"]
