#!/usr/bin/env python3
"""Benchmark: plain FP32 moving-boundary FSI at 64^3 on one MI355X.

Workload: 64^3 duct with a 4x32x24-cell elastic flap (24,576 particles,
seeded jitter), moving-wall bounce-back with the subcycle force
accumulator, 120 LBM substeps and 10 MPM substeps per FSI step.  A "step" of
this benchmark is one FSI step.  `value` is MLUPS counted over all lattice
cells, 64^3 * 120 * steps / seconds / 1e6, of the median of `--repeats`
blocks into which the K timed FSI steps are split (the blocks run back to
back between two synchronisations; each FSI step is projection, flag
rebuild, 120 fused LBM substeps, finalize, moments and 10 MPM substeps).

Also reported:
  roofline     the resident subcycle kernel (k_lbm_resident: one launch runs
               the FSI step's 120 LBM substeps with the lattice on chip):
               algorithmic bytes per launch (120 x (152 B per fluid cell +
               4 B cell word per cell), what the substeps would move through
               HBM one launch each) over its average duration, timed live
               with HIP events on the stream the kernels run on, against
               8 TB/s; the launch-per-substep kernel is timed beside it;
  copy_ceiling device-to-device copy bandwidth measured in the same run;
  cpu_baseline the reference-structured C++ oracle (OpenMP) on the host
               cores: median of 3 FSI steps of the same workload after one
               warm-up step;
  batched / served / precision   the three forms beside the plain path.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "mpm-lbm_amd"))
sys.path.insert(0, str(ROOT))

METRIC = "FSI MLUPS + % HBM peak at 64^3; batched/served vs plain at equal error; p99"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed FSI steps")
    ap.add_argument("--warmup", type=int, default=2, help="untimed FSI steps")
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--substeps", type=int, default=120)
    ap.add_argument("--kernel-launches", type=int, default=600,
                    help="launch-per-substep kernel launches timed beside the resident subcycle")
    ap.add_argument("--resident-launches", type=int, default=10, help="resident subcycle launches timed for the roofline")
    ap.add_argument("--cpu-steps", type=int, default=3,
                    help="timed FSI steps of the CPU baseline sample, after one warm-up step; median (0 = skip)")
    ap.add_argument("--traffic-json", default=str(ROOT / "profiles" / "lbm_resident_64cube_traffic.json"))
    ap.add_argument("--push-traffic-json", default=str(ROOT / "profiles" / "lbm_substep_64cube_traffic.json"))
    ap.add_argument("--batch", type=int, default=64, help="instances of the batched sweep (0 = skip)")
    ap.add_argument("--batch-exchanges", type=int, default=3, help="timed batched exchanges")
    ap.add_argument("--batch-f16", type=int, default=1,
                    help="also time the batched sweep with FP16-deviation storage (0 = skip)")
    ap.add_argument("--serve-seconds", type=float, default=8.0, help="arrival window of the served form (0 = skip)")
    ap.add_argument("--repeats", type=int, default=5,
                    help="timed blocks the K timed steps are split into; value = their median")
    ap.add_argument("--precision", type=int, default=17,
                    help="FSI steps of the 192x96x96 precision form (17 = 2,040 LBM steps), 16-bit storage against "
                         "FP32 on the GPU (0 = skip)")
    ap.add_argument("--precision-fp64-steps", type=int, default=2,
                    help="FSI steps of the precision form compared with the FP64 oracle on the host (0 = skip)")
    ap.add_argument("--serve-rate", type=float, default=20.0, help="Poisson arrival rate of served requests, per s")
    return ap.parse_args()


def copy_ceiling_gbs(torch, nbytes=1 << 30, reps=20):
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    for _ in range(3):
        b.copy_(a)
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    s.record()
    for _ in range(reps):
        b.copy_(a)
    e.record()
    torch.cuda.synchronize()
    ms = s.elapsed_time(e) / reps
    del a, b
    return 2.0 * nbytes / (ms * 1e-3) / 1e9


def time_substep_kernel(torch, lbm, launches):
    """Average duration of the fused substep as the FSI loop issues it
    (moving wall + gather of the previous substep): `launches` back-to-back
    launches captured in one HIP graph, timed with HIP events around its
    replay on the stream the kernels run on."""
    import ctypes

    from mpm_lbm import _native

    launches += launches % 2  # even: the buffers return to their start
    flags = _native.SUB_MOVING_WALL | _native.SUB_GATHER_PREV

    def issue():
        parity, cur = 0, lbm.cur
        stream = lbm._stream()
        for _ in range(launches):
            lbm.native.call("fsi_lbm_substep", ctypes.byref(lbm._params), ctypes.byref(lbm._state), cur, parity,
                            flags, stream)
            cur, parity = 1 - cur, 1 - parity

    issue()  # warm
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        issue()
    graph.replay()
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    graph.replay()
    e.record()
    torch.cuda.synchronize()
    lbm._moments_fresh = False
    return s.elapsed_time(e) / launches


def time_resident_kernel(torch, lbm, launches, n_sub, flags):
    """Average duration of one resident subcycle launch (k_lbm_resident: the
    FSI step's n_sub moving-wall substeps with the lattice on chip), issued
    with the FSI step's flags on the live lattice: HIP events around each of
    `launches` launches, on the stream the kernel runs on."""
    import ctypes

    from mpm_lbm import _native

    stream = lbm._stream()
    durations = []
    for _ in range(launches + 1):
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        lbm.native.call("fsi_lbm_subcycle_resident", ctypes.byref(lbm._params), ctypes.byref(lbm._state), lbm.cur,
                        n_sub, flags, _native.ptr(lbm._res_halo), _native.ptr(lbm._res_sync), stream)
        e.record()
        torch.cuda.synchronize()
        lbm.cur = (lbm.cur + n_sub) % 2
        durations.append(s.elapsed_time(e))
    lbm._moments_fresh = False
    lbm.check_resident_waits()
    return sum(durations[1:]) / launches  # the first launch is a warm-up


def _timed(torch, fn):
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    fn()
    e.record()
    torch.cuda.synchronize()
    return s.elapsed_time(e) / 1e3


def _batched_exchanges(torch, ws, out_dir, exchanges):
    """A BatchedFSI of `ws` after its tuning and graph-building exchanges,
    and the seconds of `exchanges` timed exchanges."""
    from mpm_lbm.batch import BatchedFSI

    t0 = time.perf_counter()
    b = BatchedFSI(ws, out_dir=out_dir)
    setup = time.perf_counter() - t0
    b.step()  # eager: builds the block tables, tunes the substep plan
    b.step()  # builds the exchange graph and runs it
    torch.cuda.synchronize()
    return b, setup, _timed(torch, lambda: [b.step() for _ in range(exchanges)])


def batched_form(torch, count, exchanges, f16=True):
    """Batched sweep (seeded mix of 32/48/64/96^3 moving-boundary and penalty
    flap instances): aggregate MLUPS of one-launch-per-operation exchanges
    replayed as a HIP graph, against the same instances run one at a time
    through FSIDriver3D (graph-captured FSI step for moving-boundary instances,
    120 one-LBM-step FSI steps for penalty instances).  With `f16`, the same
    sweep with FP16-deviation storage is timed too (reported beside it)."""
    from mpm_lbm import workloads

    ws = workloads.batch_sweep(count=count, seed=4, out_dir="/tmp/fsi_bench_batch")
    b, setup, secs = _batched_exchanges(torch, ws, "/tmp/fsi_bench_batch/run", exchanges)
    cells = b.total_cells
    batched = cells * b.n_substeps * exchanges / secs / 1e6
    log(f"[bench] batched: {count} instances, {cells} cells, {exchanges} exchanges in {secs * 1e3:.1f} ms "
        f"-> {batched:.1f} MLUPS (set-up {setup:.1f} s)")
    plain_s = 0.0
    for d in b.drivers:
        if d.config.coupling_mode == "moving_boundary":
            d.use_graph = True
            d.step_once()
            d.step_once()
            torch.cuda.synchronize()
            plain_s += _timed(torch, d.step_once)
        else:
            d.step_once()
            torch.cuda.synchronize()
            plain_s += _timed(torch, lambda: [d.step_once() for _ in range(b.n_substeps)])
    plain = cells * b.n_substeps / plain_s / 1e6
    log(f"[bench] plain (one instance at a time): {plain_s * 1e3:.1f} ms per exchange -> {plain:.1f} MLUPS")
    modes = [d.config.coupling_mode for d in b.drivers]
    groups, tuning, n_sub = len(b.subcycle_groups()), b.tuning, b.n_substeps
    b.close()
    del b
    half = None
    if f16:
        ws16 = workloads.batch_sweep(count=count, seed=4, out_dir="/tmp/fsi_bench_batch16", storage="f16")
        b16, _, secs16 = _batched_exchanges(torch, ws16, "/tmp/fsi_bench_batch16/run", exchanges)
        rate16 = cells * b16.n_substeps * exchanges / secs16 / 1e6
        half = {"batched_mlups": round(rate16, 1), "ms_per_exchange": round(secs16 * 1e3 / exchanges, 3),
                "speedup_vs_plain_f32": round(rate16 / plain, 3), "subcycle_groups": len(b16.subcycle_groups()),
                "substep_plan": b16.tuning["plan"] if b16.tuning else b16.substep_plan,
                "error": "FP16-deviation storage: the precision form's bounds against FP64 "
                         "(tests/test_gpu_batched.py::test_gpu_batched_fp16_storage_vs_fp64)"}
        log(f"[bench] batched FP16 storage: {secs16 * 1e3 / exchanges:.1f} ms per exchange -> {rate16:.1f} MLUPS")
        b16.close()
        del b16
    return {
        "instances": count,
        "moving_boundary": modes.count("moving_boundary"),
        "penalty": modes.count("penalty"),
        "total_cells": cells,
        "substeps_per_exchange": n_sub,
        "batched_mlups": round(batched, 1),
        "plain_mlups": round(plain, 1),
        "speedup_vs_plain": round(batched / plain, 3),
        "ms_per_exchange": round(secs * 1e3 / exchanges, 3),
        "substep_plan_tuning": tuning,
        "subcycle_groups": groups,
        "f16": half,
        "equal_error": "same FP32 kernels and per-instance arguments; batched runs checked against the oracle at "
                       "the plain path's bar (tests/test_gpu_batched.py)",
    }


def served_form(torch, rate, seconds):
    """Continuous batching of a Poisson request stream: lattices from
    {32,48,64,96,128}^3, modes 50/50, U(50,500) LBM steps per request rounded
    up to whole 120-substep exchanges, 200 GB admission budget; latency =
    completion - arrival, through admission, build and every exchange."""
    from mpm_lbm.serve import FSIServer, latency_summary, poisson_requests

    reqs = poisson_requests(rate=rate, duration=seconds, seed=5, out_dir="/tmp/fsi_bench_serve")
    offered = sum(r.cells * 120 * r.exchanges for r in reqs) / seconds / 1e6
    server = FSIServer(budget_bytes=200e9, out_dir="/tmp/fsi_bench_serve/run")  # slot table, requests built ahead
    t0 = time.perf_counter()
    server.run(reqs, sync=torch.cuda.synchronize)
    wall = time.perf_counter() - t0
    done = [r for r in reqs if r.done is not None]
    sustained = sum(r.cells * 120 * r.exchanges for r in done) / wall / 1e6
    lat = latency_summary(reqs)
    log(f"[bench] served: {len(done)}/{len(reqs)} requests in {wall:.1f} s, offered {offered:.0f} MLUPS, "
        f"sustained {sustained:.0f} MLUPS, p50 {lat.get('p50_s', 0):.3f} s, p99 {lat.get('p99_s', 0):.3f} s, "
        f"{server.stats()}")
    return {
        "requests": len(reqs),
        "arrival_rate_per_s": rate,
        "arrival_window_s": seconds,
        "load_definition": "request = U(50,500) LBM steps, rounded up to whole exchanges of 120 substeps",
        "offered_mlups": round(offered, 1),
        "sustained_mlups": round(sustained, 1),
        "makespan_s": round(wall, 3),
        "p50_latency_s": round(lat.get("p50_s", float("nan")), 4),
        "p99_latency_s": round(lat.get("p99_s", float("nan")), 4),
        "max_latency_s": round(lat.get("max_s", float("nan")), 4),
        **server.stats(),
    }


def precision_form(torch, long_steps, fp64_steps):
    """192x96x96 duct with a 6x48x36-cell flap (82,944 particles, seed 3),
    regularised collision at tau = 0.55, subcycled moving boundary (120 LBM
    substeps per FSI step), in FP32, FP16- and BF16-deviation storage on the
    GPU: each storage against the FP64 oracle (C++ on the host cores) after
    `fp64_steps` FSI steps, and the 16-bit storages against FP32 after
    `long_steps` FSI steps (17 = 2,040 LBM steps), with the total fluid mass
    drift of every run."""
    import numpy as np

    from mpm_lbm import precision_case
    from mpm_lbm.sim.drivers.fsi_driver import FSIDriver3D
    from oracle.oracle import OracleFSI

    long_steps = max(long_steps, fp64_steps)
    base, config, state, compare = precision_case.setup(long_steps)

    def snap(d):
        return state(d.lbm.rho.to_numpy(), d.lbm.v.to_numpy(), d.lbm.solid.to_numpy(), d.solid.x.to_numpy())

    runs, flags, gpu_ms = {}, None, {}
    for st in ("f32", "f16", "bf16"):
        d = FSIDriver3D(config(st), f"/tmp/fsi_bench_prec/{st}", particle_cloud=base.cloud, write_geometry_file=False)
        d.initialize()
        flags = d.lbm.solid.to_numpy() if flags is None else flags
        s0, mid = snap(d), None
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for k in range(long_steps):
            d.step_once()
            if k + 1 == fp64_steps:
                mid = snap(d)
        e1.record()
        torch.cuda.synchronize()
        gpu_ms[st] = round(e0.elapsed_time(e1) / long_steps, 3)
        runs[st] = (s0, mid, snap(d))
        del d
        torch.cuda.empty_cache()
    out = {"workload": precision_case.WORKLOAD, "lbm_steps_16bit_vs_fp32": 120 * long_steps,
           "ms_per_fsi_step_with_snapshots": gpu_ms}
    ref0, _, ref32 = runs["f32"]
    out["16bit_vs_fp32"] = {st: compare(runs[st][2], ref32, ref0["mass"], runs[st][0]["mass"])
                            for st in ("f16", "bf16")}
    out["fp32_mass_drift"] = (ref32["mass"] - ref0["mass"]) / ref0["mass"]
    if fp64_steps > 0:
        t0 = time.perf_counter()
        o = OracleFSI(config("f32"), flags, base.cloud, base.material, dtype=np.float64)
        f = o.fields()
        o0 = state(f["rho"], f["v"], f["solid"], f["x"])
        for _ in range(fp64_steps):
            o.step_once()
        f = o.fields()
        ref = state(f["rho"], f["v"], f["solid"], f["x"])
        out["lbm_steps_vs_fp64"] = 120 * fp64_steps
        out["vs_fp64"] = {st: compare(r[1], ref, o0["mass"], r[0]["mass"]) for st, r in runs.items()}
        out["fp64_mass_drift"] = (ref["mass"] - o0["mass"]) / o0["mass"]
        out["fp64_oracle_seconds"] = round(time.perf_counter() - t0, 1)
    log("[bench] precision: " + json.dumps(out))
    return out


def _cpu_model():
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(workload, steps):
    """The reference-structured C++/OpenMP oracle on the host cores: one
    untimed warm-up FSI step, then the median of `steps` timed FSI steps."""
    import numpy as np

    from mpm_lbm.sim.geometry.config import GeometryConfig
    from mpm_lbm.sim.geometry.duct_flap_proxy import duct_flap_proxy_static_geometry
    from oracle.oracle import OracleFSI

    flags, _ = duct_flap_proxy_static_geometry(workload.shape, GeometryConfig.from_json(workload.geometry_json))
    o = OracleFSI(workload.config, flags, workload.cloud, workload.material, dtype=np.float32)
    o.step_once()  # warm-up: first-touch of the arrays, thread pool start
    times = []
    for _ in range(steps):
        t0 = time.perf_counter()
        o.step_once()
        times.append(time.perf_counter() - t0)
    dt = float(np.median(times))
    cells = workload.n_cells * workload.config.lbm_substeps_per_fsi_step
    cores = int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))
    return {
        "value": round(cells / dt / 1e6, 3),
        "unit": "MLUPS",
        "cores": cores,
        "kind": "port",
        "cpu_model": _cpu_model(),
        "sample": f"median of {steps} FSI steps after 1 untimed warm-up step, of the same 64^3 workload "
                  f"({workload.config.lbm_substeps_per_fsi_step} LBM + {workload.config.mpm_substeps_per_lbm_step} "
                  f"MPM substeps each); C++/OpenMP reference-structured oracle, FP32; step times "
                  f"{', '.join(f'{t:.2f}' for t in times)} s",
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (the FSI path runs on a single "
              "MI355X; see the sharding design notes)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from mpm_lbm import workloads
    from mpm_lbm.sim.drivers.fsi_driver import FSIDriver3D

    n, sub = args.n, args.substeps
    w = workloads.moving_boundary_flap(n=n, steps=args.warmup + args.steps + 1, substeps=sub, out_dir="/tmp/fsi_bench")
    log(f"[bench] workload {w.name}: {w.config.n_particles} particles, {sub} LBM substeps/FSI step")
    d = FSIDriver3D(w.config, "/tmp/fsi_bench/run", particle_cloud=w.cloud, write_geometry_file=False)
    d.initialize()
    d.step_once()  # one eager FSI step: the flags and walls of the run
    launch_tuning = d.tune_launch()
    log(f"[bench] substep launch tuning: {launch_tuning}")
    for i in range(args.warmup):
        d.step_once()
    torch.cuda.synchronize()
    log(f"[bench] warmup {args.warmup} FSI steps done")
    # the K timed steps run back to back, split into `repeats` blocks each
    # bracketed by events; value = the median block's rate
    nblk = max(1, min(args.repeats, args.steps))
    per = [args.steps // nblk + (1 if b < args.steps % nblk else 0) for b in range(nblk)]
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(nblk + 1)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev[0].record()
    for b, nb in enumerate(per):
        for i in range(nb):
            d.step_once()
        ev[b + 1].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = ev[0].elapsed_time(ev[-1])
    blocks = [{"steps": nb, "ms": round(ev[b].elapsed_time(ev[b + 1]), 4)} for b, nb in enumerate(per)]
    for blk in blocks:
        blk["mlups"] = round(n ** 3 * sub * blk["steps"] / (blk["ms"] * 1e-3) / 1e6, 2)
    med = sorted(blocks, key=lambda blk: blk["mlups"])[len(blocks) // 2]
    mlups = med["mlups"]
    ms_step = med["ms"] / med["steps"]
    secs = ms_step * args.steps / 1e3
    whole = n ** 3 * sub * args.steps / (gpu_ms * 1e-3) / 1e6
    log(f"[bench] timed {args.steps} FSI steps in {nblk} blocks: {gpu_ms:.2f} ms (wall {wall:.3f} s); median block "
        f"{mlups:.1f} MLUPS, whole region {whole:.1f} MLUPS; blocks {[blk['mlups'] for blk in blocks]}")
    row = d.collect_diagnostics(d.current_lbm_step)
    if not all(np.isfinite(v) for v in row.values() if isinstance(v, float)):
        raise RuntimeError("non-finite diagnostics after the timed steps")

    fluid = int(np.count_nonzero(d.lbm.solid.to_numpy() == 0))
    algo_bytes = fluid * 152 + n ** 3 * 4  # per substep
    resident = d.lbm._res_sync is not None and not d.lbm._res_declined
    if not resident:
        raise RuntimeError("the 64^3 subcycle did not take the resident kernel")
    from mpm_lbm import _native

    res_flags = _native.SUB_MOVING_WALL | _native.SUB_GATHER_PREV | _native.SUB_DIAG
    r_ms = time_resident_kernel(torch, d.lbm, args.resident_launches, sub, res_flags)
    achieved = sub * algo_bytes / (r_ms * 1e-3) / 1e9
    log(f"[bench] resident subcycle: {r_ms * 1e3:.1f} us avg over {args.resident_launches} launches = "
        f"{r_ms * 1e3 / sub:.3f} us per substep, {achieved:.0f} GB/s")
    k_ms = time_substep_kernel(torch, d.lbm, args.kernel_launches)
    push_gbs = algo_bytes / (k_ms * 1e-3) / 1e9
    log(f"[bench] launch-per-substep kernel: {k_ms * 1e3:.2f} us avg over {args.kernel_launches} launches, "
        f"{push_gbs:.0f} GB/s")
    copy_gbs = copy_ceiling_gbs(torch)

    def _traffic(path):
        tp = Path(path)
        return json.loads(tp.read_text()).get("hbm_bytes_per_launch") if tp.is_file() else None

    traffic, push_traffic = _traffic(args.traffic_json), _traffic(args.push_traffic_json)

    batched = None
    if args.batch > 0:
        batched = batched_form(torch, args.batch, args.batch_exchanges, f16=args.batch_f16 > 0)

    served = None
    if args.serve_seconds > 0:
        served = served_form(torch, args.serve_rate, args.serve_seconds)

    precision = None
    if args.precision > 0:
        precision = precision_form(torch, args.precision, args.precision_fp64_steps)

    cpu = None
    if args.cpu_steps > 0:
        log("[bench] CPU baseline ...")
        cpu = cpu_baseline(w, args.cpu_steps)
        log(f"[bench] CPU baseline {cpu['value']} MLUPS on {cpu['cores']} threads")

    out = {
        "metric": METRIC,
        "value": round(mlups, 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"plain FP32 moving_boundary FSI, {n}^3 duct + 4x32x24-cell flap, "
                        f"{w.config.n_particles} particles, {sub} LBM substeps + "
                        f"{w.config.mpm_substeps_per_lbm_step} MPM substeps per FSI step",
            "lattice": [n, n, n],
            "timed_blocks": blocks,
            "value_whole_timed_region": round(whole, 2),
            "fluid_cells": fluid,
            "particles": w.config.n_particles,
            "lbm_substeps_per_fsi_step": sub,
            "mfluops": round(fluid * sub * args.steps / secs / 1e6, 2),
            "particle_updates_per_s": round(w.config.n_particles * w.config.mpm_substeps_per_lbm_step * args.steps / secs, 1),
            "copy_ceiling_gbs": round(copy_gbs, 1),
            "resident_us_per_substep": round(r_ms * 1e3 / sub, 3),
            "launch_per_substep_kernel": {
                "kernel": "k_lbm_substep (fused collide+stream+bounce-back+BC, one launch per substep)",
                "avg_launch_us": round(k_ms * 1e3, 3),
                "achieved_gbs": round(push_gbs, 1),
                "frac": round(push_gbs / HBM_PEAK_GBS, 4),
                "traffic": push_traffic,
                "launch_tuning": launch_tuning,
            },
            "batched": batched,
            "served": served,
            "precision": precision,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel": f"k_lbm_resident ({sub} moving-wall substeps per launch, the lattice on chip)",
            "algorithmic_bytes_per_launch": sub * algo_bytes,
            "algorithmic_bytes_per_substep": algo_bytes,
            "avg_launch_us": round(r_ms * 1e3, 3),
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
