"""Serving form (continuous batching) on the serial emulation of the kernels:
a request served in a changing mix gives bit-identically the results of the
same request run alone, admission stays within the HBM budget, first in first
out, and the footprint estimate matches what the solvers allocate."""
import numpy as np
import pytest
import torch

from mpm_lbm import workloads
from mpm_lbm.serve import FSIServer, Request, estimate_bytes, latency_summary
from mpm_lbm.sim.drivers.fsi_driver import FSIDriver3D


def _requests(tmp_path):
    mk = workloads.flap_case
    return [
        Request(0, mk("moving_boundary", 12, substeps=4, thickness=2, seed=11, mpm_substeps=2, out_dir=str(tmp_path)),
                2, 0.0),
        Request(1, mk("penalty", 12, substeps=4, thickness=2, seed=12, out_dir=str(tmp_path)), 1, 0.5),
        Request(2, mk("moving_boundary", 14, substeps=4, thickness=4, seed=13, mpm_substeps=3, E_mod=3e4,
                      out_dir=str(tmp_path)), 2, 1.2),
        Request(3, mk("penalty", 10, substeps=4, thickness=2, seed=14, E_mod=2e5, out_dir=str(tmp_path)), 2, 3.5),
    ]


def _alone(r, emu, tmp_path):
    w = r.workload
    d = FSIDriver3D(w.config, str(tmp_path / f"alone{r.rid}"), particle_cloud=w.cloud, native=emu, device="cpu",
                    write_geometry_file=False, use_graph=False)
    d.initialize()
    steps = r.exchanges if w.config.coupling_mode == "moving_boundary" else r.exchanges * 4
    for _ in range(steps):
        d.step_once()
    return d


def test_served_requests_equal_alone_runs(emu, tmp_path):
    reqs = _requests(tmp_path)
    server = FSIServer(native=emu, device="cpu", out_dir=str(tmp_path / "srv"), n_substeps=4)
    snap = {}

    def keep(r, d):  # the caller's copy of the results; the server then drops the driver
        snap[r.rid] = {("lbm", n): getattr(d.lbm, n).to_numpy() for n in ("F", "rho", "v", "hydro_force", "solid_phi")}
        snap[r.rid].update({("solid", n): getattr(d.solid, n).to_numpy() for n in ("x", "v", "F", "C")})

    server.run(reqs, sim_exchange_seconds=1.0, on_done=keep)
    assert all(r.driver is None for r in reqs)  # retired requests hold no device state
    # schedule: every request done, in a mix that changed as requests came and went
    assert all(r.done is not None for r in reqs)
    assert server.regroups >= 3
    assert reqs[0].admitted == 0.0 and reqs[0].done == 2.0
    for r in reqs:
        assert r.latency >= r.exchanges * 1.0 - 1e-12
    for r in reqs:
        a = _alone(r, emu, tmp_path)
        for (part, name), fb in snap[r.rid].items():
            fa = getattr(getattr(a, part), name).to_numpy()
            assert np.array_equal(fa, fb), (r.rid, part, name)
    s = latency_summary(reqs)
    assert s["completed"] == 4 and s["p50_s"] <= s["p99_s"] <= s["max_s"]


def test_admission_respects_budget_fifo(emu, tmp_path):
    reqs = _requests(tmp_path)
    for r in reqs:
        r.arrival = 0.0
    sizes = [estimate_bytes(r.workload.config) for r in reqs]
    budget = max(sizes) + min(sizes)  # never room for more than two at once
    server = FSIServer(budget_bytes=budget, native=emu, device="cpu", out_dir=str(tmp_path / "srv2"), n_substeps=4)
    server.run(reqs, sim_exchange_seconds=1.0)
    # without a callback each request keeps its final diagnostics row
    assert all(r.driver is None and r.result["step"] > 0 for r in reqs)
    order = sorted(reqs, key=lambda r: (r.admitted, r.rid))
    assert [r.rid for r in order] == [0, 1, 2, 3]  # first in, first out
    events = sorted([(r.admitted, r.bytes) for r in reqs] + [(r.done, -r.bytes) for r in reqs],
                    key=lambda e: (e[0], e[1]))
    used = 0
    for _, b in events:
        used += b
        assert used <= budget
    with pytest.raises(ValueError):
        FSIServer(budget_bytes=min(sizes) - 1, native=emu, device="cpu", n_substeps=4).run(
            [Request(9, reqs[0].workload, 1, 0.0)], sim_exchange_seconds=1.0)


def _device_bytes(obj, seen):
    total = 0
    for v in vars(obj).values():
        items = v if isinstance(v, (list, tuple)) else [v]
        for t in items:
            if isinstance(t, torch.Tensor) and t.data_ptr() not in seen:
                seen.add(t.data_ptr())
                total += t.numel() * t.element_size()
    return total


@pytest.mark.parametrize("mode", ["moving_boundary", "penalty"])
def test_footprint_estimate_matches_allocation(emu, tmp_path, mode):
    w = workloads.flap_case(mode, 16, substeps=4, thickness=2, seed=3, out_dir=str(tmp_path))
    d = FSIDriver3D(w.config, str(tmp_path / "fp"), particle_cloud=w.cloud, native=emu, device="cpu",
                    write_geometry_file=False, use_graph=False)
    d.initialize()
    d.step_once()  # allocates the lazily created buffers (links, accumulator, workspaces)
    seen = set()
    actual = _device_bytes(d.lbm, seen) + _device_bytes(d.solid, seen)
    est = estimate_bytes(w.config)
    assert 0.8 * actual <= est <= 1.25 * actual, (est, actual)
