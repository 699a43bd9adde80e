"""Serving form: continuous batching of FSI requests on one GPU.

A request is an FSI case (a Workload: moving-boundary or penalty flap of any
lattice size) plus a number of batched exchanges to run; one exchange is one
FSI step of `n_substeps` LBM substeps for a moving-boundary case and
`n_substeps` one-LBM-step FSI steps for a penalty case (the reference's only
penalty schedule).  The server loop, on one device stream:

  1. takes in the requests that have arrived;
  2. admits queued requests first-in first-out while the estimated HBM
     footprint of the admitted set stays within the budget (a request is
     built - geometry, particles, device state - when it is admitted);
  3. regroups the live instances into a BatchedFSI when the set changed;
  4. advances every live instance by one batched exchange;
  5. retires the requests that have run all their exchanges: their results
     go to the host (the caller's `on_done` callback, or the final diagnostics
     row in `Request.result`) and their device state is released, so the
     admission budget bounds what is resident.

A request's results are those of running it alone: batching shares launches,
not state.  Latency is completion time minus arrival time.  No reference
counterpart: the reference runs one case per process.
"""
from __future__ import annotations

import math
import time
from collections import deque
from dataclasses import dataclass, field
from typing import Any, List, Optional

import numpy as np

from . import workloads as _workloads
from .batch import BatchedFSI
from .sim.drivers.fsi_driver import FSIDriver3D

# Device bytes per unit, from the buffers the solvers allocate.
BYTES_PER_CELL = 2 * 19 * 4 + 4 + 12 + 3 + 4 + 4 + 4 + 12 + 12 + 12  # populations, moments, flags, coupling fields
BYTES_PER_CELL_MB = 2 * 18 * 4 + 18 * 4 + 4 * 8  # link scalars, link corrections, fp64 accumulator
BYTES_PER_NODE = 12 + 4 + 12  # grid_v, grid_m, grid_f
BYTES_PER_PARTICLE = 12 + 12 + 36 + 36 + 4 + 4 + 4 + 2 + 12 + 48  # state, masks, pinned x, workspace share


def estimate_bytes(config) -> int:
    """HBM footprint of one instance of `config` (FSIDriverConfig)."""
    shape = tuple(config.lattice_shape) if getattr(config, "lattice_shape", None) else (config.n_grid,) * 3
    cells = int(np.prod(shape))
    nodes = int(np.prod([n + 1 for n in shape]))
    per_cell = BYTES_PER_CELL + (BYTES_PER_CELL_MB if config.coupling_mode == "moving_boundary" else 0)
    return cells * per_cell + nodes * BYTES_PER_NODE + config.n_particles * BYTES_PER_PARTICLE


@dataclass
class Request:
    rid: int
    workload: Any
    exchanges: int
    arrival: float
    admitted: Optional[float] = None
    done: Optional[float] = None
    driver: Any = field(default=None, repr=False)
    remaining: int = 0
    bytes: int = 0
    result: Any = field(default=None, repr=False)

    @property
    def latency(self) -> Optional[float]:
        return None if self.done is None else self.done - self.arrival

    @property
    def cells(self) -> int:
        c = self.workload.config
        shape = tuple(c.lattice_shape) if getattr(c, "lattice_shape", None) else (c.n_grid,) * 3
        return int(np.prod(shape))


class FSIServer:
    def __init__(self, budget_bytes: float = 200e9, native=None, device=None, out_dir="/tmp/fsi_serve",
                 use_graph: bool = True, n_substeps: int = 120):
        self.budget = float(budget_bytes)
        self.n_substeps = int(n_substeps)  # LBM substeps per exchange, whatever the mix
        self.native = native
        self.device = device
        self.out_dir = out_dir
        self.use_graph = use_graph
        self.exchanges_run = 0
        self.regroups = 0

    def _build(self, r: Request):
        w = r.workload
        d = FSIDriver3D(w.config, f"{self.out_dir}/req{r.rid:05d}", particle_cloud=w.cloud, native=self.native,
                        device=self.device, write_geometry_file=False, use_graph=False)
        d.initialize()
        return d

    def run(self, requests: List[Request], sim_exchange_seconds: Optional[float] = None, sync=None, on_done=None):
        """Serve `requests` (arrival times in seconds from the start) to completion.

        A retiring request's driver is handed to on_done(request, driver) (or,
        without a callback, its final diagnostics row is kept in
        request.result) and then dropped, which frees its device buffers.

        With sim_exchange_seconds the clock is simulated - it advances by that
        much per batched exchange and jumps to the next arrival when the
        server is idle - so the schedule is deterministic; otherwise it is the
        wall clock, with `sync()` (e.g. torch.cuda.synchronize) called after
        each exchange so completion times are device completion times."""
        pending = deque(sorted(requests, key=lambda r: (r.arrival, r.rid)))
        queue: deque = deque()
        active: List[Request] = []
        used = 0
        batch = None
        t0 = time.perf_counter()
        sim_now = 0.0

        def now():
            return sim_now if sim_exchange_seconds is not None else time.perf_counter() - t0

        while pending or queue or active:
            t = now()
            while pending and pending[0].arrival <= t:
                queue.append(pending.popleft())
            changed = False
            while queue:
                r = queue[0]
                r.bytes = estimate_bytes(r.workload.config)
                c = r.workload.config
                if c.coupling_mode == "moving_boundary" and c.lbm_substeps_per_fsi_step != self.n_substeps:
                    raise ValueError(f"request {r.rid}: moving-boundary FSI step of {c.lbm_substeps_per_fsi_step} "
                                     f"substeps, the server runs exchanges of {self.n_substeps}")
                if r.bytes > self.budget:
                    raise ValueError(f"request {r.rid} needs {r.bytes} bytes, more than the whole budget")
                if used + r.bytes > self.budget:
                    break
                queue.popleft()
                r.driver = self._build(r)
                r.admitted = now()
                r.remaining = int(r.exchanges)
                active.append(r)
                used += r.bytes
                changed = True
            if not active:
                if pending:
                    wait = pending[0].arrival - now()
                    if sim_exchange_seconds is not None:
                        sim_now = max(sim_now, pending[0].arrival)
                    elif wait > 0:
                        time.sleep(wait)
                continue
            if changed or batch is None:
                batch = BatchedFSI(drivers=[r.driver for r in active], use_graph=self.use_graph,
                                   n_substeps=self.n_substeps)
                self.regroups += 1
            batch.step()
            if sync is not None:
                sync()
            self.exchanges_run += 1
            if sim_exchange_seconds is not None:
                sim_now += sim_exchange_seconds
            t = now()
            still = []
            for r in active:
                r.remaining -= 1
                if r.remaining <= 0:
                    r.done = t
                    used -= r.bytes
                    self._retire(r, on_done)
                else:
                    still.append(r)
            if len(still) != len(active):
                batch = None
            active = still
        return requests

    @staticmethod
    def _retire(r: Request, on_done):
        if on_done is not None:
            on_done(r, r.driver)
        else:
            r.result = r.driver.collect_diagnostics(r.driver.current_lbm_step)
        r.driver = None


def poisson_requests(rate=20.0, duration=10.0, seed=5, sizes=(32, 48, 64, 96, 128), lbm_steps=(50, 500),
                     substeps=120, out_dir="/tmp/fsi_serve_ws") -> List[Request]:
    """Poisson arrivals at `rate` per second over `duration` seconds; each
    request a flap case with a lattice from `sizes`, u_in ~ U(0.02, 0.08),
    E ~ logU(1e4, 1e6), thickness in {2, 4, 6}, modes 50/50, and a length of
    U(lbm_steps) LBM steps rounded up to whole exchanges of `substeps`."""
    rng = np.random.default_rng(seed)
    out, t, k = [], 0.0, 0
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t > duration:
            break
        n = int(rng.choice(sizes))
        mode = "moving_boundary" if rng.random() < 0.5 else "penalty"
        w = _workloads.flap_case(mode, n, substeps=substeps, u_in=float(rng.uniform(0.02, 0.08)),
                                 E_mod=float(10 ** rng.uniform(4.0, 6.0)), thickness=int(rng.choice([2, 4, 6])),
                                 seed=seed * 100000 + k, out_dir=out_dir)
        steps = int(rng.integers(lbm_steps[0], lbm_steps[1] + 1))
        out.append(Request(k, w, max(1, math.ceil(steps / substeps)), t))
        k += 1
    return out


def latency_summary(requests: List[Request]) -> dict:
    lat = np.array([r.latency for r in requests if r.latency is not None])
    if lat.size == 0:
        return {"completed": 0}
    return {
        "completed": int(lat.size),
        "p50_s": float(np.percentile(lat, 50)),
        "p99_s": float(np.percentile(lat, 99)),
        "max_s": float(lat.max()),
        "mean_s": float(lat.mean()),
    }
