"""Serving form: continuous batching of FSI requests on one GPU.

A request is an FSI case (a Workload: moving-boundary or penalty flap of any
lattice size) plus a number of batched exchanges to run; one exchange is one
FSI step of `n_substeps` LBM substeps for a moving-boundary case and
`n_substeps` one-LBM-step FSI steps for a penalty case (the reference's only
penalty schedule).  All live requests sit in one fixed-capacity slot table
(BatchedFSI.slots) whose instance table and block maps are rewritten in place,
so a change of the live set updates the exchange graph instead of building a
new batch.  The server loop, on one device stream:

  1. takes in the requests that have arrived;
  2. admits queued requests first-in first-out into free slots while the
     estimated HBM footprint of the admitted set stays within the budget;
  3. launches one batched exchange of every live request: a replay of the
     exchange graph of the current launch plan, its block maps rewritten in
     place when the live set changed (a new graph only for a new plan);
  4. while the device runs it, builds the next queued requests (geometry,
     particles, device state: at most `prebuild` built ahead of admission),
     so admission itself only fills a slot;
  5. retires the requests that have run all their exchanges: their results
     go to the host (the caller's `on_done` callback, or the final diagnostics
     row in `Request.result`), their slot is freed and their device state
     released, so the admission budget bounds what is resident (plus the
     requests built ahead).

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
                 use_graph: bool = True, n_substeps: int = 120, capacity: int = 256, prebuild: int = 8,
                 substep_plan: str = "split"):
        self.budget = float(budget_bytes)
        self.n_substeps = int(n_substeps)  # LBM substeps per exchange, whatever the mix
        self.native = native
        self.device = device
        self.out_dir = out_dir
        self.use_graph = use_graph
        self.capacity = int(capacity)  # slots of the table: requests live at once
        self.prebuild = int(prebuild)  # requests built ahead of their admission
        self.substep_plan = substep_plan
        self.table = None
        self.exchanges_run = 0
        self.regroups = 0  # exchanges whose live set differs from the previous exchange's
        self.build_seconds = 0.0  # host time spent building requests
        self.build_inline = 0  # requests built at admission (none ready ahead)

    def _build(self, r: Request):
        t0 = time.perf_counter()
        w = r.workload
        d = FSIDriver3D(w.config, f"{self.out_dir}/req{r.rid:05d}", particle_cloud=w.cloud, native=self.native,
                        device=self.device, write_geometry_file=False, use_graph=False)
        d.initialize()
        self.build_seconds += time.perf_counter() - t0
        return d

    def _table(self, d):
        if self.table is None:
            self.table = BatchedFSI.slots(self.capacity, native=d.lbm.native, device=d.lbm.device,
                                          n_substeps=self.n_substeps, substep_plan=self.substep_plan,
                                          use_graph=self.use_graph)
        return self.table

    def _check(self, r: Request):
        c = r.workload.config
        if c.coupling_mode == "moving_boundary" and c.lbm_substeps_per_fsi_step != self.n_substeps:
            raise ValueError(f"request {r.rid}: moving-boundary FSI step of {c.lbm_substeps_per_fsi_step} "
                             f"substeps, the server runs exchanges of {self.n_substeps}")
        r.bytes = estimate_bytes(c)
        if r.bytes > self.budget:
            raise ValueError(f"request {r.rid} needs {r.bytes} bytes, more than the whole budget")

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
        slot = {}
        used = 0
        t0 = time.perf_counter()
        sim_now = 0.0
        changed = True

        def now():
            return sim_now if sim_exchange_seconds is not None else time.perf_counter() - t0

        def build_ahead(limit=None):
            """Build queued requests ahead of admission (at most `limit` now); returns how many."""
            ready = sum(1 for q in queue if q.driver is not None)
            n = 0
            for q in queue:
                if ready >= self.prebuild or (limit is not None and n >= limit):
                    break
                if q.driver is None:
                    q.driver = self._build(q)
                    ready += 1
                    n += 1
            return n

        def overlap_exchange():
            """While the exchange runs on the device: take in arrivals and build them."""
            take_arrivals()
            build_ahead()
            if sim_exchange_seconds is not None or self.table.device.type != "cuda":
                return
            import torch

            done = torch.cuda.Event()
            done.record()
            while not done.query():
                take_arrivals()
                if not build_ahead(limit=1):
                    time.sleep(2e-4)

        def take_arrivals():
            t = now()
            while pending and pending[0].arrival <= t:
                r = pending.popleft()
                self._check(r)
                queue.append(r)

        while pending or queue or active:
            take_arrivals()
            while queue and len(active) < self.capacity:
                r = queue[0]
                if used + r.bytes > self.budget:
                    break
                queue.popleft()
                if r.driver is None:
                    r.driver = self._build(r)
                    self.build_inline += 1
                slot[r.rid] = self._table(r.driver).admit(r.driver)
                r.admitted = now()
                r.remaining = int(r.exchanges)
                active.append(r)
                used += r.bytes
                changed = True
            if not active:
                build_ahead()
                if pending:
                    wait = pending[0].arrival - now()
                    if sim_exchange_seconds is not None:
                        sim_now = max(sim_now, pending[0].arrival)
                    elif wait > 0:
                        time.sleep(wait)
                continue
            if changed:
                self.regroups += 1
                changed = False
            self.table.step()  # queued on the device
            overlap_exchange()
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
                    self.table.release(slot.pop(r.rid))
                    self._retire(r, on_done)
                    changed = True
                else:
                    still.append(r)
            active = still
        return requests

    def stats(self) -> dict:
        """Exchanges, how the exchange graph was run (built, rewritten in place
        for a new instance set, replayed) and the host time spent building requests."""
        t = self.table
        replays = t.replays if t else 0
        return {
            "exchanges": self.exchanges_run,
            "live_set_changes": self.regroups,
            "graph_builds": t.graph_builds if t else 0,
            "graph_updates": t.graph_updates if t else 0,
            "graph_replays": replays,
            "replay_fraction": round(replays / self.exchanges_run, 4) if self.exchanges_run else 0.0,
            "build_seconds": round(self.build_seconds, 3),
            "built_at_admission": self.build_inline,
        }

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
