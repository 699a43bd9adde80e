"""Batched form: many independent FSI instances advanced together.

Each instance is a full FSIDriver3D (its own device buffers, geometry,
particles and coupling constants).  `BatchedFSI` gathers their C-ABI
descriptors into one device table and advances all of them with one launch
per operation (`fsi_batch_launch`): every workgroup finds its instance in a
block-prefix table, so instances of different sizes and modes share each
launch.  One batched exchange - 120 LBM substeps for moving-boundary
instances, 120 consecutive one-LBM-step FSI steps for penalty / none
instances (the reference rejects penalty with subcycling,
fsi_config.py:342-343) - is captured once as a HIP graph and replayed.

Results are identical to running each instance alone: the kernels and their
per-instance arguments are the same, only the launch is shared.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _native
from ._device import stream_handle, struct_buffer
from .sim.drivers.fsi_driver import FSIDriver3D

TPB = 256
# Launch plans of the batched LBM substep.  Both run the same cell function
# with the same flags per instance, so their results are identical; which is
# faster depends on the mix (one launch per mode compiles the mode's flags in
# but leaves each launch a share of the chip).
SUBSTEP_PLANS = ("split", "mixed")
# plan chosen per batch signature (instance shapes, modes, storage), this process
_TUNED = {}
_MODE = {"none": _native.BATCH_MODE_NONE, "penalty": _native.BATCH_MODE_PENALTY,
         "moving_boundary": _native.BATCH_MODE_MOVING_BOUNDARY}


class BatchedFSI:
    def __init__(self, workloads=None, out_dir="/tmp/fsi_batch", native=None, device=None, use_graph=True,
                 drivers=None, n_substeps=None, substep_plan=None):
        """Batch `workloads` (a driver is built for each), or batch initialized
        `drivers` that already exist (they keep their state; a serving loop
        regroups its live instances this way).  n_substeps: LBM substeps per
        exchange (default: the moving-boundary instances' FSI step, else 120);
        moving-boundary instances must use it as their FSI step.
        substep_plan: how each LBM substep is launched, "split" (one launch
        per coupling mode, the mode's flags compiled in) or "mixed" (one
        launch over every instance, flags read per instance); None tunes it
        on the first exchange (`SUBSTEP_PLANS`)."""
        if substep_plan is not None and substep_plan not in SUBSTEP_PLANS:
            raise ValueError(f"substep_plan must be one of {SUBSTEP_PLANS} or None")
        self.substep_plan = substep_plan
        self.tuning = None
        if drivers is not None:
            self.workloads = None
            self.drivers = list(drivers)
        else:
            if not workloads:
                raise ValueError("BatchedFSI needs at least one workload")
            self.workloads = list(workloads)
            self.drivers = []
            for k, w in enumerate(self.workloads):
                d = FSIDriver3D(w.config, f"{out_dir}/inst{k:03d}", particle_cloud=w.cloud, native=native,
                                device=device, write_geometry_file=False, use_graph=False)
                d.initialize()
                self.drivers.append(d)
        if not self.drivers:
            raise ValueError("BatchedFSI needs at least one instance")
        for d in self.drivers:
            if d.config.coupling_mode == "moving_boundary" and d.config.reaction_transfer_mode != "engineering":
                raise NotImplementedError("the batched form runs the engineering reaction transfer only "
                                          f"(reaction_transfer_mode={d.config.reaction_transfer_mode!r})")
            if d.config.lbm_open_boundary_semantics != "equilibrium_all_population_reset":
                raise NotImplementedError("the batched form runs the legacy x-face reset only "
                                          f"(lbm_open_boundary_semantics={d.config.lbm_open_boundary_semantics!r})")
        self.native = self.drivers[0].lbm.native
        self.device = self.drivers[0].lbm.device
        mb = [d for d in self.drivers if d.config.coupling_mode == "moving_boundary"]
        subs = {d.config.lbm_substeps_per_fsi_step for d in mb}
        if n_substeps is not None:
            subs.add(int(n_substeps))
        if len(subs) > 1:
            raise ValueError("moving-boundary instances of one batch must share lbm_substeps_per_fsi_step")
        self.n_substeps = subs.pop() if subs else 120
        if self.n_substeps % 2:
            raise ValueError("the batched exchange needs an even number of LBM substeps")
        for d in self.drivers:
            if d.config.coupling_mode == "moving_boundary" and d.config.fsi_exchange_mode != "lbm_subcycled_per_fsi_step":
                raise ValueError("moving-boundary instances of a batch must be subcycled")
        self._counters = [struct_buffer(_native.CouplingCounters, self.device) for _ in self.drivers]
        self._react_ws = [torch.zeros(max(1, _native.coupled_workspace_doubles(d.solid.n_particles)),
                                      dtype=torch.float64, device=self.device) for d in self.drivers]
        self._prepare_instances()
        self.use_graph = use_graph and self.device.type == "cuda"
        self._graph = None
        self.exchanges = 0

    # ----------------------------------------------------------------- set-up
    def _prepare_instances(self):
        inst = (_native.BatchInstance * len(self.drivers))()
        for k, d in enumerate(self.drivers):
            lbm = d.lbm
            lbm._ensure_ready()  # static link masks for penalty / none instances
            if d.config.coupling_mode == "moving_boundary":
                lbm._ensure_link_buffers()
                lbm.update_dynamic_solid(d.config.dynamic_solid_threshold)
            lbm._sync_moments()
            I = inst[k]
            I.lp = lbm._params
            I.ls = lbm._state
            I.mp = d.solid._params
            I.ms = d.solid._state
            coupler = d.mb_coupler or d.penalty_coupler or d.projector
            I.cp = coupler.params
            I.cc = _native.ptr(self._counters[k])
            I.mode = _MODE[d.config.coupling_mode]
            I.cur0 = lbm.cur
            I.n_substeps = self.n_substeps if d.config.coupling_mode == "moving_boundary" else 1
            I.mpm_substeps = d.config.mpm_substeps_per_lbm_step
            I.active = 1
            I.cell_force = 1 if d.config.coupling_mode == "penalty" else 0
            I.react_ws = _native.ptr(self._react_ws[k])
            lbm._cell_force_active = bool(I.cell_force)
        self._inst_host = inst
        raw = bytes(memoryview(inst).cast("B"))
        self._inst_dev = torch.frombuffer(bytearray(raw), dtype=torch.uint8).to(self.device)
        self._tables = {}

    def _block_map(self, domain, modes, min_mpm=0):
        """Device block map (instance, workgroup) of `domain` over the instances
        of `modes` with more than min_mpm MPM substeps (fsi_batch_block_map)."""
        key = (domain, tuple(sorted(modes)), min_mpm)
        if key not in self._tables:
            mask = sum(1 << m for m in modes)
            addr = ctypes.addressof(self._inst_host)
            n = self.native.lib.fsi_batch_block_map(addr, len(self.drivers), domain, mask, min_mpm, None)
            if n < 0:
                raise _native.FsiError(self.native.lib.fsi_last_error().decode())
            host = np.zeros((max(n, 1), 2), dtype=np.int32)
            if n:
                self.native.lib.fsi_batch_block_map(addr, len(self.drivers), domain, mask, min_mpm,
                                                    host.ctypes.data_as(ctypes.c_void_p))
            self._tables[key] = (torch.from_numpy(host).to(self.device), n)
        return self._tables[key]

    def _launch(self, op, modes, t=0, min_mpm=0):
        table, nblocks = self._block_map(_native.BATCH_OP_DOMAIN[op], modes, min_mpm)
        if nblocks == 0:
            return
        self.native.call("fsi_batch_launch", op, _native.ptr(self._inst_dev), len(self.drivers), _native.ptr(table),
                         nblocks, t, stream_handle(self.device))

    def _clear_counters(self, which):
        self.native.call("fsi_batch_launch", _native.BATCH_OP_COUNTERS_CLEAR, _native.ptr(self._inst_dev),
                         len(self.drivers), None, 0, which, stream_handle(self.device))

    def _mpm_substeps(self, modes):
        """MPM substeps with each instance's reaction (fsi_mpm_substep_coupled per instance)."""
        max_m = max(I.mpm_substeps for I in self._inst_host if I.mode in modes)
        for j in range(max_m):
            self._launch(_native.BATCH_OP_MPM_CLEAR, modes, min_mpm=j)
            self._launch(_native.BATCH_OP_MPM_P2G_REACT, modes, min_mpm=j)
            self._launch(_native.BATCH_OP_MPM_GRID_UPDATE_REACT, modes, min_mpm=j)
            self._launch(_native.BATCH_OP_MPM_G2P, modes, min_mpm=j)

    def _project(self, modes):
        self._clear_counters(0)
        self._launch(_native.BATCH_OP_PROJ_CLEAR, modes)
        self._launch(_native.BATCH_OP_PROJ_PARTICLES, modes)
        self._launch(_native.BATCH_OP_PROJ_NORMALIZE, modes)

    def _issue_exchange(self):
        """One batched exchange: n substeps of every instance (see the module docstring)."""
        MB = (_native.BATCH_MODE_MOVING_BOUNDARY,)
        ONE = (_native.BATCH_MODE_PENALTY, _native.BATCH_MODE_NONE)
        PEN = (_native.BATCH_MODE_PENALTY,)
        has_mb = any(I.mode in MB for I in self._inst_host)
        has_one = any(I.mode in ONE for I in self._inst_host)
        if has_mb:
            self._project(MB)
            self._launch(_native.BATCH_OP_LBM_MB_BEGIN, MB)
            self._launch(_native.BATCH_OP_LBM_REBUILD, MB)
        for t in range(self.n_substeps):
            if has_one:
                self._project(ONE)
                self._clear_counters(1)
                self._launch(_native.BATCH_OP_PENALTY, PEN)
            self._issue_substep(t)
            if has_one:
                self._launch(_native.BATCH_OP_LBM_MOMENTS, ONE, t=t)
                self._mpm_substeps(ONE)
        if has_mb:
            self._launch(_native.BATCH_OP_LBM_MB_FINALIZE, MB)
            self._launch(_native.BATCH_OP_LBM_MB_FINALIZE_FOLD, MB)
            self._launch(_native.BATCH_OP_LBM_MOMENTS, MB, t=self.n_substeps - 1)
            self._mpm_substeps(MB)

    def _issue_substep(self, t):
        plan = self.substep_plan
        timer = None
        if plan is None:
            # tuning exchange: the plans take turns (t = 0 runs the moving-boundary
            # instances' first-substep variant and is not timed)
            plan = SUBSTEP_PLANS[t % len(SUBSTEP_PLANS)]
            if t > 0:
                timer = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
                timer[0].record()
        if plan == "mixed":
            self._launch(_native.BATCH_OP_LBM_SUBSTEP, tuple(_MODE.values()), t=t)
        else:
            self._launch(_native.BATCH_OP_LBM_SUBSTEP_MB, (_native.BATCH_MODE_MOVING_BOUNDARY,), t=t)
            self._launch(_native.BATCH_OP_LBM_SUBSTEP_PENALTY, (_native.BATCH_MODE_PENALTY,), t=t)
            self._launch(_native.BATCH_OP_LBM_SUBSTEP_NONE, (_native.BATCH_MODE_NONE,), t=t)
        if timer is not None:
            timer[1].record()
            self._timers.append((plan, timer))

    def _signature(self):
        """What the substep timings depend on: per instance its lattice, mode,
        storage, collision, particle count and MPM substeps, and the exchange length."""
        return (self.n_substeps,) + tuple(
            (I.lp.nx, I.lp.ny, I.lp.nz, int(I.mode), int(I.lp.storage), int(I.lp.collision), int(I.mp.n_particles),
             int(I.mpm_substeps)) for I in self._inst_host)

    def _tune_begin(self):
        """Before an exchange: take a plan chosen earlier for the same batch, or
        tune on this exchange (GPU only; on the CPU the plan is "split")."""
        if self.substep_plan is not None:
            return
        key = self._signature()
        if key in _TUNED:
            self.substep_plan = _TUNED[key]
            self.tuning = {"plan": self.substep_plan, "cached": True}
        elif self.device.type != "cuda":
            self.substep_plan = "split"
        else:
            self._timers = []

    def _tune_end(self):
        if self.substep_plan is not None:
            return
        torch.cuda.synchronize(self.device)
        per = {p: [] for p in SUBSTEP_PLANS}
        for plan, (a, b) in self._timers:
            per[plan].append(a.elapsed_time(b) * 1000.0)
        self._timers = []
        mean = {p: float(np.mean(v)) if v else float("inf") for p, v in per.items()}
        self.substep_plan = min(SUBSTEP_PLANS, key=lambda p: mean[p])
        _TUNED[self._signature()] = self.substep_plan
        self.tuning = {"plan": self.substep_plan, "cached": False,
                       "us_per_substep": {p: round(v, 2) for p, v in mean.items()},
                       "samples": {p: len(v) for p, v in per.items()}}

    # ------------------------------------------------------------------- run
    def step(self):
        """Advance every instance by one batched exchange."""
        if self.use_graph and self.exchanges > 0:
            # the first exchange runs eagerly: it builds the block tables
            # (host-to-device copies) that the captured launches then read
            if self._graph is None:
                torch.cuda.synchronize(self.device)
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g):
                    self._issue_exchange()
                self._graph = g
            self._graph.replay()
        else:
            self._tune_begin()
            self._issue_exchange()
            self._tune_end()
        self.exchanges += 1
        self._sync_host_state()

    def _sync_host_state(self):
        """Host bookkeeping of each driver after a batched exchange."""
        for d, I in zip(self.drivers, self._inst_host):
            lbm = d.lbm
            lbm.cur = I.cur0
            lbm._moments_fresh = True
            lbm._pending = None
            lbm._hydro_from_links = False
            if I.mode == _native.BATCH_MODE_MOVING_BOUNDARY:
                lbm.mb_sample_count = self.n_substeps
                d.current_lbm_step += 1
                d.total_lbm_substeps += self.n_substeps
                d.total_mpm_substeps += I.mpm_substeps
                lbm._rebuilt_phi_version = None
            else:
                d.current_lbm_step += self.n_substeps
                d.total_lbm_substeps += self.n_substeps
                d.total_mpm_substeps += self.n_substeps * I.mpm_substeps

    @property
    def total_cells(self):
        return sum(I.lp.nx * I.lp.ny * I.lp.nz for I in self._inst_host)

    def cell_updates_per_exchange(self):
        return self.total_cells * self.n_substeps

    def coupling_stats(self, k):
        from ._device import read_struct

        return read_struct(self._counters[k], _native.CouplingCounters)
