"""Batched form: many independent FSI instances advanced together.

Each instance is a full FSIDriver3D (its own device buffers, geometry,
particles and coupling constants).  `BatchedFSI` gathers their C-ABI
descriptors into one device table and advances all of them with one launch
per operation (`fsi_batch_launch`): every workgroup finds its instance in a
block map, so instances of different sizes and modes share each launch.  One
batched exchange - 120 LBM substeps for moving-boundary instances, 120
consecutive one-LBM-step FSI steps for penalty / none instances (the
reference rejects penalty with subcycling, fsi_config.py:342-343) - is
captured once as a HIP graph and replayed.

The table is a fixed array of slots.  A fixed batch fills every slot once; a
slot table (`BatchedFSI.slots`, the serving form's) admits and releases
instances in place: the instance table lives at a fixed device address and is
rewritten between exchanges, and the block maps are rebuilt for the live
instances.  The exchange graph is built from its launch list in the native
library (fsi_batch_graph_create, no stream capture), and a change of the
instance set rewrites the block maps and workgroup counts of its kernel nodes
in place (fsi_batch_graph_update) as long as the launch plan - which
operations run, over which instance groups - is the same; one graph is kept
per launch plan.

Moving-boundary instances run their whole subcycle before the particle
passes, so their substeps need no other instance's data.  They are not put in
the per-substep launches with the penalty instances: they run in groups whose
populations fit in the device's last-level cache (`subcycle_group_bytes`),
one group after the other, each through all its substeps.  What a substep
writes is then still in that cache when the next one reads it.  A 96^3
lattice batched with three others costs 46 us per substep, against 34 us
alone or with one other.  In the exchange graph these subcycles and the rest
of the moving-boundary exchange form a side chain that runs beside the
penalty instances' steps (side chain 1 of `FsiBatchOp.branch`): the two touch
no common instance, and the penalty chain's short, latency-bound launches
leave the device room for the substeps.  The penalty instances can be split
further (`penalty_chains`): sets of about equal lattice size, each set's 120
steps on a chain of its own.  On the 64-instance sweep that measured no faster
(164.4, 165.9 and 167.5 ms per exchange for one, two and three sets,
interleaved repeats), so one set is the default.  The resident subcycles run before the fork,
alone on the device, since each needs all its workgroups resident at once.

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
# One copy of the populations per subcycle group, at most.  On MI355X (256 MB
# of last-level cache) a batched substep over one or two 96^3 lattices (67
# and 134 MB) costs 34 us per lattice, over three or four 46 us.  Run one
# after the other, groups of one 96^3 lattice and a few small ones gave the
# shortest exchange of the 64-instance sweep (179 ms; 188 ms at 160 MiB, 192
# ms without groups); beside the penalty chain, 48 to 96 MiB gave the same.
SUBCYCLE_GROUP_BYTES = 72 << 20
Q = 19  # D3Q19 populations per cell
# plan chosen per batch signature (instance shapes, modes, storage), this process
_TUNED = {}
_MODE = {"none": _native.BATCH_MODE_NONE, "penalty": _native.BATCH_MODE_PENALTY,
         "moving_boundary": _native.BATCH_MODE_MOVING_BOUNDARY}
MB = (_native.BATCH_MODE_MOVING_BOUNDARY,)
ONE = (_native.BATCH_MODE_PENALTY, _native.BATCH_MODE_NONE)
PEN = (_native.BATCH_MODE_PENALTY,)

def _full_device(device):
    """`device` with its index (cuda -> cuda:<current>); None -> the GPU if there is one."""
    if device is None:
        device = "cuda" if torch.cuda.is_available() else "cpu"
    device = torch.device(device)
    if device.type == "cuda" and device.index is None:
        device = torch.device("cuda", torch.cuda.current_device())
    return device


def check_batchable(d):
    """Raise unless driver `d` can run in a batched exchange."""
    if d.config.coupling_mode == "moving_boundary" and d.config.reaction_transfer_mode != "engineering":
        raise NotImplementedError("the batched form runs the engineering reaction transfer only "
                                  f"(reaction_transfer_mode={d.config.reaction_transfer_mode!r})")
    if d.config.lbm_open_boundary_semantics != "equilibrium_all_population_reset":
        raise NotImplementedError("the batched form runs the legacy x-face reset only "
                                  f"(lbm_open_boundary_semantics={d.config.lbm_open_boundary_semantics!r})")
    if d.config.coupling_mode == "moving_boundary" and d.config.fsi_exchange_mode != "lbm_subcycled_per_fsi_step":
        raise ValueError("moving-boundary instances of a batch must be subcycled")


class BatchedFSI:
    def __init__(self, workloads=None, out_dir="/tmp/fsi_batch", native=None, device=None, use_graph=True,
                 drivers=None, n_substeps=None, substep_plan=None, capacity=None, resident_lbm=True,
                 subcycle_group_bytes=SUBCYCLE_GROUP_BYTES, penalty_chains=1):
        """Batch `workloads` (a driver is built for each), or batch initialized
        `drivers` that already exist (they keep their state).  n_substeps: LBM
        substeps per exchange (default: the moving-boundary instances' FSI
        step, else 120); moving-boundary instances must use it as their FSI
        step.  substep_plan: how each LBM substep is launched, "split" (one
        launch per coupling mode, the mode's flags compiled in) or "mixed"
        (one launch over every instance, flags read per instance); None tunes
        it on the first exchange (`SUBSTEP_PLANS`).  capacity: slots of the
        table (default: exactly the instances given; more leaves free slots
        for `admit`).  use_graph: run each exchange as a native graph (the
        tuning exchange runs eagerly).  resident_lbm: moving-boundary instances
        whose lattice the resident subcycle covers (FP32, legacy x faces,
        nz = 64, nx and ny multiples of 4, every block resident at once) run
        their subcycle as one LBM_RESIDENT launch each instead of inside the
        batched substep launches.  subcycle_group_bytes: the other
        moving-boundary instances run their subcycles in groups whose
        populations (one copy) take at most this many bytes, one group at a
        time (an instance larger than that is a group of its own); 0 = inside
        the per-substep launches of the penalty instances.  penalty_chains:
        graph chains the penalty / none instances are split over (with
        subcycle groups only; 1 = one chain)."""
        if substep_plan is not None and substep_plan not in SUBSTEP_PLANS:
            raise ValueError(f"substep_plan must be one of {SUBSTEP_PLANS} or None")
        self.substep_plan = substep_plan
        self.tuning = None
        if drivers is not None:
            self.workloads = None
            given = list(drivers)
        else:
            self.workloads = list(workloads or [])
            given = []
            for k, w in enumerate(self.workloads):
                d = FSIDriver3D(w.config, f"{out_dir}/inst{k:03d}", particle_cloud=w.cloud, native=native,
                                device=device, write_geometry_file=False, use_graph=False)
                d.initialize()
                given.append(d)
        if capacity is None and not given:
            raise ValueError("BatchedFSI needs at least one instance")
        self.capacity = int(capacity) if capacity is not None else len(given)
        if self.capacity < max(1, len(given)):
            raise ValueError("capacity below the number of instances given")
        for d in given:
            check_batchable(d)
        if given:
            native, device = given[0].lbm.native, given[0].lbm.device
        if native is None:
            native = _native.load_hip()
        self.native = native
        self.device = _full_device(device)
        mb = [d for d in given if d.config.coupling_mode == "moving_boundary"]
        subs = {d.config.lbm_substeps_per_fsi_step for d in mb}
        if n_substeps is not None:
            subs.add(int(n_substeps))
        if len(subs) > 1:
            raise ValueError("moving-boundary instances of one batch must share lbm_substeps_per_fsi_step")
        self.n_substeps = subs.pop() if subs else 120
        if self.n_substeps % 2:
            raise ValueError("the batched exchange needs an even number of LBM substeps")
        self.drivers = [None] * self.capacity
        self.resident_lbm = bool(resident_lbm)
        self.subcycle_group_bytes = int(subcycle_group_bytes or 0)
        # set c > 0 runs on side chain c + 1 (chain 1 is the subcycle groups')
        if not 1 <= int(penalty_chains) <= _native.BATCH_SIDE_CHAINS:
            raise ValueError(f"penalty_chains must be in 1..{_native.BATCH_SIDE_CHAINS}")
        self.penalty_chains = int(penalty_chains)
        self._res = [None] * self.capacity  # (halo, sync) of resident instances
        self._cus = (torch.cuda.get_device_properties(self.device).multi_processor_count
                     if self.device.type == "cuda" else None)
        self._counters = [None] * self.capacity
        self._react_ws = [None] * self.capacity
        self._boxes = [None] * self.capacity
        self._inst_host = (_native.BatchInstance * self.capacity)()
        nbytes = ctypes.sizeof(self._inst_host)
        self._inst_dev = torch.zeros(nbytes, dtype=torch.uint8, device=self.device)
        self._maps = {}  # key -> [device map, live workgroups]
        self._graphs = {}  # launch plan -> [native graph handle, its launch list]
        self._plan = None
        self._dirty = True
        self.graph_builds = 0  # exchanges run by a newly built graph
        self.graph_updates = 0  # graphs rewritten in place for a new instance set
        self.replays = 0  # exchanges run by a graph built earlier (updated or not)
        for d in given:
            self._fill_slot(self.drivers.index(None), d)
        self.use_graph = bool(use_graph)
        self._recording = None
        self._chain = 0  # graph chain of the launches recorded now (0: main)
        self._set = None  # penalty set the ONE-mode launches issued now cover (None: all)
        self._join = False  # the next recorded main-chain launch waits for the side chain
        self._plan_for = None
        self.exchanges = 0

    @classmethod
    def slots(cls, capacity, native=None, device=None, n_substeps=120, substep_plan="split", use_graph=True):
        """An empty slot table of `capacity` instances (the serving form).  No
        instance runs LBM_RESIDENT, and the moving-boundary instances form one
        subcycle group (on the graph's side chain), so the exchange's launch
        list keeps its shape as requests come and go."""
        return cls(drivers=[], native=native, device=device, n_substeps=n_substeps, substep_plan=substep_plan,
                   use_graph=use_graph, capacity=capacity, resident_lbm=False, subcycle_group_bytes=1 << 62,
                   penalty_chains=1)

    # ----------------------------------------------------------------- slots
    def _fill_slot(self, k, d):
        lbm = d.lbm
        lbm._ensure_ready()  # static link masks for penalty / none instances
        if d.config.coupling_mode == "moving_boundary":
            lbm._ensure_link_buffers()
            lbm.update_dynamic_solid(d.config.dynamic_solid_threshold)
        lbm._sync_moments()
        self._counters[k] = struct_buffer(_native.CouplingCounters, self.device)
        self._react_ws[k] = torch.zeros(max(1, _native.coupled_workspace_doubles(d.solid.n_particles)),
                                        dtype=torch.float64, device=self.device)
        # zero: not valid, the first passes cover the whole lattice / grid
        self._boxes[k] = torch.zeros(_native.BATCH_BOX_INTS, dtype=torch.int32, device=self.device)
        I = self._inst_host[k]
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
        I.boxes = _native.ptr(self._boxes[k])
        I.resident = 0
        if self.resident_lbm and d.config.coupling_mode == "moving_boundary" and self._resident_ok(lbm):
            halo = torch.empty((_native.res_halo_floats(lbm.N),), dtype=torch.float32, device=self.device)
            sync = torch.zeros((_native.res_sync_words(lbm.nx, lbm.ny),), dtype=torch.int32, device=self.device)
            self._res[k] = (halo, sync)
            I.res_halo, I.res_sync, I.resident = _native.ptr(halo), _native.ptr(sync), 1
        lbm._cell_force_active = bool(I.cell_force)
        self.drivers[k] = d
        self._dirty = True

    def admit(self, d) -> int:
        """Put initialized driver `d` into a free slot; returns the slot.  It
        runs from the next exchange on; the other instances are untouched."""
        check_batchable(d)
        if _full_device(d.lbm.device) != self.device:
            raise ValueError("the instance lives on another device than the batch")
        if d.config.coupling_mode == "moving_boundary" and d.config.lbm_substeps_per_fsi_step != self.n_substeps:
            raise ValueError(f"moving-boundary FSI step of {d.config.lbm_substeps_per_fsi_step} substeps, "
                             f"the batch runs exchanges of {self.n_substeps}")
        try:
            k = self.drivers.index(None)
        except ValueError:
            raise RuntimeError(f"all {self.capacity} slots are taken") from None
        self._fill_slot(k, d)
        return k

    def release(self, k):
        """Free slot k (its instance leaves the batch; the driver keeps its state)."""
        if self.drivers[k] is None:
            raise ValueError(f"slot {k} is free")
        self._inst_host[k] = _native.BatchInstance()
        self.drivers[k] = None
        self._counters[k] = None
        self._react_ws[k] = None
        self._boxes[k] = None
        self._res[k] = None
        self._dirty = True

    def _resident_ok(self, lbm):
        """Whether a moving-boundary instance's subcycle can run as LBM_RESIDENT:
        the lattice the resident subcycle covers, with every block (one per CU)
        resident at once."""
        if not lbm._resident_shape_ok():
            return False
        return self._cus is None or _native.res_blocks(lbm.nx, lbm.ny) <= self._cus

    def check_resident_waits(self):
        """Raise if an LBM_RESIDENT launch went on without a neighbour block
        (a wait that timed out: that instance's results are invalid)."""
        for k in self.live_slots:
            if self._res[k] is None:
                continue
            lbm = self.drivers[k].lbm
            sync = self._res[k][1]
            timeouts = int(sync[64 * _native.res_blocks(lbm.nx, lbm.ny)].item())
            if timeouts:
                sync.zero_()
                raise RuntimeError(f"slot {k}: {timeouts} neighbour waits of the resident subcycle timed out")

    def subcycle_groups(self):
        """The live moving-boundary slots not run resident, in subcycle groups:
        first fit, largest lattice first, within subcycle_group_bytes of
        populations per group."""
        if not self.subcycle_group_bytes:
            return []
        items = []
        for k in self.live_slots:
            I = self._inst_host[k]
            if I.mode == _native.BATCH_MODE_MOVING_BOUNDARY and not I.resident:
                size = 4 if I.lp.storage == _native.STORE_F32 else 2
                items.append((Q * size * I.lp.nx * I.lp.ny * I.lp.nz, k))
        groups = []  # [bytes, slots]
        for nbytes, k in sorted(items, key=lambda e: (-e[0], e[1])):
            for g in groups:
                if g[0] + nbytes <= self.subcycle_group_bytes:
                    g[0] += nbytes
                    g[1].append(k)
                    break
            else:
                groups.append([nbytes, [k]])
        return [sorted(g[1]) for g in groups]

    def penalty_sets(self):
        """The live penalty / none slots split into the sets whose steps run on
        chains of their own: largest lattice first, each to the set with the
        fewest cells so far.  One set unless subcycle groups are on."""
        one = [k for k in self.live_slots if self._inst_host[k].mode in ONE]
        n = self.penalty_chains if self.subcycle_group_bytes else 1
        sets = [[0, []] for _ in range(min(n, max(1, len(one))))]
        cells = lambda k: self._inst_host[k].lp.nx * self._inst_host[k].lp.ny * self._inst_host[k].lp.nz
        for k in sorted(one, key=lambda k: (-cells(k), k)):
            s = min(sets, key=lambda e: e[0])
            s[0] += cells(k)
            s[1].append(k)
        return [sorted(s[1]) for s in sets]

    @property
    def live_slots(self):
        return [k for k, d in enumerate(self.drivers) if d is not None]

    def _refresh(self):
        """After admissions / releases: the instance table rewritten in place,
        every known block map rebuilt for the live instances, and the launch
        plan recorded again on the next exchange."""
        if self.device.type == "cuda":
            torch.cuda.synchronize(self.device)  # the last exchange has read the old tables
        raw = torch.frombuffer(bytearray(memoryview(self._inst_host).cast("B")), dtype=torch.uint8)
        self._inst_dev.copy_(raw)
        for key in list(self._maps):
            self._write_map(key)
        self._plan = None
        self._dirty = False

    def _write_map(self, key):
        """Block map of `key` for the live instances, into its device buffer
        (grown when too small: the graphs take the new address on update)."""
        if key[0] == "resident":  # the blocks of one resident instance, in order
            k = key[1]
            d = self.drivers[k]
            n = _native.res_blocks(d.lbm.nx, d.lbm.ny) if d is not None and self._res[k] is not None else 0
            host = np.stack([np.full(n, k, dtype=np.int32), np.arange(n, dtype=np.int32)], axis=1)
            entry = self._maps.get(key)
            if entry is None or entry[0].shape[0] < n:
                entry = [torch.empty((max(n, 1), 2), dtype=torch.int32, device=self.device), 0]
                self._maps[key] = entry
            if n:
                entry[0][:n].copy_(torch.from_numpy(host))
            entry[1] = n
            return
        if key[0] in ("group", "set"):
            if key[0] == "group":  # the substep tiles of one subcycle group's instances
                groups, sub = self.subcycle_groups(), (_native.BATCH_DOMAIN_PLANES, MB, 0)
            else:  # ("set", c, domain, modes, min_mpm): that map cut to penalty set c
                groups, sub = self.penalty_sets(), key[2:]
            members = groups[key[1]] if key[1] < len(groups) else []
            self._write_map(sub)
            table, n = self._maps[sub]
            rows = table[:n]
            host = rows[torch.isin(rows[:, 0], torch.tensor(members, dtype=torch.int32, device=rows.device))]
            n = int(host.shape[0])
            entry = self._maps.get(key)
            if entry is None or entry[0].shape[0] < n:
                entry = [torch.empty((max(n, 1), 2), dtype=torch.int32, device=self.device), 0]
                self._maps[key] = entry
            if n:
                entry[0][:n].copy_(host)
            entry[1] = n
            return
        domain, modes, min_mpm = key
        mask = sum(1 << m for m in modes)
        addr = ctypes.addressof(self._inst_host)
        lib = self.native.lib
        n = lib.fsi_batch_block_map(addr, self.capacity, domain, mask, min_mpm, None)
        if n < 0:
            raise _native.FsiError(lib.fsi_last_error().decode())
        entry = self._maps.get(key)
        if entry is None or entry[0].shape[0] < n:
            entry = [torch.empty((max(n, 1), 2), dtype=torch.int32, device=self.device), 0]
            self._maps[key] = entry
        if n:
            host = np.empty((n, 2), dtype=np.int32)
            lib.fsi_batch_block_map(addr, self.capacity, domain, mask, min_mpm, host.ctypes.data_as(ctypes.c_void_p))
            entry[0][:n].copy_(torch.from_numpy(host))
        entry[1] = n

    def _block_map(self, domain, modes, min_mpm=0):
        """(device map, workgroups) of `domain` over the live instances of
        `modes` with more than min_mpm MPM substeps."""
        key = (domain, tuple(sorted(modes)), min_mpm)
        if key not in self._maps:
            self._write_map(key)
        return self._maps[key]

    # ------------------------------------------------------------ launches
    def _launch(self, op, modes, t=0, min_mpm=0):
        key = (_native.BATCH_OP_DOMAIN[op], tuple(sorted(modes)), min_mpm)
        if self._set is not None and any(m in ONE for m in modes):
            key = ("set", self._set) + key
            if key not in self._maps:
                self._write_map(key)
            table, live = self._maps[key]
        else:
            table, live = self._block_map(*key)
        if live == 0:
            return
        if self._recording is not None:
            self._record(op, t, key)
            return
        self.native.call("fsi_batch_launch", op, _native.ptr(self._inst_dev), self.capacity, _native.ptr(table),
                         live, t, stream_handle(self.device))

    def _launch_resident(self, k):
        """The subcycle of resident slot k as one LBM_RESIDENT launch."""
        key = ("resident", k)
        if key not in self._maps:
            self._write_map(key)
        table, live = self._maps[key]
        if live == 0:
            return
        if self._recording is not None:
            self._record(_native.BATCH_OP_LBM_RESIDENT, 0, key)
            return
        self.native.call("fsi_batch_launch", _native.BATCH_OP_LBM_RESIDENT, _native.ptr(self._inst_dev),
                         self.capacity, _native.ptr(table), live, 0, stream_handle(self.device))

    def _launch_group(self, g, t):
        """Substep t of subcycle group g."""
        key = ("group", g)
        if key not in self._maps:
            self._write_map(key)
        table, live = self._maps[key]
        if live == 0:
            return
        if self._recording is not None:
            self._record(_native.BATCH_OP_LBM_SUBSTEP_MB, t, key)
            return
        self.native.call("fsi_batch_launch", _native.BATCH_OP_LBM_SUBSTEP_MB, _native.ptr(self._inst_dev),
                         self.capacity, _native.ptr(table), live, t, stream_handle(self.device))

    def _record(self, op, t, key):
        """Append a launch to the plan being recorded, on the graph chain
        `_chain` (0: the main chain)."""
        if self._chain:
            branch = self._chain
        elif self._join:
            branch, self._join = _native.BATCH_BRANCH_JOIN, False
        else:
            branch = _native.BATCH_BRANCH_MAIN
        self._recording.append((op, t, key, branch))

    def _clear_counters(self, which, modes):
        """Counters `which` (0 projection, 1 penalty) of the instances of `modes`."""
        self._launch(_native.BATCH_OP_COUNTERS_CLEAR, modes, t=which)

    def _live_inst(self):
        return [self._inst_host[k] for k in self.live_slots]

    def _mpm_substeps(self, modes):
        """MPM substeps with each instance's reaction (fsi_mpm_substep_coupled per instance)."""
        slots = self.live_slots if self._set is None else self.penalty_sets()[self._set]
        max_m = max(self._inst_host[k].mpm_substeps for k in slots if self._inst_host[k].mode in modes)
        for j in range(max_m):
            self._launch(_native.BATCH_OP_BOX_ROLL, modes, t=1, min_mpm=j)
            self._launch(_native.BATCH_OP_MPM_CLEAR, modes, min_mpm=j)
            self._launch(_native.BATCH_OP_MPM_P2G_REACT, modes, min_mpm=j)
            self._launch(_native.BATCH_OP_MPM_GRID_UPDATE_REACT, modes, min_mpm=j)
            self._launch(_native.BATCH_OP_MPM_G2P, modes, min_mpm=j)

    def _project(self, modes):
        self._clear_counters(0, modes)
        self._launch(_native.BATCH_OP_BOX_ROLL, modes, t=0)
        self._launch(_native.BATCH_OP_PROJ_CLEAR, modes)
        self._launch(_native.BATCH_OP_PROJ_PARTICLES, modes)
        self._launch(_native.BATCH_OP_PROJ_NORMALIZE, modes)

    def _issue_exchange(self):
        """One batched exchange: n substeps of every instance (see the module docstring)."""
        live = self._live_inst()
        has_mb = any(I.mode in MB for I in live)
        has_one = any(I.mode in ONE for I in live)
        groups = self.subcycle_groups() if has_mb else []
        self._join = False
        if has_mb:
            self._project(MB)
            self._launch(_native.BATCH_OP_LBM_MB_BEGIN, MB)
            self._launch(_native.BATCH_OP_LBM_REBUILD, MB)
            # resident subcycles alone on the device: each needs all its workgroups at once
            for k in self.live_slots:  # the substep launches leave them out
                if self._inst_host[k].resident:
                    self._launch_resident(k)
            if groups:
                # the rest of the moving-boundary exchange on side chain 1 of the
                # graph, beside the penalty instances' steps (other instances);
                # the launch after those joins it
                self._chain = 1
                for g in range(len(groups)):
                    for t in range(self.n_substeps):
                        self._launch_group(g, t)
                self._mb_exchange_end()
                self._chain = 0
        per_substep = ONE if self.subcycle_group_bytes else tuple(_MODE.values())
        sets = self.penalty_sets() if has_one else [[]]
        for c in reversed(range(len(sets))):
            # penalty set c on side chain c + 1, set 0 on the main chain: the side
            # chains are recorded first, so that each forks where the main chain
            # leaves off (after the resident subcycles)
            self._set, self._chain = (c if len(sets) > 1 else None), (c + 1 if c else 0)
            for t in range(self.n_substeps):
                if has_one:
                    self._project(ONE)
                    self._clear_counters(1, PEN)
                    self._launch(_native.BATCH_OP_PENALTY, PEN, t=t)  # stores the moments it reads (t >= 1)
                self._issue_substep(t, per_substep)
                if has_one:
                    self._mpm_substeps(ONE)
        self._set, self._chain = None, 0
        self._join = bool(groups) or len(sets) > 1
        if has_one:
            # no moments pass between the steps (batch_moment_source): one after the last
            self._launch(_native.BATCH_OP_LBM_MOMENTS, ONE, t=self.n_substeps - 1)
        if has_mb and not groups:
            self._mb_exchange_end()

    def _mb_exchange_end(self):
        """After the moving-boundary subcycles: wall force, moments, MPM substeps."""
        self._launch(_native.BATCH_OP_LBM_MB_FINALIZE, MB)
        self._launch(_native.BATCH_OP_LBM_MB_FINALIZE_FOLD, MB)
        self._launch(_native.BATCH_OP_LBM_MOMENTS, MB, t=self.n_substeps - 1)
        self._mpm_substeps(MB)

    def _issue_substep(self, t, modes):
        """The batched substep t of the instances of `modes`."""
        plan = self.substep_plan
        timer = None
        if plan is None:
            # tuning exchange: the plans take turns (t = 0 runs the moving-boundary
            # instances' first-substep variant and is not timed)
            plan = SUBSTEP_PLANS[t % len(SUBSTEP_PLANS)]
            if t > 0 and self._recording is None:
                timer = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
                timer[0].record()
        if plan == "mixed":
            self._launch(_native.BATCH_OP_LBM_SUBSTEP, modes, t=t)
        else:
            if _native.BATCH_MODE_MOVING_BOUNDARY in modes:
                self._launch(_native.BATCH_OP_LBM_SUBSTEP_MB, (_native.BATCH_MODE_MOVING_BOUNDARY,), t=t)
            self._launch(_native.BATCH_OP_LBM_SUBSTEP_PENALTY, (_native.BATCH_MODE_PENALTY,), t=t)
            self._launch(_native.BATCH_OP_LBM_SUBSTEP_NONE, (_native.BATCH_MODE_NONE,), t=t)
        if timer is not None:
            timer[1].record()
            self._timers.append((plan, timer))

    def _record_plan(self):
        """The launch plan of one exchange - (operation, substep index, block
        map key) per launch, for the live instances - from a dry run that also
        builds every block map it needs."""
        self._recording = []
        try:
            self._issue_exchange()
            return tuple(self._recording)
        finally:
            self._recording = None

    def _launch_list(self):
        """The plan's launches with the current block maps (FsiBatchOp array)."""
        ops = (_native.BatchOp * len(self._plan))()
        for o, (op, t, key, branch) in zip(ops, self._plan):
            o.op, o.t, o.branch = op, t, branch
            if key is not None:
                table, live = self._maps[key]
                o.block_map, o.n_blocks = _native.ptr(table), live
        return ops

    def _run_graph(self):
        """The exchange as a native graph: built once per launch plan, its
        block maps and workgroup counts rewritten in place when they changed."""
        ops = self._launch_list()
        raw = bytes(ops)
        entry = self._graphs.get(self._plan)
        if entry is None:
            handle = ctypes.c_void_p()
            self.native.call("fsi_batch_graph_create", ops, len(ops), _native.ptr(self._inst_dev), self.capacity,
                             ctypes.byref(handle))
            entry = [handle, raw]
            self._graphs[self._plan] = entry
            self.graph_builds += 1
        else:
            if entry[1] != raw:
                if self.device.type == "cuda":
                    torch.cuda.synchronize(self.device)  # no launch of the graph in flight while it changes
                self.native.call("fsi_batch_graph_update", entry[0], ops, len(ops))
                entry[1] = raw
                self.graph_updates += 1
            self.replays += 1
        self.native.call("fsi_batch_graph_launch", entry[0], stream_handle(self.device))

    def close(self):
        """Release the native graphs (also done when the batch is collected)."""
        for handle, _ in self._graphs.values():
            self.native.lib.fsi_batch_graph_destroy(handle)
        self._graphs = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --------------------------------------------------------------- tuning
    def _signature(self):
        """What the substep timings depend on: per instance its lattice, mode,
        storage, collision, particle count and MPM substeps, the exchange length
        and which instances the per-substep launches hold."""
        return (self.n_substeps, self.resident_lbm, self.subcycle_group_bytes) + tuple(
            (I.lp.nx, I.lp.ny, I.lp.nz, int(I.mode), int(I.lp.storage), int(I.lp.collision), int(I.mp.n_particles),
             int(I.mpm_substeps)) for I in self._live_inst())

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

    def _ensure_plan(self):
        if self.substep_plan is None and self.device.type != "cuda":
            self.substep_plan = "split"  # nothing to time on the CPU
        if self._plan is None or self._plan_for != self.substep_plan:
            self._plan = self._record_plan()
            self._plan_for = self.substep_plan

    # ------------------------------------------------------------------- run
    def step(self):
        """Advance every instance by one batched exchange."""
        if self._dirty:
            self._refresh()
        if not self.live_slots:
            return
        if self.use_graph and self.substep_plan is not None:
            self._ensure_plan()
            self._run_graph()
        else:
            # eager: use_graph=False, or the exchange that tunes the substep plan
            self._tune_begin()
            self._issue_exchange()
            self._tune_end()
        self.exchanges += 1
        self._sync_host_state()

    def _sync_host_state(self):
        """Host bookkeeping of each driver after a batched exchange."""
        for k in self.live_slots:
            d, I = self.drivers[k], self._inst_host[k]
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
        return sum(I.lp.nx * I.lp.ny * I.lp.nz for I in self._live_inst())

    def cell_updates_per_exchange(self):
        return self.total_cells * self.n_substeps

    def coupling_stats(self, k):
        from ._device import read_struct

        return read_struct(self._counters[k], _native.CouplingCounters)
