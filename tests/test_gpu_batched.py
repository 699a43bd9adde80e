"""Batched form on the GPU: a mixed batch (moving-boundary and penalty
instances of different sizes and MPM substep counts) advanced with one launch
per operation, the exchange replayed as a HIP graph, each instance checked
against its own oracle run to the single-instance parity bar."""
import numpy as np
import pytest

from mpm_lbm import _native, workloads
from mpm_lbm.batch import BatchedFSI
from oracle.oracle import OracleFSI
from tests.parity import check_history, field_errors

pytestmark = pytest.mark.gpu
TOL = 1.0e-5


def _mixed(tmp_path, storage="f32"):
    kw = {"out_dir": str(tmp_path), "storage": storage}
    return [
        workloads.flap_case("moving_boundary", 24, substeps=8, thickness=2, seed=1, mpm_substeps=2, **kw),
        workloads.flap_case("penalty", 16, substeps=8, thickness=2, seed=2, **kw),
        workloads.flap_case("moving_boundary", 16, substeps=8, thickness=4, seed=3, mpm_substeps=3, E_mod=3e4, **kw),
        workloads.flap_case("penalty", 20, substeps=8, thickness=4, seed=4, E_mod=2e5, **kw),
        # coupling "none": the split plan's NONE launch
        workloads.flap_case("none", 12, substeps=8, thickness=2, seed=5, **kw),
    ]


def test_gpu_batched_matches_oracle(hip, tmp_path):
    ws = _mixed(tmp_path)
    batch = BatchedFSI(ws, out_dir=str(tmp_path / "batch"), native=hip, penalty_chains=2)
    oracles = []
    for w, d in zip(ws, batch.drivers):
        flags = d.lbm.static_solid.to_numpy()
        o = OracleFSI(w.config, flags, w.cloud, w.material, dtype=np.float32)
        o.fp64 = OracleFSI(w.config, flags, w.cloud, w.material, dtype=np.float64)
        oracles.append(o)
    n_exchanges = 3  # eager (tunes the substep launch plan), capture + replay, replay
    for _ in range(n_exchanges):
        batch.step()
    from mpm_lbm.batch import SUBSTEP_PLANS

    print("substep plan tuning", batch.tuning, flush=True)
    assert batch.substep_plan in SUBSTEP_PLANS and batch.tuning["plan"] == batch.substep_plan
    # the replayed graphs ran the moving-boundary subcycles on side chain 1,
    # and the penalty / none instances in two sets on the main chain and side chain 2
    assert batch.subcycle_groups() == [[0, 2]] and batch.replays + batch.graph_builds == n_exchanges - 1
    assert len(batch.penalty_sets()) == 2
    assert {br for *_, br in batch._plan} == {_native.BATCH_BRANCH_MAIN, 1, 2, _native.BATCH_BRANCH_JOIN}
    for k, (d, o) in enumerate(zip(batch.drivers, oracles)):
        steps = n_exchanges if d.config.coupling_mode == "moving_boundary" else n_exchanges * batch.n_substeps
        for _ in range(steps):
            o.step_once()
            o.fp64.step_once()
        errs = field_errors(d, o)
        print("batched instance", k, d.config.coupling_mode, errs, flush=True)
        check_history([errs], f"batched instance {k} ({d.config.coupling_mode})", TOL)


def test_gpu_batched_fp16_storage_vs_fp64(hip, tmp_path):
    """The mixed batch with FP16-deviation storage (the precision form's
    storage inside the batched launches and the graph's side chain): each
    instance against its own FP64 oracle run, to the FP16 bounds of the
    precision form."""
    from tests.test_precision_forms import BOUNDS

    ws = _mixed(tmp_path, storage="f16")
    batch = BatchedFSI(ws, out_dir=str(tmp_path / "batch16"), native=hip)
    assert {int(I.lp.storage) for I in batch._live_inst()} == {_native.STORE_F16_DEV}
    oracles = [OracleFSI(w.config, d.lbm.static_solid.to_numpy(), w.cloud, w.material, dtype=np.float64)
               for w, d in zip(ws, batch.drivers)]
    n_exchanges = 3
    for _ in range(n_exchanges):
        batch.step()
    assert batch.subcycle_groups() == [[0, 2]] and batch.graph_builds == 1
    for k, (d, o) in enumerate(zip(batch.drivers, oracles)):
        steps = n_exchanges if d.config.coupling_mode == "moving_boundary" else n_exchanges * batch.n_substeps
        for _ in range(steps):
            o.step_once()
        errs = field_errors(d, o)
        print("batched f16 instance", k, d.config.coupling_mode, {q: errs[q] for q in BOUNDS["f16"]}, flush=True)
        assert errs["solid_equal"], k
        for q, bound in BOUNDS["f16"].items():
            assert errs[q] <= bound, (k, d.config.coupling_mode, q, errs[q])


def test_gpu_slot_table_replays_across_admissions(hip, tmp_path):
    """Serving-form slot table on the GPU: instances admitted and released
    between exchanges while one exchange graph is replayed, its block maps
    and workgroup counts rewritten in place (the launch plan is unchanged);
    each instance against its own oracle run for the exchanges it took part in."""
    from mpm_lbm.sim.drivers.fsi_driver import FSIDriver3D

    ws = _mixed(tmp_path)[:4]
    drivers, oracles = [], []
    for k, w in enumerate(ws):
        d = FSIDriver3D(w.config, str(tmp_path / f"slot{k}"), particle_cloud=w.cloud, native=hip,
                        write_geometry_file=False, use_graph=False)
        d.initialize()
        flags = d.lbm.static_solid.to_numpy()
        o = OracleFSI(w.config, flags, w.cloud, w.material, dtype=np.float32)
        o.fp64 = OracleFSI(w.config, flags, w.cloud, w.material, dtype=np.float64)
        drivers.append(d)
        oracles.append(o)
    table = BatchedFSI.slots(8, native=hip, n_substeps=8)
    runs = [0] * 4
    # MB (2 MPM substeps), penalty 20^3, MB (3)
    slot = {k: table.admit(drivers[k]) for k in (0, 3, 2)}
    table.step()
    for k in (0, 2, 3):
        runs[k] += 1
    slot[1] = table.admit(drivers[1])  # a second, smaller penalty instance: same launch plan
    table.step()
    for k in (0, 1, 2, 3):
        runs[k] += 1
    table.release(slot.pop(3))
    table.step()
    for k in (0, 1, 2):
        runs[k] += 1
    print("slot table", {"builds": table.graph_builds, "updates": table.graph_updates, "replays": table.replays},
          flush=True)
    # one launch plan throughout: one graph, rewritten in place for each new instance set
    assert table.graph_builds == 1 and table.replays == 2 and table.graph_updates == 2
    for k, (d, o) in enumerate(zip(drivers, oracles)):
        steps = runs[k] if d.config.coupling_mode == "moving_boundary" else runs[k] * 8
        for _ in range(steps):
            o.step_once()
            o.fp64.step_once()
        errs = field_errors(d, o)
        print("slot instance", k, d.config.coupling_mode, runs[k], errs, flush=True)
        check_history([errs], f"slot instance {k} ({d.config.coupling_mode})", TOL)
