"""Batched form: one launch per operation for many instances of mixed sizes
and modes gives the same results as running each instance alone (checked on
the serial emulation of the kernels; the GPU version is in test_gpu_batched)."""
import numpy as np
import pytest

from mpm_lbm import _native, workloads
from mpm_lbm.batch import BatchedFSI
from mpm_lbm.sim.drivers.fsi_driver import FSIDriver3D
from tests.parity import rel_max_err


def _mixed(tmp_path, storage="f32"):
    return [
        workloads.flap_case("moving_boundary", 16, substeps=6, thickness=2, seed=1, mpm_substeps=2,
                            out_dir=str(tmp_path), storage=storage),
        workloads.flap_case("penalty", 12, substeps=6, thickness=2, seed=2, out_dir=str(tmp_path), storage=storage),
        workloads.flap_case("moving_boundary", 12, substeps=6, thickness=4, seed=3, mpm_substeps=3, E_mod=3e4,
                            out_dir=str(tmp_path), storage=storage),
        workloads.flap_case("penalty", 16, substeps=6, thickness=4, seed=4, E_mod=2e5, out_dir=str(tmp_path),
                            storage=storage),
    ]


def _alone(w, tmp_path, emu, k):
    d = FSIDriver3D(w.config, str(tmp_path / f"alone{k}"), particle_cloud=w.cloud, native=emu, device="cpu",
                    write_geometry_file=False, use_graph=False)
    d.initialize()
    return d


def test_batched_equals_alone(emu, tmp_path):
    ws = _mixed(tmp_path)
    batch = BatchedFSI(ws, out_dir=str(tmp_path / "batch"), native=emu, device="cpu")
    alone = [_alone(w, tmp_path, emu, k) for k, w in enumerate(ws)]
    for _ in range(2):
        batch.step()
        for d in alone:
            steps = 1 if d.config.coupling_mode == "moving_boundary" else batch.n_substeps
            for _ in range(steps):
                d.step_once()
    for k, (b, a) in enumerate(zip(batch.drivers, alone)):
        assert b.total_lbm_substeps == a.total_lbm_substeps, k
        assert b.total_mpm_substeps == a.total_mpm_substeps, k
        for name, fb, fa in (("F", b.lbm.F, a.lbm.F), ("rho", b.lbm.rho, a.lbm.rho), ("v", b.lbm.v, a.lbm.v),
                             ("hydro", b.lbm.hydro_force, a.lbm.hydro_force), ("x", b.solid.x, a.solid.x),
                             ("pv", b.solid.v, a.solid.v), ("Fp", b.solid.F, a.solid.F)):
            err = rel_max_err(fb.to_numpy(), fa.to_numpy())
            assert err == 0.0, (k, name, err)  # same kernels, same order: bit-identical
        assert np.array_equal(b.lbm.solid.to_numpy(), a.lbm.solid.to_numpy())
        # the projection counters of the last projection: a moving-boundary
        # instance's survive the penalty instances' projections after it
        got, want = batch.coupling_stats(k), a.projector._counters()
        assert got.active_cell_count == want.active_cell_count > 0, k
        assert got.max_phi_raw_bits == want.max_phi_raw_bits, k
        assert got.projected_mass == pytest.approx(want.projected_mass, rel=1e-12), k


def test_batched_resident_subcycle_equals_alone(emu, tmp_path):
    """A moving-boundary instance the resident subcycle covers (16 x 12 x 64)
    runs its subcycle as one LBM_RESIDENT launch; the batch still equals each
    instance run alone, and the same batch without that launch, bit for bit."""
    from mpm_lbm import _native

    ws = [workloads.moving_boundary_flap(shape=(16, 12, 64), steps=2, substeps=6, flap=(2, 6, 6), mpm_substeps=2,
                                         out_dir=str(tmp_path)),
          *_mixed(tmp_path)[:2]]
    batch = BatchedFSI(ws, out_dir=str(tmp_path / "batch"), native=emu, device="cpu")
    flat = BatchedFSI(ws, out_dir=str(tmp_path / "flat"), native=emu, device="cpu", resident_lbm=False)
    assert [I.resident for I in batch._live_inst()] == [1, 0, 0]
    assert [I.resident for I in flat._live_inst()] == [0, 0, 0]
    alone = [_alone(w, tmp_path, emu, k) for k, w in enumerate(ws)]
    for _ in range(2):
        batch.step()
        flat.step()
        for d in alone:
            steps = 1 if d.config.coupling_mode == "moving_boundary" else batch.n_substeps
            for _ in range(steps):
                d.step_once()
    batch.check_resident_waits()
    assert batch._maps[("resident", 0)][1] == _native.res_blocks(16, 12)
    # the resident instance has no workgroup in the substep launches
    planes = [v for key, v in batch._maps.items() if key[0] == _native.BATCH_DOMAIN_PLANES]
    assert planes and all(0 not in v[0][:v[1], 0].tolist() for v in planes)
    for k in range(3):
        for other in (flat.drivers[k], alone[k]):
            b = batch.drivers[k]
            for name, fb, fo in (("F", b.lbm.F, other.lbm.F), ("rho", b.lbm.rho, other.lbm.rho),
                                 ("v", b.lbm.v, other.lbm.v), ("hydro", b.lbm.hydro_force, other.lbm.hydro_force),
                                 ("x", b.solid.x, other.solid.x), ("pv", b.solid.v, other.solid.v)):
                assert rel_max_err(fb.to_numpy(), fo.to_numpy()) == 0.0, (k, name)
            assert np.array_equal(b.lbm.solid.to_numpy(), other.lbm.solid.to_numpy())


def _expected_map(inst, domain, modes, min_mpm):
    """Block map restated in Python: (instance, workgroup) pairs, instances in index order."""
    from mpm_lbm import _native

    rows = []
    for k, I in enumerate(inst):
        if not (I.active and I.mode in modes and I.mpm_substeps > min_mpm):
            continue
        if domain == _native.BATCH_DOMAIN_INSTANCES:
            nb = 1
        elif domain == _native.BATCH_DOMAIN_BOX:
            nb = min(_native.BATCH_BOX_BLOCKS, -(-(I.lp.nx * I.lp.ny * I.lp.nz) // 256))
        elif domain == _native.BATCH_DOMAIN_PLANES:
            nb = I.lp.nx * (-(-(I.lp.ny * I.lp.nz) // 256))
        elif domain == _native.BATCH_DOMAIN_CELLS:
            nb = -(-(I.lp.nx * I.lp.ny * I.lp.nz) // 256)
        elif domain == _native.BATCH_DOMAIN_NODES:
            nb = -(-(I.mp.gnx * I.mp.gny * I.mp.gnz) // 256)
        else:
            nb = -(-I.mp.n_particles // _native.PARTICLE_BLOCK)
        rows += [(k, b) for b in range(nb)]
    return rows


def test_block_map_helper_matches_python(emu, tmp_path):
    import ctypes

    from mpm_lbm import _native

    ws = _mixed(tmp_path)
    batch = BatchedFSI(ws, out_dir=str(tmp_path / "batch2"), native=emu, device="cpu")
    batch._inst_host[2].mpm_substeps = 1  # exercise the min_mpm filter
    addr = ctypes.addressof(batch._inst_host)
    for domain in (_native.BATCH_DOMAIN_CELLS, _native.BATCH_DOMAIN_NODES, _native.BATCH_DOMAIN_PARTICLES,
                   _native.BATCH_DOMAIN_PLANES, _native.BATCH_DOMAIN_INSTANCES, _native.BATCH_DOMAIN_BOX):
        for modes in ((_native.BATCH_MODE_MOVING_BOUNDARY,), (_native.BATCH_MODE_PENALTY, _native.BATCH_MODE_NONE),
                      (0, 1, 2)):
            for min_mpm in (0, 1):
                mask = sum(1 << m for m in modes)
                n = emu.lib.fsi_batch_block_map(addr, len(ws), domain, mask, min_mpm, None)
                out = np.zeros((max(n, 1), 2), dtype=np.int32)
                emu.lib.fsi_batch_block_map(addr, len(ws), domain, mask, min_mpm, out.ctypes.data_as(ctypes.c_void_p))
                got = [tuple(r) for r in out[:n].tolist()]
                assert got == _expected_map(batch._inst_host, domain, modes, min_mpm), (domain, modes, min_mpm)


def test_substep_launch_plans_agree(emu, tmp_path):
    """The two launch plans of the batched substep (one launch per coupling
    mode, flags compiled in; or one launch over every instance, flags read
    per instance) give bit-identical instances; an unknown plan is refused."""
    ws = _mixed(tmp_path)
    runs = {}
    for plan in ("split", "mixed"):
        b = BatchedFSI(ws, out_dir=str(tmp_path / plan), native=emu, device="cpu", substep_plan=plan)
        b.step()
        b.step()
        runs[plan] = b
    for da, db in zip(runs["split"].drivers, runs["mixed"].drivers):
        for name in ("F", "rho", "v", "hydro_force", "cell_force", "solid_phi"):
            assert np.array_equal(getattr(da.lbm, name).to_numpy(), getattr(db.lbm, name).to_numpy()), name
        for name in ("x", "v", "F", "C"):
            assert np.array_equal(getattr(da.solid, name).to_numpy(), getattr(db.solid, name).to_numpy()), name
    with pytest.raises(ValueError, match="substep_plan"):
        BatchedFSI(ws, out_dir=str(tmp_path / "bad"), native=emu, device="cpu", substep_plan="fast")


def test_slot_table_admit_release_in_place(emu, tmp_path):
    """A slot table admits and releases instances between exchanges: the
    instance table keeps its address, the exchange graph of an unchanged
    launch plan is rewritten in place (block maps, workgroup counts) instead
    of built again, a new plan gets a graph of its own, and every instance
    still matches its run alone bit for bit."""
    ws = _mixed(tmp_path)
    drivers = [_alone(w, tmp_path / "slot", emu, k) for k, w in enumerate(ws)]
    alone = [_alone(w, tmp_path, emu, k) for k, w in enumerate(ws)]
    table = BatchedFSI.slots(8, native=emu, device="cpu", n_substeps=6)
    runs = [0] * len(ws)
    s0 = table.admit(drivers[0])
    s1 = table.admit(drivers[1])
    table.step()
    runs[0] += 1
    runs[1] += 1
    plan = table._plan
    inst_addr = table._inst_dev.data_ptr()
    s3 = table.admit(drivers[3])  # another penalty instance: same plan, the graph is updated
    table.step()
    for k in (0, 1, 3):
        runs[k] += 1
    assert table._plan == plan and (table.graph_builds, table.graph_updates, table.replays) == (1, 1, 1)
    table.release(s0)
    table.release(s1)
    s2 = table.admit(drivers[2])  # moving-boundary with 3 MPM substeps: a new plan, a new graph
    assert s2 in (s0, s1)
    table.step()
    runs[2] += 1
    runs[3] += 1
    assert table._plan != plan and table.graph_builds == 2
    assert table._inst_dev.data_ptr() == inst_addr
    table.release(s3)
    table.release(s2)
    assert table.live_slots == []
    table.step()  # nothing live: nothing runs
    table.close()
    for k, (b, a) in enumerate(zip(drivers, alone)):
        steps = runs[k] if a.config.coupling_mode == "moving_boundary" else runs[k] * 6
        for _ in range(steps):
            a.step_once()
        assert b.total_lbm_substeps == a.total_lbm_substeps, k
        for name in ("F", "rho", "v", "hydro_force"):
            assert np.array_equal(getattr(b.lbm, name).to_numpy(), getattr(a.lbm, name).to_numpy()), (k, name)
        for name in ("x", "v", "F", "C"):
            assert np.array_equal(getattr(b.solid, name).to_numpy(), getattr(a.solid, name).to_numpy()), (k, name)


def test_batch_graph_rejects_other_launch_lists(emu, tmp_path):
    """fsi_batch_graph_update takes only the launch list the graph was built
    from (same operations and substep indices); an operation without
    workgroups is refused at build."""
    import ctypes

    ws = _mixed(tmp_path)[:2]
    batch = BatchedFSI(ws, out_dir=str(tmp_path / "g"), native=emu, device="cpu", substep_plan="split")
    batch.step()
    handle, _ = next(iter(batch._graphs.values()))
    ops = batch._launch_list()
    other = (_native.BatchOp * len(ops))(*ops)
    other[3].t = other[3].t + 1
    with pytest.raises(_native.FsiError, match="other operations"):
        emu.call("fsi_batch_graph_update", handle, other, len(other))
    with pytest.raises(_native.FsiError, match="number of operations"):
        emu.call("fsi_batch_graph_update", handle, ops, len(ops) - 1)
    empty = (_native.BatchOp * 1)()
    empty[0].op = _native.BATCH_OP_PENALTY
    h = ctypes.c_void_p()
    with pytest.raises(_native.FsiError, match="without a block map"):
        emu.call("fsi_batch_graph_create", empty, 1, _native.ptr(batch._inst_dev), batch.capacity, ctypes.byref(h))
    batch.close()


def test_subcycle_groups_and_chains_equal_flat(emu, tmp_path):
    """Moving-boundary instances run their subcycles in groups (here one
    instance per group, and all in one group) outside the per-substep
    launches of the penalty instances, on side chain 1 of the exchange graph;
    the penalty instances are split over the main chain and side chain 2 (or
    all on the main chain).  Every instance still equals the same batch with
    every instance in the per-substep launches of one chain, bit for bit."""
    ws = _mixed(tmp_path)
    runs = {}
    for name, budget, chains in (("own", 1, 2), ("one", 1 << 30, 2), ("one1", 1 << 30, 1), ("flat", 0, 2)):
        b = BatchedFSI(ws, out_dir=str(tmp_path / name), native=emu, device="cpu", subcycle_group_bytes=budget,
                       penalty_chains=chains)
        b.step()
        b.step()
        runs[name] = b
    assert runs["own"].subcycle_groups() == [[0], [2]]
    assert runs["one"].subcycle_groups() == [[0, 2]]
    assert runs["flat"].subcycle_groups() == []
    assert runs["one"].penalty_sets() == [[3], [1]]  # 16^3 and 12^3
    assert runs["one1"].penalty_sets() == [[1, 3]] and runs["flat"].penalty_sets() == [[1, 3]]
    mb = _native.BATCH_MODE_MOVING_BOUNDARY
    for name in ("own", "one", "one1"):
        plan = runs[name]._plan
        # the grouped batches' per-substep launches hold the penalty instances alone
        assert not [key for _, _, key, _ in plan if key[0] == _native.BATCH_DOMAIN_PLANES and mb in key[1]]
        # side chain 1: the groups' substeps, then the rest of the moving-boundary exchange
        c1 = [i for i, (*_, br) in enumerate(plan) if br == 1]
        groups = [i for i in c1 if plan[i][2][0] == "group"]
        assert {plan[i][0] for i in groups} == {_native.BATCH_OP_LBM_SUBSTEP_MB}
        assert len(groups) == len(runs[name].subcycle_groups()) * runs[name].n_substeps == c1.index(groups[-1]) + 1
        assert all(plan[i][2][0] == "group" or plan[i][2][1] == (mb,) for i in c1)
        assert _native.BATCH_OP_MPM_G2P in {plan[i][0] for i in c1}
        # side chain 2 and the main chain after the fork: penalty sets 1 and 0
        c2 = [i for i, (*_, br) in enumerate(plan) if br == 2]
        fork = [i for i, (_, _, key, br) in enumerate(plan) if br == 0 and key[0] == "set"]
        if name == "one1":
            assert not c2 and not fork
        else:
            assert c2 and {plan[i][2][:2] for i in c2} == {("set", 1)}
            assert fork and {plan[i][2][:2] for i in fork} == {("set", 0)} and fork[0] > c2[-1] > c1[-1]
            assert len(c2) == len(fork)  # the same launches over the other set
        joins = [i for i, (*_, br) in enumerate(plan) if br == _native.BATCH_BRANCH_JOIN]
        assert len(joins) == 1 and joins[0] == len(plan) - 1
        assert plan[joins[0]][0] == _native.BATCH_OP_LBM_MOMENTS  # the penalty instances' last moments
    assert {br for *_, br in runs["flat"]._plan} == {_native.BATCH_BRANCH_MAIN}
    for k in range(4):
        for name in ("own", "one", "one1"):
            a, b = runs["flat"].drivers[k], runs[name].drivers[k]
            for field in ("F", "rho", "v", "hydro_force", "cell_force"):
                assert np.array_equal(getattr(a.lbm, field).to_numpy(), getattr(b.lbm, field).to_numpy()), (k, field)
            for field in ("x", "v", "F", "C"):
                assert np.array_equal(getattr(a.solid, field).to_numpy(), getattr(b.solid, field).to_numpy()), (k, field)
            ca, cb = runs["flat"].coupling_stats(k), runs[name].coupling_stats(k)
            assert bytes(ca) == bytes(cb), (k, name)
    with pytest.raises(ValueError, match="penalty_chains"):
        BatchedFSI(ws, out_dir=str(tmp_path / "bad"), native=emu, device="cpu", penalty_chains=4)


def test_subcycle_groups_first_fit(emu, tmp_path):
    """Groups are packed first fit, largest lattice first, within the budget."""
    ws = _mixed(tmp_path)
    b = BatchedFSI(ws, out_dir=str(tmp_path / "g"), native=emu, device="cpu",
                   subcycle_group_bytes=19 * 4 * 16 ** 3 + 19 * 4 * 12 ** 3 - 1)
    assert b.subcycle_groups() == [[0], [2]]
    b.subcycle_group_bytes = 19 * 4 * (16 ** 3 + 12 ** 3)
    assert b.subcycle_groups() == [[0, 2]]
