"""LBM restart files in the reference's format and whole-run FSI checkpoints,
on the serial emulation of the kernels.

The restart format follows the reference's lbm/restart.py (keys, shapes, the
JSON metadata record, the n_grid / boundary-condition / target-velocity guard,
required vs optional files; its tests/test_lbm_restart_*.py hold no vectors,
only this contract).  A checkpointed run resumed in a fresh driver must end
bit-identical to the same run taken without interruption, in both coupling
modes; writes are atomic and only the newest `keep_last` files stay.
"""
import json

import numpy as np
import pytest

from mpm_lbm import workloads
from mpm_lbm.checkpoint import (CHECKPOINT_METADATA_KEY, latest_fsi_checkpoint, list_fsi_checkpoints,
                                load_fsi_checkpoint, save_fsi_checkpoint)
from mpm_lbm.sim.drivers.fsi_config import FSIDriverConfig
from mpm_lbm.sim.drivers.fsi_driver import FSIDriver3D
from mpm_lbm.sim.lbm.restart import (RESTART_ARRAY_KEYS, expected_restart_metadata_from_config,
                                     read_lbm_restart_npz, save_lbm_restart_from_lbm, validate_lbm_restart_metadata,
                                     write_lbm_restart_npz)


def _driver(w, out, emu, **cfg):
    config = FSIDriverConfig(**{**w.config.to_dict(), **cfg}) if cfg else w.config
    return FSIDriver3D(config, str(out), particle_cloud=w.cloud, native=emu, device="cpu", write_geometry_file=False,
                       use_graph=False)


def _state(d):
    lbm, s = d.lbm, d.solid
    out = {f"lbm_{k}": getattr(lbm, k).to_numpy() for k in ("F", "rho", "v", "solid", "solid_phi", "solid_vel",
                                                           "cell_force", "hydro_force")}
    out.update({f"mpm_{k}": getattr(s, k).to_numpy() for k in ("x", "v", "C", "F", "Jp")})
    return out


def _same(a, b):
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def _rows(rows):
    """Diagnostics rows without the wall-clock column."""
    return json.dumps([{k: v for k, v in r.items() if k != "elapsed_seconds"} for r in rows])


def _penalty(tmp_path, steps=4):
    return workloads.penalty_block(n=12, steps=steps, block=(4, 4, 2), seed=3, out_dir=str(tmp_path))


def _moving_boundary(tmp_path, steps=3):
    return workloads.moving_boundary_flap(n=12, steps=steps, substeps=4, mpm_substeps=2, flap=(2, 6, 4), seed=5,
                                          out_dir=str(tmp_path))


# ---------------------------------------------------------------- LBM restart
def _restart_arrays(n, rng):
    F = (rng.random((n, n, n, 19)) * 0.05 + 0.01).astype(np.float32)
    solid = np.zeros((n, n, n), dtype=np.int8)
    solid[:, 0, :] = 1
    return {"rho": F.sum(axis=3), "v": (rng.standard_normal((n, n, n, 3)) * 1e-3).astype(np.float32), "f": F,
            "F": F, "solid": solid, "static_solid": solid.copy()}


def test_restart_file_layout_and_stats(tmp_path):
    rng = np.random.default_rng(0)
    a = _restart_arrays(6, rng)
    meta = {"n_grid": 6, "lbm_boundary_condition_mode": "duct_velocity_inlet_pressure_outlet",
            "target_u_lbm": [0.04, 0.0, 0.0]}
    stats = write_lbm_restart_npz(tmp_path / "r.npz", a, meta)
    with np.load(tmp_path / "r.npz", allow_pickle=False) as z:
        assert sorted(z.files) == sorted(RESTART_ARRAY_KEYS + ("metadata_json",))
        assert json.loads(str(z["metadata_json"])) == meta
    fluid = a["solid"] == 0
    assert stats["rho_min"] == pytest.approx(float(a["rho"][fluid].min()))
    assert stats["rho_max"] == pytest.approx(float(a["rho"][fluid].max()))
    assert stats["velocity_max_norm"] == pytest.approx(float(np.linalg.norm(a["v"][fluid], axis=1).max()), rel=1e-6)
    assert not stats["has_nan"] and not stats["has_inf"]
    back, meta2 = read_lbm_restart_npz(tmp_path / "r.npz")
    assert meta2 == meta
    for k in RESTART_ARRAY_KEYS:
        np.testing.assert_array_equal(back[k], a[k])
    with pytest.raises(ValueError, match="missing LBM restart arrays: static_solid"):
        write_lbm_restart_npz(tmp_path / "bad.npz", {k: v for k, v in a.items() if k != "static_solid"}, meta)
    a["v"][1, 1, 1, 0] = np.nan
    assert write_lbm_restart_npz(tmp_path / "nan.npz", a, meta)["has_nan"]


def test_restart_metadata_guard():
    expected = {"n_grid": 12, "lbm_boundary_condition_mode": "duct_velocity_inlet_pressure_outlet",
                "target_u_lbm": [0.04, 0.0, 0.0]}
    assert all(validate_lbm_restart_metadata(dict(expected), expected).values())
    g = validate_lbm_restart_metadata({**expected, "n_grid": 16}, expected)
    assert not g["restart_n_grid_matches"] and g["restart_target_u_lbm_matches"]
    g = validate_lbm_restart_metadata({**expected, "target_u_lbm": [0.05, 0.0, 0.0]}, expected)
    assert not g["restart_target_u_lbm_matches"]
    g = validate_lbm_restart_metadata({**expected, "lbm_boundary_condition_mode": "periodic"}, expected)
    assert not g["restart_lbm_boundary_condition_mode_matches"]


def test_driver_loads_restart_state(emu, tmp_path):
    """A restart written by one driver and loaded by a fresh one through
    lbm_restart_path reproduces the writer's LBM state, and the next step is
    bit-identical to the writer's next step."""
    w = _penalty(tmp_path, steps=3)
    a = _driver(w, tmp_path / "a", emu)
    a.initialize()
    a.step_once()
    a.step_once()
    path = tmp_path / "restart" / "lbm_restart.npz"
    stats = save_lbm_restart_from_lbm(a.lbm, path, expected_restart_metadata_from_config(a.config))
    assert stats["array_keys"] == list(RESTART_ARRAY_KEYS)
    # LBM only: reset the particles of the writer to the initial cloud so both
    # drivers step the same solid from here
    b = _driver(w, tmp_path / "b", emu, lbm_restart_path=str(path), lbm_restart_required=True)
    b.initialize()
    assert b.lbm_restart_load_report["restart_loaded"] is True
    assert json.loads((tmp_path / "b" / "lbm_restart_load_report.json").read_text())["restart_loaded"] is True
    for k in ("F", "rho", "v", "solid", "static_solid"):
        np.testing.assert_array_equal(getattr(b.lbm, k).to_numpy(), getattr(a.lbm, k).to_numpy(), err_msg=k)
    fresh = _driver(w, tmp_path / "c", emu)
    fresh.initialize()
    for k in ("x", "v", "C", "F", "Jp"):
        getattr(a.solid, k).from_numpy(getattr(fresh.solid, k).to_numpy())
    a.lbm.cell_force.from_numpy(np.zeros_like(a.lbm.cell_force.to_numpy()))
    a.step_once()
    b.step_once()
    for k in ("F", "rho", "v", "cell_force"):
        np.testing.assert_array_equal(getattr(b.lbm, k).to_numpy(), getattr(a.lbm, k).to_numpy(), err_msg=k)


def test_driver_restart_guard_and_missing_file(emu, tmp_path):
    w = _penalty(tmp_path)
    a = _driver(w, tmp_path / "a", emu)
    a.initialize()
    path = tmp_path / "r.npz"
    meta = expected_restart_metadata_from_config(a.config)
    save_lbm_restart_from_lbm(a.lbm, path, {**meta, "target_u_lbm": [0.09, 0.0, 0.0]})
    with pytest.raises(ValueError, match="metadata mismatch"):
        _driver(w, tmp_path / "b", emu, lbm_restart_path=str(path)).initialize()
    # a wrongly shaped payload under matching metadata
    arrays = _restart_arrays(8, np.random.default_rng(1))
    write_lbm_restart_npz(tmp_path / "shape.npz", arrays, meta)
    with pytest.raises(ValueError, match="shape mismatch"):
        _driver(w, tmp_path / "c", emu, lbm_restart_path=str(tmp_path / "shape.npz")).initialize()
    missing = str(tmp_path / "none.npz")
    d = _driver(w, tmp_path / "d", emu, lbm_restart_path=missing, lbm_restart_required=False)
    d.initialize()
    assert d.lbm_restart_load_report["restart_loaded"] is False
    assert d.lbm_restart_load_report["restart_missing"] is True
    with pytest.raises(FileNotFoundError):
        _driver(w, tmp_path / "e", emu, lbm_restart_path=missing, lbm_restart_required=True).initialize()


# ------------------------------------------------------------ FSI checkpoints
@pytest.mark.parametrize("mode", ["penalty", "moving_boundary", "moving_boundary_unsubcycled"])
def test_resumed_run_is_bit_identical(emu, tmp_path, mode):
    """Run interrupted after k steps and resumed in a new driver == the same
    run without interruption (every field, diagnostics rows, outputs)."""
    if mode == "penalty":
        w = _penalty(tmp_path)
    elif mode == "moving_boundary":
        w = _moving_boundary(tmp_path)
    else:
        w = workloads.moving_boundary_unsubcycled(n=12, steps=4, flap=(2, 6, 4), seed=6, out_dir=str(tmp_path))
    n = w.config.n_lbm_steps
    ref = _driver(w, tmp_path / "ref", emu)
    ref.run()
    ck = tmp_path / "ck"
    first = _driver(w, tmp_path / "run", emu)
    first.run(checkpoint_dir=ck, checkpoint_every=1, stop_after_steps=2)
    assert first.current_lbm_step == 2
    assert not (tmp_path / "run" / "diagnostics_timeseries.csv").exists()
    assert latest_fsi_checkpoint(ck).name == "ckpt_00000002.npz"
    second = _driver(w, tmp_path / "run", emu)
    second.run(checkpoint_dir=ck, checkpoint_every=1, resume=True)
    assert second.current_lbm_step == n
    assert second.total_lbm_substeps == ref.total_lbm_substeps
    assert second.total_mpm_substeps == ref.total_mpm_substeps
    _same(_state(second), _state(ref))
    assert _rows(second.diagnostics_rows) == _rows(ref.diagnostics_rows)
    assert second.flap_tip_monitor_rows == ref.flap_tip_monitor_rows
    assert second.fluent_like_monitor_rows == ref.fluent_like_monitor_rows
    with np.load(tmp_path / "run" / "diagnostics_timeseries.npz") as a, \
            np.load(tmp_path / "ref" / "diagnostics_timeseries.npz") as b:
        assert a.files == b.files
        for k in a.files:
            if k != "elapsed_seconds":
                np.testing.assert_array_equal(a[k], b[k], err_msg=k)
    # writing checkpoints does not change the run that writes them
    full = _driver(w, tmp_path / "full", emu)
    full.run(checkpoint_dir=tmp_path / "ck_full", checkpoint_every=1)
    _same(_state(full), _state(ref))
    # a resume of a finished run steps nothing
    third = _driver(w, tmp_path / "run", emu)
    third.run(checkpoint_dir=ck, resume=True)
    assert third.current_lbm_step == n
    _same(_state(third), _state(ref))


def test_checkpoint_keep_last_and_atomic_write(emu, tmp_path):
    w = _penalty(tmp_path, steps=5)
    d = _driver(w, tmp_path / "run", emu)
    ck = tmp_path / "ck"
    d.run(checkpoint_dir=ck, checkpoint_every=1, keep_last=2)
    assert [p.name for p in list_fsi_checkpoints(ck)] == ["ckpt_00000004.npz", "ckpt_00000005.npz"]
    assert not list(ck.glob(".tmp_ckpt_*"))
    with np.load(latest_fsi_checkpoint(ck), allow_pickle=False) as z:
        meta = json.loads(str(z[CHECKPOINT_METADATA_KEY]))
        assert meta["current_lbm_step"] == 5
        # the reference's restart arrays are part of every checkpoint
        for k in RESTART_ARRAY_KEYS:
            assert f"lbm_{k}" in z.files
    # a failed write leaves the directory as it was
    class Boom(Exception):
        pass

    import mpm_lbm.checkpoint as ckmod

    real = ckmod.np.savez

    def broken(f, **arrays):
        f.write(b"partial")
        raise Boom()

    ckmod.np.savez = broken
    try:
        with pytest.raises(Boom):
            save_fsi_checkpoint(d, ck, keep_last=2)
    finally:
        ckmod.np.savez = real
    assert [p.name for p in ck.iterdir()] and not list(ck.glob(".tmp_ckpt_*"))
    assert [p.name for p in list_fsi_checkpoints(ck)] == ["ckpt_00000004.npz", "ckpt_00000005.npz"]


def test_checkpoint_guard(emu, tmp_path):
    w = _penalty(tmp_path, steps=2)
    d = _driver(w, tmp_path / "a", emu)
    d.initialize()
    d.step_once()
    path = save_fsi_checkpoint(d, tmp_path / "ck")
    other = _driver(w, tmp_path / "b", emu, target_u_lbm=(0.05, 0.0, 0.0))
    with pytest.raises(ValueError, match="config mismatch"):
        load_fsi_checkpoint(other, path)
    with pytest.raises(FileNotFoundError):
        load_fsi_checkpoint(other, tmp_path / "ck" / "ckpt_00000099.npz")
    with pytest.raises(ValueError, match="resume=True needs checkpoint_dir"):
        _driver(w, tmp_path / "c", emu).run(resume=True)


def test_checkpoint_guards_material_and_time_step(emu, tmp_path):
    """The material comes from the geometry file, not a config entry: a run
    whose resolved material differs refuses the checkpoint, and so does one
    with another MPM time step or reaction scale."""
    w = _penalty(tmp_path / "w1", steps=2)
    d = _driver(w, tmp_path / "a", emu)
    d.initialize()
    d.step_once()
    path = save_fsi_checkpoint(d, tmp_path / "ck")
    stiffer = workloads.penalty_block(n=12, steps=2, block=(4, 4, 2), seed=3, E_mod=9.0e4, out_dir=str(tmp_path / "w2"))
    with pytest.raises(ValueError, match="material/time-step mismatch"):
        load_fsi_checkpoint(_driver(stiffer, tmp_path / "b", emu), path)
    with pytest.raises(ValueError, match="config mismatch"):
        load_fsi_checkpoint(_driver(w, tmp_path / "c", emu, mpm_dt=2.0e-4), path)
    with pytest.raises(ValueError, match="config mismatch"):
        load_fsi_checkpoint(_driver(w, tmp_path / "e", emu, beta_lbm=2.0e-3), path)


def test_checkpoint_restores_an_inactive_cell_force(emu, tmp_path):
    """A checkpoint written before any penalty force existed, loaded into a
    driver that has run penalty steps, leaves no stale force behind."""
    w = _penalty(tmp_path, steps=3)
    fresh = _driver(w, tmp_path / "a", emu)
    fresh.initialize()
    path = save_fsi_checkpoint(fresh, tmp_path / "ck")
    busy = _driver(w, tmp_path / "b", emu)
    busy.initialize()
    busy.solid.set_uniform_velocity(0.05, 0.0, 0.0)  # a moving block: a nonzero penalty force
    for _ in range(2):
        busy.step_once()
    assert np.abs(busy.lbm.cell_force.to_numpy()).max() > 0
    load_fsi_checkpoint(busy, path)
    assert np.abs(busy.lbm.cell_force.to_numpy()).max() == 0
    assert not busy.lbm._cell_force_active
    busy.step_once()
    fresh.step_once()
    _same(_state(busy), _state(fresh))
