"""CPU check of the kernel code: the serial emulation build of the very same
per-cell / per-particle bodies the GPU runs, against the oracle, on small
lattices.  Bar: every field within 1e-5 of the oracle's field magnitude
(max |a - b| / max |b|) after every FSI step; solid flags bit-exact."""
import numpy as np
import pytest

from mpm_lbm import workloads
from tests.parity import check_history, make_pair, run_and_compare

TOL = 1.0e-5


def _check(history):
    check_history(history, "emulation", TOL, strict=True)


def test_moving_boundary_subcycled(emu, tmp_path):
    w = workloads.moving_boundary_flap(n=16, steps=3, substeps=6, flap=(2, 6, 6), mpm_substeps=3,
                                       out_dir=str(tmp_path))
    d, o = make_pair(w, tmp_path / "run", emu, "cpu")
    _check(run_and_compare(d, o, 3))
    acc = d.lbm.get_moving_boundary_accumulation_stats()
    assert acc["mb_subcycle_force_sample_count"] == 6
    assert acc["mb_subcycle_force_mean_norm_max"] == pytest.approx(o.mean_norm_max, rel=1e-5)
    assert acc["mb_subcycle_force_accum_norm_max"] == pytest.approx(o.accum_norm_max, rel=1e-5)
    mbs = d.lbm.get_moving_boundary_stats()
    assert mbs["bb_link_count"] == o.mbstats.link_count
    assert mbs["bb_max_correction"] == pytest.approx(o.mbstats.max_correction, rel=1e-5, abs=1e-12)


def test_moving_boundary_flap_spanning_z(emu, tmp_path):
    """A flap across the whole z extent: particle stencils leave the lattice
    and the MPM grid through both z faces, so the row gathers read shifted
    rows and mask the cells outside."""
    w = workloads.moving_boundary_flap(n=16, steps=2, substeps=4, flap=(2, 6, 16), mpm_substeps=2,
                                       out_dir=str(tmp_path))
    x = np.asarray(w.cloud["x"]) * 16
    base = np.floor(x[:, 2] - 0.5).astype(int)
    assert base.min() == -1 and base.max() == 15  # shifts of +1, -1 and -2 against the 16-cell lattice
    d, o = make_pair(w, tmp_path / "run", emu, "cpu")
    _check(run_and_compare(d, o, 2))


def test_moving_boundary_one_step_per_fsi_step(emu, tmp_path):
    w = workloads.moving_boundary_unsubcycled(n=16, steps=3, flap=(2, 6, 6), out_dir=str(tmp_path))
    d, o = make_pair(w, tmp_path / "run", emu, "cpu")
    _check(run_and_compare(d, o, 4))
    assert d.lbm.get_moving_boundary_accumulation_stats()["mb_subcycle_force_sample_count"] == 0


def test_penalty(emu, tmp_path):
    w = workloads.penalty_block(n=16, steps=4, block=(4, 4, 2), out_dir=str(tmp_path))
    d, o = make_pair(w, tmp_path / "run", emu, "cpu")
    _check(run_and_compare(d, o, 4))


def test_none_mode_periodic(emu, tmp_path):
    from mpm_lbm.sim.drivers.fsi_config import FSIDriverConfig

    w = workloads.penalty_block(n=12, steps=2, block=(4, 4, 2), out_dir=str(tmp_path))
    cfg = FSIDriverConfig(coupling_mode="none", n_grid=12, n_particles=w.config.n_particles, n_lbm_steps=2,
                          mpm_substeps_per_lbm_step=2, mpm_dt=1e-4, gravity=(0.0, -9.8, 0.0))
    w2 = workloads.Workload("none", cfg, w.cloud, w.geometry_json, (400.0, 0.2, 1.0))
    d, o = make_pair(w2, tmp_path / "run2", emu, "cpu")
    _check(run_and_compare(d, o, 3))
