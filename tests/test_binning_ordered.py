"""Particle binning and the particle-ordered (deterministic) transfers on the
kernels' own code (the serial emulation build):

  * the bin keys, offsets and lists of fsi_bin_particles equal, bit for bit,
    those built from the oracle's stencil bases (tests/binning.py);
  * FSIDriver3D with mpm_transfer_order="particle" (binned projection and
    P2G, every node sum in particle order) meets the strict parity bar
    against the oracle in every coupling schedule.
The GPU runs the same checks in tests/test_gpu_ordered.py."""
import dataclasses

import numpy as np
import pytest

from mpm_lbm import workloads
from mpm_lbm.sim.drivers.fsi_config import FSIDriverConfig
from tests import binning
from tests.parity import check_history, make_pair, run_and_compare


def ordered(w):
    return dataclasses.replace(w, config=FSIDriverConfig(**{**w.config.to_dict(), "mpm_transfer_order": "particle"}))


@pytest.mark.parametrize("shape,seed", [((16, 16, 16), 0), ((20, 12, 9), 1)])
def test_binning_matches_oracle_bases(emu, shape, seed):
    inv_dx = float(max(shape))
    x = binning.cloud_for_binning(seed, shape, 5000)
    got = binning.native_binning(emu, x, inv_dx, shape, "cpu")
    want = binning.oracle_binning(x, inv_dx, shape)
    for name, a, b in zip(("key", "start", "order"), got, want):
        assert np.array_equal(a, b), name
    key = want[0]
    nb = (shape[0] + 3) * (shape[1] + 3) * (shape[2] + 3)
    assert (key == nb).any() and np.bincount(key[key < nb]).max() >= 8  # overflow and crowded bins exercised


def test_binning_of_no_particles(emu):
    key, start, order = binning.native_binning(emu, np.zeros((0, 3), np.float32), 8.0, (8, 8, 8), "cpu")
    assert not start.any()


@pytest.mark.parametrize("kind", ["moving_boundary_subcycled", "penalty", "moving_boundary_one_step", "flap_across_z"])
def test_particle_order_mode_strict_parity(emu, tmp_path, kind):
    if kind == "moving_boundary_subcycled":
        w = workloads.moving_boundary_flap(n=16, steps=3, substeps=6, flap=(2, 6, 6), mpm_substeps=3,
                                           out_dir=str(tmp_path))
    elif kind == "penalty":
        w = workloads.penalty_block(n=16, steps=4, block=(4, 4, 2), out_dir=str(tmp_path))
    elif kind == "moving_boundary_one_step":
        w = workloads.moving_boundary_unsubcycled(n=16, steps=3, flap=(2, 6, 6), out_dir=str(tmp_path))
    else:
        w = workloads.moving_boundary_flap(n=16, steps=2, substeps=4, flap=(2, 6, 16), mpm_substeps=2,
                                           out_dir=str(tmp_path))
    d, o = make_pair(ordered(w), tmp_path / "run", emu, "cpu")
    assert d.solid.transfer_order == "particle" and d.projector.transfer_order == "particle"
    check_history(run_and_compare(d, o, 3), f"particle order {kind}", 1.0e-5, strict=True)


def test_particle_order_mode_step10_rows(emu, tmp_path):
    """The Step-10 moving-boundary run (a free box crossing cells every step)
    in particle order: rows against the oracle and the reference outputs."""
    from tests import diag, step10

    cfg = step10.config("moving_boundary")
    cfg = FSIDriverConfig(**{**cfg.to_dict(), "mpm_transfer_order": "particle"})
    orig = step10.config
    try:
        step10.config = lambda mode: cfg
        d, _, problems = diag.run_step10("moving_boundary", emu, "cpu", tmp_path / "out", with_fp64=True)
    finally:
        step10.config = orig
    assert not problems, problems
