"""The driver's diagnostics contract on the kernels' own code (the serial
emulation build of the HIP bodies): every row of the reference's Step-10 mode
matrix, reduced on the device path, against the oracle's rows at every step
(1e-5 of each column's scale) and against the reference's committed rows and
final particles (tests/golden/step10_mode_matrix, 2e-3).  The GPU library
runs the same checks in tests/test_gpu_diagnostics.py."""
import numpy as np
import pytest

from tests import diag, step10


@pytest.mark.parametrize("mode", ["none", "penalty", "moving_boundary"])
def test_emulated_step10_rows(emu, tmp_path, mode):
    d, o, problems = diag.run_step10(mode, emu, "cpu", tmp_path / "out", with_fp64=True)
    assert not problems, problems
    if mode == "moving_boundary":
        x0 = step10.box_cloud(step10.config(mode))["x"]
        bad, _, _ = diag.golden_particle_problems(d, x0)
        assert not bad, bad
    if mode != "none":
        # the in-grid reaction sum (moving_boundary.py:101-111, penalty.py:129),
        # against the FP64 oracle: the FP32 oracle sums its ~10^5 node
        # contributions sequentially in FP32, ~2e-5 off by itself
        st = (d.mb_coupler or d.penalty_coupler).get_stats()
        key = "net_grid_reaction_force" if mode == "moving_boundary" else "net_reaction_grid_force"
        want = np.asarray(o.fp64.reaction_stats[6:9], np.float64)
        np.testing.assert_allclose(st[key], want, rtol=0, atol=1e-5 * np.abs(want).max())


def test_emulated_net_grid_reaction_leaves_out_nodes_outside(emu, tmp_path):
    """A flap across the whole z extent: stencils leave the MPM grid, so the
    in-grid reaction sum is smaller than the particle sum, as the reference's
    per-node atomics give (moving_boundary.py:101-111)."""
    from mpm_lbm import workloads
    from tests.parity import make_pair

    w = workloads.moving_boundary_flap(n=16, steps=2, substeps=4, flap=(2, 6, 16), mpm_substeps=2,
                                       out_dir=str(tmp_path))
    d, o = make_pair(w, tmp_path / "run", emu, "cpu", with_fp64=True)
    for _ in range(2):
        d.step_once()
        o.step_once()
        o.fp64.step_once()
    st = d.mb_coupler.get_stats()
    grid, part = np.asarray(st["net_grid_reaction_force"]), np.asarray(st["net_particle_reaction_force"])
    want = np.asarray(o.fp64.reaction_stats[6:9], np.float64)
    np.testing.assert_allclose(grid, want, rtol=0, atol=1e-5 * np.abs(want).max())
    np.testing.assert_allclose(part, o.fp64.reaction_stats[3:6], rtol=0, atol=1e-5 * np.abs(want).max())
    assert np.abs(part - grid).max() > 1e-3 * np.abs(part).max()


@pytest.mark.parametrize("kind", ["penalty", "moving_boundary_subcycled"])
def test_emulated_rows_of_workloads(emu, tmp_path, kind):
    """Rows of the penalty block and the subcycled flap workloads (small
    lattices) against the oracle rows after every FSI step."""
    from mpm_lbm import workloads

    if kind == "penalty":
        w = workloads.penalty_block(n=16, steps=4, block=(4, 4, 2), out_dir=str(tmp_path))
    else:
        w = workloads.moving_boundary_flap(n=16, steps=3, substeps=6, flap=(2, 6, 6), mpm_substeps=3,
                                           out_dir=str(tmp_path))
    d, _, problems = diag.run_workload_rows(w, emu, "cpu", tmp_path / "run", 3)
    assert not problems, problems
    if kind != "penalty":
        assert d.diagnostics_rows[-1]["mb_subcycle_force_sample_count"] == 6
