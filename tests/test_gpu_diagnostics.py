"""The driver's diagnostics contract on the GPU: every column of
FSIDriver3D.collect_diagnostics (reduced on the device by libfsi_hip.so)
against the oracle's rows at every step, and against the reference's own
committed Step-10 rows and final particles (tests/golden/step10_mode_matrix).
Column bar as tests/diag.py states; the reference rows within 2e-3 (its
fast-math FP32 with atomics), integer columns exact."""
import numpy as np
import pytest

from mpm_lbm import workloads
from tests import diag, step10

pytestmark = pytest.mark.gpu


def _report(problems):
    for step, p in problems.items():
        print(step, p, flush=True)


@pytest.mark.parametrize("mode", ["none", "penalty", "moving_boundary"])
def test_gpu_step10_rows_match_reference_outputs(hip, tmp_path, mode):
    d, o, problems = diag.run_step10(mode, hip, None, tmp_path / "out", with_fp64=True)
    _report(problems)
    assert not problems, problems
    if mode == "moving_boundary":
        x0 = step10.box_cloud(step10.config(mode))["x"]
        bad, ex, eJ = diag.golden_particle_problems(d, x0)
        print("golden particles: x", ex, "J", eJ, flush=True)
        assert not bad, bad
    if mode != "none":
        st = (d.mb_coupler or d.penalty_coupler).get_stats()
        key = "net_grid_reaction_force" if mode == "moving_boundary" else "net_reaction_grid_force"
        want = np.asarray(o.fp64.reaction_stats[6:9], np.float64)
        np.testing.assert_allclose(st[key], want, rtol=0, atol=1e-4 * np.abs(want).max())


def test_gpu_rows_penalty_32(hip, tmp_path):
    """The 32^3 penalty workload: rows after every one of 10 FSI steps."""
    w = workloads.penalty_block(n=32, steps=10, out_dir=str(tmp_path))
    _, _, problems = diag.run_workload_rows(w, hip, None, tmp_path / "run", 10)
    _report(problems)
    assert not problems, problems


def test_gpu_rows_moving_boundary_64(hip, tmp_path):
    """The 64^3 subcycled moving-boundary workload (120 LBM substeps per FSI
    step): rows after the initial state and one FSI step."""
    w = workloads.moving_boundary_flap(n=64, steps=1, substeps=120, out_dir=str(tmp_path))
    d, _, problems = diag.run_workload_rows(w, hip, None, tmp_path / "run", 1)
    _report(problems)
    assert not problems, problems
    row = d.diagnostics_rows[-1]
    assert row["mb_subcycle_force_sample_count"] == 120
    assert row["mb_subcycle_force_accumulation_mode"] == "mean_force_accumulator"
