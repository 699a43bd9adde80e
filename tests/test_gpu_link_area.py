"""GPU parity of the experimental link-area reaction transfer
(link_area.py:84-194) against the oracle: the per-direction link tables of
the last substep, the area scale and whole runs (1e-5 bar as in
tests/test_gpu_parity.py)."""
import numpy as np
import pytest

from mpm_lbm import workloads
from tests.parity import check_history, make_pair, run_and_compare
from tests.test_link_area import _link_area, _tables

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("policy", ["inverse_length", "length"])
@pytest.mark.parametrize("schedule", ["subcycled", "one_step"])
def test_gpu_link_area_matches_oracle(hip, tmp_path, schedule, policy):
    if schedule == "subcycled":
        w = workloads.moving_boundary_flap(n=32, steps=2, substeps=20, flap=(2, 16, 12), out_dir=str(tmp_path))
    else:
        w = workloads.moving_boundary_unsubcycled(n=32, steps=3, flap=(2, 16, 12), out_dir=str(tmp_path))
    w = _link_area(w, policy, mb_reaction_scale=1.0e-8)
    d, o = make_pair(w, tmp_path / "run", hip, None, with_fp64=True)
    history = run_and_compare(d, o, w.config.n_lbm_steps)
    for step, errs in enumerate(history):
        print(schedule, policy, step, {k: v for k, v in errs.items() if k not in ("vs_fp64", "method")}, flush=True)
    check_history(history, f"link_area/{schedule}/{policy}", 1.0e-5, strict=schedule == "one_step")
    ds = d.lbm.get_moving_boundary_directional_stats()
    count, _, _ = _tables(o)
    np.testing.assert_array_equal(ds["link_count_by_dir"], count)
    # the tables are fp64 sums here and fp32 sums in the FP32 oracle, whose
    # own rounding over ~10^3 links reaches 2e-5: compare with the FP64 oracle
    _, solid, _ = _tables(o.fp64)
    np.testing.assert_allclose(ds["solid_force_by_dir"], solid, rtol=0, atol=1e-5 * np.abs(solid).max())
    st = d.link_area_coupler.get_stats()
    print("scale", st["area_scale"], o.ca.area_scale, flush=True)
    assert st["area_scale"] == pytest.approx(o.ca.area_scale, rel=1e-5)
    rs = np.asarray(o.reaction_stats, np.float64)
    assert st["active_reaction_particle_count"] == int(rs[0]) > 0
    np.testing.assert_allclose(st["net_particle_reaction_force"], rs[3:6], rtol=0, atol=1e-4 * np.abs(rs[3:6]).max())
