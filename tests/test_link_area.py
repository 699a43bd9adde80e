"""The experimental link-area reaction transfer (reaction_transfer_mode =
"link_area_experimental"; the reference's coupling/link_area.py and
link_area_accounting.py) on the serial emulation of the kernels.

What pins the oracle here:
* the per-direction link tables of the last substep add up to the totals
  that the Step-10 artifacts pin (link count) and to the net solid force, and
  the fluid impulse of each direction is minus its solid force;
* with the uniform policy the area scale is 1 (up to the rounding of a ratio
  of two sums of the same forces), and the link-area run then follows the
  engineering run;
* under any policy the scale is the weighted ratio restated from the tables
  (axis weight 1, face-diagonal weight 1/sqrt(2) or sqrt(2)), clipped; a
  zero force gives the lower bound;
* one reaction pass with scale s scatters s times the engineering force.
The product is then held to the oracle: tables, scale and whole runs.
"""
import dataclasses

import numpy as np
import pytest

from mpm_lbm import workloads
from mpm_lbm.sim.coupling.link_area_accounting import LinkAreaMomentumAccounting3D
from mpm_lbm.sim.drivers.fsi_config import FSIDriverConfig
from oracle import oracle as orc
from tests.parity import check_history, make_pair, run_and_compare

POLICIES = ("inverse_length", "uniform", "length")


def _link_area(w, policy="inverse_length", **kw):
    return dataclasses.replace(w, config=FSIDriverConfig(**{
        **w.config.to_dict(), "reaction_transfer_mode": "link_area_experimental", "link_area_policy": policy, **kw}))


def _subcycled(tmp_path, steps=2):
    return workloads.moving_boundary_flap(n=12, steps=steps, substeps=6, mpm_substeps=2, flap=(2, 6, 4),
                                          out_dir=str(tmp_path))


def _one_step(tmp_path, steps=3):
    return workloads.moving_boundary_unsubcycled(n=12, steps=steps, flap=(2, 6, 4), out_dir=str(tmp_path))


def _tables(o):
    Q = orc.Q
    st = o.mbstats
    return (np.asarray(st.link_count_by_dir[:], np.int64), np.asarray(st.solid_force_by_dir[:], np.float64).reshape(Q, 3),
            np.asarray(st.fluid_impulse_by_dir[:], np.float64).reshape(Q, 3))


def _expected_scale(o, policy, lo=0.25, hi=2.0):
    e = orc.E.astype(np.float64)
    length = np.linalg.norm(e, axis=1)
    w = {"uniform": np.where(length > 0, 1.0, 0.0), "inverse_length": np.where(length > 0, 1.0 / np.maximum(length, 1e-30), 0.0),
         "length": length}[policy]
    _, solid, _ = _tables(o)
    num = abs(float((solid[:, 0] * w).sum()))
    den = abs(float(o.mbstats.net_solid_force[0])) + 1.0e-12
    return float(np.clip(num / den, lo, hi))


# -------------------------------------------------------------- oracle pins
@pytest.mark.parametrize("make", [_subcycled, _one_step], ids=["subcycled", "one_step"])
def test_oracle_direction_tables_add_up(emu, tmp_path, make):
    w = _link_area(make(tmp_path))
    d, o = make_pair(w, tmp_path / "run", emu, "cpu")
    for _ in range(w.config.n_lbm_steps):
        o.step_once()
    count, solid, fluid = _tables(o)
    assert count[0] == 0 and count.sum() == o.mbstats.link_count > 0
    np.testing.assert_array_equal(fluid, -solid)
    # the oracle sums its net force in float32: rounding of O(max |table|)
    np.testing.assert_allclose(solid.sum(0), np.asarray(o.mbstats.net_solid_force[:]), rtol=0,
                               atol=2e-6 * np.abs(solid).max())
    # axis directions carry one nonzero force component, diagonals two
    e = orc.E
    for s in range(1, orc.Q):
        assert np.all(solid[s][e[s] == 0] == 0)


@pytest.mark.parametrize("make", [_subcycled, _one_step], ids=["subcycled", "one_step"])
def test_oracle_uniform_policy_is_the_engineering_transfer(emu, tmp_path, make):
    """In FP64 the uniform scale is 1 to rounding (a ratio of two sums of the
    same link forces), so the link-area run follows the engineering run."""
    base = make(tmp_path)
    _, eng = make_pair(base, tmp_path / "eng", emu, "cpu", dtype=np.float64)
    _, la = make_pair(_link_area(base, "uniform"), tmp_path / "la", emu, "cpu", dtype=np.float64)
    for step in range(base.config.n_lbm_steps):
        eng.step_once()
        la.step_once()
        if step > 0:  # the first step from rest has no net force (lower bound)
            assert la.ca.area_scale == pytest.approx(1.0, abs=1e-10)
    for name in ("x", "pv", "F", "hydro"):
        a, b = getattr(la, name), getattr(eng, name)
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-9 * np.abs(b).max(), err_msg=name)


@pytest.mark.parametrize("policy", POLICIES)
def test_oracle_scale_is_the_clipped_weighted_ratio(emu, tmp_path, policy):
    w = _link_area(_subcycled(tmp_path), policy)
    _, o = make_pair(w, tmp_path / "run", emu, "cpu", dtype=np.float64)
    o.step_once()
    o.step_once()
    want = _expected_scale(o, policy)
    assert o.ca.area_scale == pytest.approx(want, rel=1e-12)
    if policy == "inverse_length":
        assert 0.25 < want < 1.0
    if policy == "length":
        assert 1.0 < want < 2.0
    # clip bounds
    w2 = _link_area(_subcycled(tmp_path), policy, link_area_scale_min=0.9, link_area_scale_max=0.95)
    _, o2 = make_pair(w2, tmp_path / "run2", emu, "cpu", dtype=np.float64)
    o2.step_once()
    assert 0.9 <= o2.ca.area_scale <= 0.95


def test_oracle_zero_force_gives_the_lower_bound(emu, tmp_path):
    w = _link_area(_one_step(tmp_path, steps=1))
    _, o = make_pair(w, tmp_path / "run", emu, "cpu")
    o.step_once()  # the first step from rest: no net force on the walls
    assert float(o.mbstats.net_solid_force[0]) == 0.0
    assert o.ca.area_scale == pytest.approx(0.25)


def test_oracle_reaction_scales_linearly(emu, tmp_path):
    """One reaction pass with scale s scatters s times the unscaled force
    (the force cap not reached)."""
    w = _link_area(_subcycled(tmp_path), mb_force_cap_norm=1.0e6)
    _, o = make_pair(w, tmp_path / "run", emu, "cpu", dtype=np.float64)
    o.step_once()
    o.p2g()
    o.ca.area_scale = 1.0
    o.grid_f[:] = 0
    o.reaction(2)
    base = o.grid_f.copy()
    o.ca.area_scale = 0.625
    o.grid_f[:] = 0
    o.reaction(2)
    assert np.abs(base).max() > 0
    np.testing.assert_allclose(o.grid_f, 0.625 * base, rtol=0, atol=1e-12 * np.abs(base).max())


# ------------------------------------------------------- product vs oracle
@pytest.mark.parametrize("policy", POLICIES)
@pytest.mark.parametrize("make", [_subcycled, _one_step], ids=["subcycled", "one_step"])
def test_link_area_runs_match_oracle(emu, tmp_path, make, policy):
    # a small reaction scale, so that many particle forces stay under the cap
    # and show the area scale (capped forces would not)
    w = _link_area(make(tmp_path), policy, mb_reaction_scale=1.0e-8)
    d, o = make_pair(w, tmp_path / "run", emu, "cpu", with_fp64=True)
    history = run_and_compare(d, o, w.config.n_lbm_steps)
    check_history(history, f"link_area/{policy}", 1.0e-5, strict=make is _one_step)
    ds = d.lbm.get_moving_boundary_directional_stats()
    count, _, _ = _tables(o)
    np.testing.assert_array_equal(ds["link_count_by_dir"], count)
    # fp64 sums here, fp32 sums in the FP32 oracle: compare with the FP64 oracle
    _, solid, _ = _tables(o.fp64)
    np.testing.assert_allclose(ds["solid_force_by_dir"], solid, rtol=0, atol=1e-5 * np.abs(solid).max())
    np.testing.assert_array_equal(ds["fluid_impulse_by_dir"], -ds["solid_force_by_dir"])
    np.testing.assert_allclose(ds["correction_abs_max_by_dir"], np.asarray(o.mbstats.corr_abs_max_by_dir[:]),
                               rtol=1e-6, atol=1e-12)
    st = d.link_area_coupler.get_stats()
    assert st["area_scale"] == pytest.approx(o.ca.area_scale, rel=1e-6)
    assert st["area_policy"] == policy
    # the last reaction pass: particle forces carry the scale
    rs = np.asarray(o.reaction_stats, np.float64)
    assert st["active_reaction_particle_count"] == int(rs[0]) > 0
    assert st["max_particle_reaction_norm"] == pytest.approx(rs[1], rel=1e-5)
    np.testing.assert_allclose(st["net_particle_reaction_force"], rs[3:6], rtol=0, atol=1e-5 * np.abs(rs[3:6]).max())


def test_link_accounting_summary(emu, tmp_path):
    """summarize_link_accounting: the reference's keys, and its consistency
    errors (scalar vs directional sums, hydro vs directional force for a
    one-LBM-step schedule) at rounding level."""
    w = _link_area(_one_step(tmp_path, steps=2))
    d, o = make_pair(w, tmp_path / "run", emu, "cpu")
    d.step_once()
    d.step_once()
    s = LinkAreaMomentumAccounting3D.summarize_link_accounting(d.lbm, policy="inverse_length")
    keys = {"policy", "total_link_count", "axis_link_count", "face_diagonal_link_count", "area_proxy_total",
            "scalar_vs_directional_impulse_error_norm", "scalar_vs_directional_solid_error_x",
            "scalar_vs_directional_solid_error_norm", "hydro_vs_directional_solid_error_x",
            "hydro_vs_directional_solid_error_norm", "correction_abs_sum", "correction_abs_max"}
    for base in ("bb_net_fluid_impulse", "bb_net_solid_force", "directional_fluid_impulse",
                 "directional_solid_force", "area_weighted_fluid_impulse", "area_weighted_solid_force",
                 "area_weighted_balance_error", "hydro_force_sum"):
        keys |= {f"{base}_{a}" for a in "xyz"}
    keys |= {f"scalar_vs_directional_impulse_error_{a}" for a in "xyz"}
    assert keys <= set(s)
    assert s["total_link_count"] == s["axis_link_count"] + s["face_diagonal_link_count"] > 0
    scale = max(1.0, abs(s["directional_solid_force_x"]))
    assert s["scalar_vs_directional_solid_error_norm"] < 1e-9 * scale
    assert s["hydro_vs_directional_solid_error_norm"] < 1e-5 * scale
    assert s["area_weighted_balance_error_x"] == 0.0  # fluid impulse = -solid force, link by link
    path = tmp_path / "dir" / "link.npz"
    LinkAreaMomentumAccounting3D.save_directional_npz(d.lbm, str(path))
    with np.load(path, allow_pickle=False) as z:
        assert z["link_count_by_dir"].sum() == s["total_link_count"]


def test_link_area_mode_runs_eagerly_and_batched_form_refuses(emu, tmp_path):
    from mpm_lbm.batch import BatchedFSI
    from mpm_lbm.sim.drivers.fsi_driver import FSIDriver3D

    w = _link_area(_subcycled(tmp_path))
    d = FSIDriver3D(w.config, str(tmp_path / "d"), particle_cloud=w.cloud, native=emu, device="cpu",
                    write_geometry_file=False)
    d.initialize()
    d.step_once()
    assert not d._graph_ready()
    with pytest.raises(NotImplementedError, match="engineering reaction transfer only"):
        BatchedFSI([w], native=emu, device="cpu", out_dir=str(tmp_path / "b"))
