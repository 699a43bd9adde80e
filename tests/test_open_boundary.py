"""Open-boundary treatments of the x faces beyond the legacy reset:
regularized_velocity_pressure, its limited form and the convective pressure
outlet (the reference's fluid.py:1027-1111, :1349-1493).

The oracle's bc_open is pinned by
* the reference's own counter contract for the limiter
  (tests/test_step120_actual_limiter_counter_contract.py:12-61: a 5x4x4
  lattice, every reconstructed population clipped in rho, |u| and the
  non-equilibrium part),
* closed forms: a neighbour at equilibrium leaves exactly feq(target); a
  population field linear in x is continued exactly by the convective outlet;
  the limited form without its limiter is the plain regularized form; the
  limiter's clips land on their bounds,
and the product (the serial emulation of the kernels here, libfsi_hip.so in
tests/test_gpu_open_boundary.py) is held to the oracle: the standalone
Boundary_condition and whole FSI runs in every coupling schedule.
"""
import ctypes
import dataclasses
from types import SimpleNamespace

import numpy as np
import pytest

from mpm_lbm import workloads
from mpm_lbm.sim.drivers.fsi_config import FSIDriverConfig
from mpm_lbm.sim.lbm.config import LBMConfig
from mpm_lbm.sim.lbm.fluid import LBMFluid3D
from oracle import oracle as orc
from tests.parity import check_history, make_pair, run_and_compare

Q = 19
UNKNOWN_MIN = (1, 7, 9, 11, 13)
UNKNOWN_MAX = (2, 8, 10, 12, 14)
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)

LIMITED = dict(open_boundary_limiter_enabled=True, open_boundary_u_max=0.045, open_boundary_noneq_cap=0.002,
               open_boundary_population_floor=0.0)
MODES = [("regularized_velocity_pressure", {}), ("regularized_velocity_pressure_limited", LIMITED),
         ("convective_pressure_outlet_experimental", {})]


def feq64(s, rho, u):
    """D3Q19 second-order equilibrium (textbook form, Qian et al. 1992)."""
    e = orc.E[s].astype(np.float64)
    eu = float(e @ u)
    return W[s] * rho * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * float(u @ u))


def _ob_config(sem, **kw):
    base = dict(open_boundary_limiter_enabled=False, open_boundary_rho_min=0.8, open_boundary_rho_max=1.2,
                open_boundary_u_max=0.1, open_boundary_noneq_cap=0.05, open_boundary_population_floor=None)
    base.update(kw)
    return SimpleNamespace(lbm_open_boundary_semantics=sem, **base)


def _bc_open(shape, sem, F, rho, v, solid=None, bc=(1, 0), rho_bc=(1.0, 1.0), vel_bc=((0, 0, 0), (0, 0, 0)),
             dtype=np.float64, **kw):
    """The oracle's Boundary_condition on (nx,ny,nz,19) populations; returns (F, counts)."""
    LbmArgs = orc._STRUCTS[dtype][0]
    la = LbmArgs()
    la.nx, la.ny, la.nz = shape
    la.rho0 = 1.0
    la.bc_x_left, la.bc_x_right = bc
    la.rho_bc_x_left, la.rho_bc_x_right = rho_bc
    la.vel_bc_x_left[:] = list(vel_bc[0])
    la.vel_bc_x_right[:] = list(vel_bc[1])
    ob = orc.ob_args(_ob_config(sem, **kw), dtype)
    Fa = np.array(F, dtype).reshape(-1)  # copies: the oracle rewrites F in place
    ra = np.array(rho, dtype).reshape(-1)
    va = np.array(v, dtype).reshape(-1)
    sa = np.zeros(int(np.prod(shape)), np.int8) if solid is None else np.ascontiguousarray(solid.reshape(-1), np.int8)
    k = orc.ObCounts()
    sfx = "f32" if dtype == np.float32 else "f64"
    getattr(orc.lib(), f"orc_bc_open_{sfx}")(ctypes.byref(la), ctypes.byref(ob), orc._p(Fa), orc._p(ra), orc._p(va),
                                             orc._p(sa), ctypes.byref(k))
    return Fa.reshape(tuple(shape) + (Q,)), np.asarray(k.c[:], np.int64)


def _reference_counter_case():
    """tests/test_step120_actual_limiter_counter_contract.py:14-42: 5x4x4,
    limited, pressure inlet at rho 2 (clipped to 1.1), cell 1 at u = (1,0,0)
    (clipped to 0.01) and F = 2 (non-equilibrium clipped to 0.001)."""
    shape = (5, 4, 4)
    F = np.zeros(shape + (Q,))
    for s in range(Q):
        F[..., s] = W[s]
    F[1] = 2.0
    rho = np.ones(shape)
    v = np.zeros(shape + (3,))
    v[1, :, :, 0] = 1.0
    kw = dict(open_boundary_limiter_enabled=True, open_boundary_rho_min=0.9, open_boundary_rho_max=1.1,
              open_boundary_u_max=0.01, open_boundary_noneq_cap=0.001, open_boundary_population_floor=0.09)
    return shape, F, rho, v, kw


def test_oracle_reference_limiter_counter_contract():
    shape, F, rho, v, kw = _reference_counter_case()
    out, k = _bc_open(shape, "regularized_velocity_pressure_limited", F, rho, v, bc=(1, 0), rho_bc=(2.0, 1.0),
                      dtype=np.float32, **kw)
    expected = 5 * 4 * 4
    rho_clip, vel_clip, noneq_clip, floor, recon = k
    assert recon == expected and rho_clip == expected and vel_clip == expected and noneq_clip == expected
    assert floor > 0
    # every reconstructed value is feq(1.1, (0.01,0,0)) + 0.001, floored at 0.09
    u = np.array([0.01, 0.0, 0.0])
    for s in UNKNOWN_MIN:
        want = max(feq64(s, 1.1, u) + 0.001, 0.09)
        np.testing.assert_allclose(out[0, :, :, s], want, rtol=2e-6)
    # known populations and the other planes are untouched
    known = [s for s in range(Q) if s not in UNKNOWN_MIN]
    np.testing.assert_array_equal(out[0][..., known], F[0][..., known].astype(np.float32))
    np.testing.assert_array_equal(out[1:], F[1:].astype(np.float32))


def test_oracle_regularized_of_equilibrium_neighbour_is_target_equilibrium():
    """Neighbour populations at feq(rho_n, u_n): the non-equilibrium part is 0
    and the unknowns become feq(target) exactly, on both faces and bc kinds."""
    shape = (6, 3, 4)
    rng = np.random.default_rng(3)
    rho = 1.0 + 0.02 * rng.standard_normal(shape)
    v = 0.03 * rng.standard_normal(shape + (3,))
    F = np.empty(shape + (Q,))
    for idx in np.ndindex(shape):
        for s in range(Q):
            F[idx + (s,)] = feq64(s, rho[idx], v[idx])
    for bc, rho_bc, vel_bc in (((2, 1), (1.0, 0.97), ((0.04, 0.01, 0.0), (0, 0, 0))),
                               ((1, 2), (1.03, 1.0), ((0, 0, 0), (-0.02, 0.0, 0.01)))):
        out, _ = _bc_open(shape, "regularized_velocity_pressure", F, rho, v, bc=bc, rho_bc=rho_bc, vel_bc=vel_bc)
        for face, inner, unknown, kind, rb, vb in ((0, 1, UNKNOWN_MIN, bc[0], rho_bc[0], vel_bc[0]),
                                                   (5, 4, UNKNOWN_MAX, bc[1], rho_bc[1], vel_bc[1])):
            for j in range(3):
                for kk in range(4):
                    n = (inner, j, kk)
                    tr, tu = (rb, v[n]) if kind == 1 else (rho[n], np.asarray(vb, float))
                    for s in unknown:
                        assert out[face, j, kk, s] == pytest.approx(feq64(s, tr, tu), abs=1e-15)


def test_oracle_convective_outlet_continues_linear_fields():
    shape = (7, 3, 3)
    rng = np.random.default_rng(5)
    a, b = rng.random((3, 3, Q)), 0.01 * rng.random((3, 3, Q))
    F = np.stack([a + i * b for i in range(7)])
    out, _ = _bc_open(shape, "convective_pressure_outlet_experimental", F, np.ones(shape), np.zeros(shape + (3,)),
                      bc=(0, 1))
    for s in range(Q):
        want = a[..., s] + 6 * b[..., s] if s in UNKNOWN_MAX else F[6, ..., s]
        np.testing.assert_allclose(out[6, ..., s], want, rtol=1e-13, atol=1e-15)
    # solid cell at nx-2: the face copies itself (2F - F); solid at nx-3: F[nx-2]
    solid = np.zeros(shape, np.int8)
    solid[5, 0, 0] = 1
    solid[4, 1, 1] = 1
    out, _ = _bc_open(shape, "convective_pressure_outlet_experimental", F, np.ones(shape), np.zeros(shape + (3,)),
                      solid=solid, bc=(0, 1))
    for s in UNKNOWN_MAX:
        assert out[6, 0, 0, s] == F[6, 0, 0, s]
        assert out[6, 1, 1, s] == F[5, 1, 1, s]


def test_oracle_limiter_bounds_and_off_switch():
    shape = (4, 3, 3)
    rng = np.random.default_rng(7)
    F = 0.05 + 0.1 * rng.random(shape + (Q,))
    rho = 1.0 + 0.3 * rng.standard_normal(shape)
    v = 0.2 * rng.standard_normal(shape + (3,))
    args = dict(bc=(1, 2), rho_bc=(1.5, 1.0), vel_bc=((0, 0, 0), (0.3, 0.0, 0.0)))
    plain, k0 = _bc_open(shape, "regularized_velocity_pressure", F, rho, v, **args)
    off, k1 = _bc_open(shape, "regularized_velocity_pressure_limited", F, rho, v, **args)
    np.testing.assert_array_equal(plain, off)
    assert not k0.any() and not k1.any()
    lim = dict(open_boundary_limiter_enabled=True, open_boundary_rho_min=0.9, open_boundary_rho_max=1.1,
               open_boundary_u_max=0.05, open_boundary_noneq_cap=0.01, open_boundary_population_floor=0.0)
    out, k = _bc_open(shape, "regularized_velocity_pressure_limited", F, rho, v, **args, **lim)
    assert k[4] == 2 * 3 * 3 * 5
    for face, inner, unknown, kind in ((0, 1, UNKNOWN_MIN, 1), (3, 2, UNKNOWN_MAX, 2)):
        for j in range(3):
            for kk in range(3):
                n = (inner, j, kk)
                tr = 1.5 if kind == 1 else rho[n]
                tu = v[n] if kind == 1 else np.array([0.3, 0.0, 0.0])
                tr = min(max(tr, 0.9), 1.1)
                if np.linalg.norm(tu) > 0.05:
                    tu = tu * (0.05 / np.linalg.norm(tu))
                for s in unknown:
                    ne = np.clip(F[n + (s,)] - feq64(s, rho[n], v[n]), -0.01, 0.01)
                    assert out[face, j, kk, s] == pytest.approx(max(feq64(s, tr, tu) + ne, 0.0), abs=1e-14)


# ------------------------------------------------------- product vs oracle
def test_lbm_boundary_condition_reference_counter_contract(emu):
    """The reference's counter test through LBMFluid3D.Boundary_condition."""
    _, F, rho, v, kw = _reference_counter_case()
    lbm = LBMFluid3D(LBMConfig(nx=5, ny=4, nz=4, open_boundary_semantics="regularized_velocity_pressure_limited",
                               bc_x_left=1, rho_bc_x_left=2.0, **kw), native=emu, device="cpu")
    lbm.init_simulation()
    vv = lbm.v.to_numpy()
    vv[1, :, :, 0] = 1.0
    lbm.v.from_numpy(vv.astype(np.float32))
    FF = lbm.F.to_numpy()
    FF[1, :, :, :] = 2.0
    lbm.F.from_numpy(FF.astype(np.float32))
    lbm.clear_open_boundary_limiter_run_counters()
    lbm.Boundary_condition()
    stats = lbm.get_open_boundary_limiter_stats()
    expected = 5 * 4 * 4
    assert stats["reconstructed_population_count_step"] == expected
    assert stats["reconstructed_population_count_run"] == expected
    assert stats["rho_clip_count_step"] == expected
    assert stats["velocity_clip_count_step"] == expected
    assert stats["noneq_clip_count_step"] == expected
    assert stats["population_floor_count_step"] > 0
    assert stats["limiter_activation_denominator"] == expected
    assert stats["limiter_activation_count"] >= (stats["rho_clip_count_step"] + stats["velocity_clip_count_step"]
                                                 + stats["noneq_clip_count_step"]
                                                 + stats["population_floor_count_step"])
    want, k = _bc_open((5, 4, 4), "regularized_velocity_pressure_limited", F, rho, v, bc=(1, 0), rho_bc=(2.0, 1.0),
                       dtype=np.float32, **kw)
    np.testing.assert_allclose(lbm.F.to_numpy(), want, rtol=1e-6, atol=1e-7)
    assert [stats[f"{n}_count_step"] for n in ("rho_clip", "velocity_clip", "noneq_clip", "population_floor",
                                               "reconstructed_population")] == list(k)
    lbm.Boundary_condition()
    assert lbm.get_open_boundary_limiter_stats()["reconstructed_population_count_run"] == 2 * expected
    lbm.clear_open_boundary_limiter_step_counters()
    st = lbm.get_open_boundary_limiter_stats()
    assert st["reconstructed_population_count_step"] == 0 and st["reconstructed_population_count_run"] == 2 * expected
    lbm.set_open_boundary_limiter_run_counters(1, 2, 3, 4, 5)
    st = lbm.get_open_boundary_limiter_stats()
    assert [st[f"{n}_count_run"] for n in ("rho_clip", "velocity_clip", "noneq_clip", "population_floor",
                                           "reconstructed_population")] == [1, 2, 3, 4, 5]


@pytest.mark.parametrize("sem,kw", MODES, ids=[m[0] for m in MODES])
def test_lbm_boundary_condition_matches_oracle_on_random_state(emu, sem, kw):
    """Standalone Boundary_condition on random populations and stored
    moments, both faces, a solid cell next to each face."""
    shape = (6, 5, 4)
    rng = np.random.default_rng(11)
    F = (W * (1.0 + 0.2 * rng.standard_normal(shape + (Q,)))).astype(np.float32)
    v = (0.05 * rng.standard_normal(shape + (3,))).astype(np.float32)
    rho = F.sum(-1).astype(np.float32)
    solid = np.zeros(shape, np.int8)
    solid[1, 2, 1] = 1
    solid[4, 3, 2] = 1
    for bc, rb in (((2, 1), (1.0, 0.98)), ((1, 2), (1.02, 1.0))):
        lbm = LBMFluid3D(LBMConfig(nx=6, ny=5, nz=4, open_boundary_semantics=sem, bc_x_left=bc[0], bc_x_right=bc[1],
                                   rho_bc_x_left=rb[0], rho_bc_x_right=rb[1], vel_bc_x_left=(0.04, 0.0, 0.01),
                                   vel_bc_x_right=(0.03, -0.01, 0.0), **kw), native=emu, device="cpu")
        lbm.init_simulation()
        lbm.set_solid_flags(solid)
        lbm.F.from_numpy(F)
        lbm.rho.from_numpy(rho)
        lbm.v.from_numpy(v)
        lbm.Boundary_condition()
        want, k = _bc_open(shape, sem, F, rho, v, solid=solid, bc=bc, rho_bc=rb,
                           vel_bc=((0.04, 0.0, 0.01), (0.03, -0.01, 0.0)), dtype=np.float32, **kw)
        got = lbm.F.to_numpy()
        np.testing.assert_allclose(got, want, rtol=2e-6, atol=1e-7)
        assert not np.array_equal(want, F)  # the pass did rewrite the faces
        st = lbm.get_open_boundary_limiter_stats()
        assert [st[f"{n}_count_step"] for n in ("rho_clip", "velocity_clip", "noneq_clip", "population_floor",
                                                "reconstructed_population")] == list(k)


@pytest.mark.parametrize("sem,kw", MODES, ids=[m[0] for m in MODES])
def test_step_then_boundary_condition_uses_the_steps_moments(emu, sem, kw):
    """step() and then a standalone Boundary_condition(): the pass reads the
    rho, v that the step's streaming3 left (fluid.py:1027-1030, 2099-2103),
    i.e. the moments of the populations after the step, not older ones."""
    shape = (6, 5, 4)
    rng = np.random.default_rng(5)
    F0 = (W * (1.0 + 0.05 * rng.standard_normal(shape + (Q,)))).astype(np.float32)
    lbm = LBMFluid3D(LBMConfig(nx=6, ny=5, nz=4, open_boundary_semantics=sem, bc_x_left=2, bc_x_right=1,
                               rho_bc_x_right=1.0, vel_bc_x_left=(0.04, 0.0, 0.0), **kw), native=emu, device="cpu")
    lbm.init_simulation()
    lbm.F.from_numpy(F0)
    lbm.rho.from_numpy(F0.sum(-1))
    lbm.v.from_numpy((F0 @ orc.E.astype(np.float32)) / F0.sum(-1)[..., None])
    lbm.step()
    F1 = lbm.F.to_numpy()
    lbm.Boundary_condition()
    rho1 = F1.astype(np.float64).sum(-1)
    v1 = (F1.astype(np.float64) @ orc.E.astype(np.float64)) / rho1[..., None]
    want, _ = _bc_open(shape, sem, F1, rho1, v1, bc=(2, 1), rho_bc=(1.0, 1.0),
                       vel_bc=((0.04, 0.0, 0.0), (0.0, 0.0, 0.0)), **kw)
    np.testing.assert_allclose(lbm.F.to_numpy(), want, rtol=0, atol=2e-6)
    assert np.abs(want - F1).max() > 1e-4  # the pass rewrote the faces


def _with_ob(w, sem, kw):
    return dataclasses.replace(w, config=FSIDriverConfig(**{**w.config.to_dict(), "lbm_open_boundary_semantics": sem,
                                                            **kw}))


@pytest.mark.parametrize("schedule", ["penalty", "moving_boundary", "moving_boundary_unsubcycled"])
@pytest.mark.parametrize("sem,kw", MODES, ids=[m[0] for m in MODES])
def test_fsi_runs_match_oracle(emu, tmp_path, schedule, sem, kw):
    """Whole FSI steps with an open boundary: the face cells stream like inner
    cells and the pass after each substep rewrites their unknowns, as the
    reference's streaming then Boundary_condition; the limiter's run counters
    equal the oracle's."""
    if schedule == "penalty":
        w = workloads.penalty_block(n=12, steps=4, block=(4, 4, 2), out_dir=str(tmp_path))
    elif schedule == "moving_boundary":
        w = workloads.moving_boundary_flap(n=12, steps=2, substeps=6, mpm_substeps=2, flap=(2, 6, 4),
                                           out_dir=str(tmp_path))
    else:
        w = workloads.moving_boundary_unsubcycled(n=12, steps=4, flap=(2, 6, 4), out_dir=str(tmp_path))
    w = _with_ob(w, sem, kw)
    d, o = make_pair(w, tmp_path / "run", emu, "cpu", with_fp64=True)
    history = run_and_compare(d, o, w.config.n_lbm_steps)
    check_history(history, f"{schedule}/{sem}", 1.0e-5, strict=schedule != "moving_boundary")
    st = d.lbm.get_open_boundary_limiter_stats()
    run = [st[f"{n}_count_run"] for n in ("rho_clip", "velocity_clip", "noneq_clip", "population_floor",
                                          "reconstructed_population")]
    assert run == [int(x) for x in o.ob_run]
    if sem.endswith("limited"):
        assert run[4] > 0 and sum(run[:4]) > 0  # the limiter was exercised


def test_batched_form_refuses_open_boundaries(emu, tmp_path):
    from mpm_lbm.batch import BatchedFSI

    w = _with_ob(workloads.penalty_block(n=12, steps=1, block=(4, 4, 2), out_dir=str(tmp_path)),
                 "regularized_velocity_pressure", {})
    with pytest.raises(NotImplementedError, match="legacy x-face reset"):
        BatchedFSI([w], native=emu, device="cpu", out_dir=str(tmp_path / "b"))
