"""Checks that pin the oracle where the reference's artifacts do not:
closed forms, invariants and library routines (numpy.linalg)."""
import ctypes

import numpy as np
import pytest

from oracle import oracle as orc


def _lbm(nx, ny, nz, solid=None, dtype=np.float64, tau=0.6, bc=(0, 0), vel_l=(0, 0, 0), force=(0, 0, 0)):
    from types import SimpleNamespace

    cfg = SimpleNamespace(
        n_grid=nx, lattice_shape=(nx, ny, nz), lbm_dt_phys_override_s=None, mpm_substeps_per_lbm_step=1, mpm_dt=1e-4,
        lbm_viscosity_semantics="physical_nu_mapping", physical_duct_length_m=1.0,
        fluid_kinematic_viscosity_m2_s=(tau - 0.5) / 3.0 * (1.0 / nx) ** 2 / 1e-4,
        lbm_boundary_condition_mode="default_periodic", velocity_inlet_side="min", target_u_lbm=(0, 0, 0),
        gravity=(0, 0, 0), initial_solid_velocity_norm=(0, 0, 0), beta_lbm=1e-3, penalty_force_cap_lbm=1e-4,
        coupling_mode="none", mb_reaction_scale=1.0, mb_force_cap_norm=1e-4, fsi_exchange_mode="one_lbm_step_per_fsi_step",
        mpm_substeps_per_lbm_step_=1, dynamic_solid_threshold=0.5, lbm_substeps_per_fsi_step=1)
    flags = np.zeros((nx, ny, nz), np.int8) if solid is None else solid
    particles = {"x": np.full((1, 3), 0.5, dtype), "mass": np.ones(1, dtype), "vol0": np.ones(1, dtype)}
    o = orc.OracleFSI(cfg, flags, particles, (400.0, 0.2, 1.0), dtype=dtype)
    o.la.bc_x_left, o.la.bc_x_right = bc
    o.la.vel_bc_x_left[:] = list(vel_l)
    o.la.fext[:] = list(force)
    return o


def test_inverse_moment_matrix():
    iM = orc.inv_M(np.float64)
    # reconstruct M from the oracle's source-of-truth table via M = inv(inv_M)
    Mr = np.linalg.inv(iM)
    assert np.allclose(Mr, np.round(Mr), atol=1e-9)
    Mr = np.round(Mr)
    assert np.allclose(Mr @ iM, np.eye(19), atol=1e-13)
    e = orc.E.astype(np.float64)
    # rows 0, 3, 5, 7 are density and momentum
    assert np.array_equal(Mr[0], np.ones(19))
    assert np.array_equal(Mr[3], e[:, 0]) and np.array_equal(Mr[5], e[:, 1]) and np.array_equal(Mr[7], e[:, 2])
    # inv_M has the sparsity of M^T and entries k/144
    assert np.array_equal(np.abs(iM) > 1e-12, np.abs(Mr.T) > 0)
    assert np.allclose(iM * 144, np.round(iM * 144), atol=1e-9)
    # the f32 table is the f64 inverse narrowed (fluid.py:248, :277)
    assert np.array_equal(orc.inv_M(np.float32), iM.astype(np.float32))


def test_d3q19_tables():
    e = orc.E
    LR = [0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17]
    w = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
    assert all(LR[LR[s]] == s and np.array_equal(e[LR[s]], -e[s]) for s in range(19))
    assert abs(w.sum() - 1.0) < 1e-15
    assert np.allclose((w[:, None, None] * e[:, :, None] * e[:, None, :]).sum(0), np.eye(3) / 3)


def test_guo_term_mass_and_momentum_change_closed_form():
    """Sum_l g_l = -(8/27) u.F and sum_l e_l g_l = F/9 for the reference's Guo
    term w((e-u).F/3 + (e.u)(e.F)/9) (fluid.py:408); conserved moments relax
    with S = 0, so collision changes mass and momentum by exactly these."""
    o = _lbm(3, 3, 3)
    u = np.array([0.031, -0.017, 0.011])
    Fc = np.array([2.0e-4, 1.0e-4, -3.0e-4])
    o.v[:] = np.tile(u, 27)
    o.rho[:] = 1.0
    o.cell_force[:] = np.tile(Fc, 27)
    F0 = o.F.copy()
    o.colission()
    dm = (o.f - F0).reshape(27, 19)
    assert np.allclose(dm.sum(1), -8.0 / 27.0 * u.dot(Fc), rtol=0, atol=1e-14)
    assert np.allclose(dm @ orc.E.astype(np.float64), Fc / 9.0, rtol=0, atol=1e-14)


def test_rest_state_is_a_fixed_point():
    o = _lbm(5, 4, 6)
    F0 = o.F.copy()
    for _ in range(5):
        o.lbm_step()
    assert np.allclose(o.F, F0, rtol=0, atol=1e-15)


def test_mass_conserved_with_static_bounce_back():
    rng = np.random.default_rng(0)
    solid = np.zeros((8, 7, 6), np.int8)
    solid[:, 0, :] = solid[:, -1, :] = 1
    solid[3:5, 2:4, 2:4] = 1
    o = _lbm(8, 7, 6, solid=solid)
    o.F *= 1.0 + 0.01 * rng.standard_normal(o.F.shape)
    o._call("streaming3", ctypes.byref(o.la), orc._p(o.F), orc._p(o.f), orc._p(o.rho), orc._p(o.v), orc._p(o.solid),
            orc._p(o.cell_force))
    fluid = (solid.reshape(-1) == 0)
    m0 = o.rho[fluid].sum()
    for _ in range(20):
        o.lbm_step()
    assert abs(o.rho[fluid].sum() - m0) < 1e-11 * m0


def test_static_bounce_back_equals_moving_with_zero_wall_velocity():
    """run_step8_static_bounceback_regression.py:107-118: bit-exact in the reference."""
    rng = np.random.default_rng(1)
    solid = np.zeros((6, 6, 6), np.int8)
    solid[:, 0, :] = solid[:, -1, :] = 1
    a = _lbm(6, 6, 6, solid=solid)
    a.F *= 1.0 + 0.01 * rng.standard_normal(a.F.shape)
    b = _lbm(6, 6, 6, solid=solid)
    b.F[:] = a.F
    for _ in range(5):
        a.lbm_step()
        b.lbm_step_moving_bounceback()
    assert np.array_equal(a.F, b.F)
    # link force balance: fluid impulse = -solid force (fluid.py:510-519)
    st = b.mbstats
    assert np.allclose(np.array(st.net_fluid_impulse), -np.array(st.net_solid_force), rtol=0, atol=1e-15)
    assert st.link_count == 2 * 6 * 6 * 5  # 5 directions leave each wall-adjacent cell towards its wall


def test_velocity_inlet_faces_are_equilibrium():
    solid = np.zeros((6, 5, 5), np.int8)
    o = _lbm(6, 5, 5, solid=solid, bc=(2, 1), vel_l=(0.04, 0.0, 0.0))
    o.lbm_step()
    F = o.F.reshape(6, 5, 5, 19)
    e = orc.E.astype(np.float64)
    w = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
    u = np.array([0.04, 0, 0])
    eu = e @ u
    feq = w * (1 + 3 * eu + 4.5 * eu ** 2 - 1.5 * u.dot(u))
    assert np.allclose(F[0], feq, rtol=0, atol=1e-15)


def test_accumulator_known_answer():
    """tests/test_step115_...:20-45: samples (1,0,0) and (3,0,0) -> mean (2,0,0), accum max 4, mean max 2."""
    o = _lbm(4, 4, 4, dtype=np.float32)
    o.clear_accumulator()
    for fx in (1.0, 3.0):
        o.hydro[:] = 0
        c = (1 * 4 + 1) * 4 + 1
        o.hydro[3 * c] = fx
        o.accumulate()
    o.finalize_accumulator()
    assert np.allclose(o.hydro.reshape(4, 4, 4, 3)[1, 1, 1], [2.0, 0, 0], atol=1e-6)
    assert o.mb_sample_count == 2
    assert o.accum_norm_max == pytest.approx(4.0, abs=1e-6)
    assert o.mean_norm_max == pytest.approx(2.0, abs=1e-6)


@pytest.mark.parametrize("seed", range(4))
def test_polar_rotation_against_numpy_svd(seed):
    rng = np.random.default_rng(seed)
    F = np.eye(3) + 0.3 * rng.standard_normal((3, 3))
    if seed == 3:
        F[:, 0] *= -1.0  # inverted element: det F < 0
    R = orc.polar(F)
    U, s, Vt = np.linalg.svd(F)
    if np.linalg.det(U) < 0:
        U[:, 2] *= -1
    if np.linalg.det(Vt) < 0:
        Vt[2] *= -1
    R_ref = U @ Vt  # ti.svd convention: det U = det V = +1
    assert np.allclose(R, R_ref, atol=1e-12)
    assert np.allclose(R.T @ R, np.eye(3), atol=1e-12) and abs(np.linalg.det(R) - 1) < 1e-12
    if np.linalg.det(F) > 0:
        S = R.T @ F
        assert np.allclose(S, S.T, atol=1e-12) and np.all(np.linalg.eigvalsh((S + S.T) / 2) > 0)

