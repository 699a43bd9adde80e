"""The resident MPM subcycle (fsi_mpm_subcycle_coupled: n MPM substeps in one
launch, particles on chip, three grid buffers rotated with one barrier per
substep) against n calls of fsi_mpm_substep_coupled, both through the serial
CPU emulation.

The emulation runs the resident form phase by phase as the GPU does (every
particle block's scatter into buffer (n - 1 - s) mod 3, its clear of the
nodes of its stencil box two substeps back in buffer (n + 1 - s) mod 3, then
every block's gather with the grid update applied per node), so a wrong
buffer, a missed or early clear, or a wrong final conversion changes the
result.  Every field must agree bit for bit: particles, Jp, the state's grid
after the call, the reaction counters."""
import ctypes

import numpy as np
import pytest

from mpm_lbm import _native, workloads
from mpm_lbm.sim.drivers.fsi_config import FSIDriverConfig
from mpm_lbm.sim.drivers.fsi_driver import FSIDriver3D
from tests.parity import driver_fields


def _driver(w, out, emu, resident, mpm_substeps):
    cfg = FSIDriverConfig(**{**w.config.to_dict(), "mpm_substeps_per_lbm_step": mpm_substeps})
    d = FSIDriver3D(cfg, str(out), particle_cloud=w.cloud, native=emu, device="cpu", write_geometry_file=False)
    d.initialize()
    d.solid.resident_subcycle = resident
    d.solid.resident_min_substeps = 1
    return d


def _state(d):
    f = driver_fields(d)
    s = d.solid
    for k in ("_grid_v", "_grid_m", "_grid_f", "_Jp"):
        f[k] = getattr(s, k).numpy().copy()
    coupler = d.mb_coupler or d.penalty_coupler
    if coupler is not None:
        f["counters"] = coupler._counters_buf.numpy().copy()
    return f


def _assert_same(a, b, where):
    for k in a:
        x, y = np.asarray(a[k]), np.asarray(b[k])
        same = np.array_equal(x.view(np.uint8), y.view(np.uint8)) if x.dtype.kind == "f" else np.array_equal(x, y)
        if not same:
            diff = np.argwhere(x != y)
            raise AssertionError(f"{where}: {k} differs at {len(diff)} entries, first {diff[:3].tolist()}")


def _case(mode, tmp_path):
    if mode == "moving_boundary":
        return workloads.moving_boundary_flap(n=16, steps=2, substeps=4, flap=(2, 6, 6), mpm_substeps=2,
                                              out_dir=str(tmp_path))
    return workloads.flap_case(mode, 16, substeps=4, thickness=2, out_dir=str(tmp_path))


@pytest.mark.parametrize("mode", ["moving_boundary", "penalty", "none"])
@pytest.mark.parametrize("n", [1, 2, 3, 7])
def test_mpm_resident_matches_substeps(emu, tmp_path, mode, n):
    w = _case(mode, tmp_path)
    if mode == "none":
        w.config = FSIDriverConfig(**{**w.config.to_dict(), "coupling_mode": "none"})
    a = _driver(w, tmp_path / "launch", emu, False, n)
    b = _driver(w, tmp_path / "res", emu, True, n)
    for step in range(2):
        a.step_once()
        b.step_once()
        _assert_same(_state(a), _state(b), f"{mode}, n = {n}, FSI step {step}")
    assert b.solid._res_sync is not None and not b.solid._res_declined
    assert a.solid._res_sync is None


def test_mpm_resident_entry_point_rejects(emu, tmp_path):
    """FSI_E_UNSUPPORTED (nothing run) without particles or with a reaction
    stencil on another spacing than the MPM grid; FSI_E_INVALID for n < 1 or
    missing scratch / sync."""
    w = _case("moving_boundary", tmp_path)
    d = _driver(w, tmp_path / "d", emu, True, 2)
    d.step_once()
    s, lbm, c = d.solid, d.lbm, d.mb_coupler
    ws = np.zeros(_native.coupled_workspace_doubles(s.n_particles))
    scratch = np.zeros(_native.mpm_res_scratch_floats(int(np.prod(s.grid_shape))), dtype=np.float32)
    sync = np.zeros(_native.MPM_RES_SYNC_WORDS, dtype=np.int32)
    P = ctypes.c_void_p

    def call(mp, cp, n, scr=scratch, syn=sync):
        return emu.lib.fsi_mpm_subcycle_coupled(ctypes.byref(lbm._params), ctypes.byref(lbm._state), ctypes.byref(mp),
                                                ctypes.byref(s._state), ctypes.byref(cp),
                                                _native.ptr(c._counters_buf), 2, n, P(ws.ctypes.data),
                                                P(scr.ctypes.data) if scr is not None else None,
                                                P(syn.ctypes.data) if syn is not None else None, None)

    mp = _native.MpmParams.from_buffer_copy(s._params)
    cp = _native.CouplingParams.from_buffer_copy(c.params)
    cp.inv_dx_norm = mp.inv_dx * 2.0
    assert call(mp, cp, 2) == _native.E_UNSUPPORTED
    cp.inv_dx_norm = mp.inv_dx
    empty = _native.MpmParams.from_buffer_copy(s._params)
    empty.n_particles = 0
    assert call(empty, cp, 2) == _native.E_UNSUPPORTED
    assert call(mp, cp, 0) == _native.E_INVALID
    assert call(mp, cp, 2, scr=None) == _native.E_INVALID
    assert call(mp, cp, 2, syn=None) == _native.E_INVALID
