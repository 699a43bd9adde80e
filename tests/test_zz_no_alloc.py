"""fsi_mpm_substep_coupled_noalloc (the allocation-free variant, reaction
records in the library's fixed per-GPU scratch) against
fsi_mpm_substep_coupled with a caller workspace, through the C ABI, on the same
inputs: two calls in a row, the second on a smaller particle set, so the
second reads scratch the first left behind.  Bar: the integer counters equal,
every float output within 1e-5 of the other entry point's field magnitude
(the P2G tile flush adds into the grid with float atomics, so two runs of
either entry point differ in the last bits)."""
import ctypes

import numpy as np
import pytest
import torch

from mpm_lbm import _native, workloads
from mpm_lbm.sim.drivers.fsi_driver import FSIDriver3D
from tests.parity import rel_max_err

pytestmark = pytest.mark.gpu

SOLID_TENSORS = ("_x", "_v", "_C", "_F", "_Jp", "_grid_v", "_grid_m", "_grid_f")
COUNTER_INTS = ("active_reaction_particle_count",)
COUNTER_FLOATS = ("max_particle_reaction_bits", "max_grid_reaction_bits")
COUNTER_VECS = ("net_particle_reaction", "net_grid_reaction")


def _driver(kind, tmp_path):
    if kind == "moving_boundary":
        w = workloads.moving_boundary_flap(n=32, steps=1, substeps=20, flap=(2, 16, 12), out_dir=str(tmp_path))
        mode = 2
    else:
        w = workloads.penalty_block(n=16, steps=1, block=(4, 4, 2), out_dir=str(tmp_path))
        mode = 1
    d = FSIDriver3D(w.config, str(tmp_path / kind), particle_cloud=w.cloud, write_geometry_file=False)
    d.initialize()
    if mode == 1:
        d.solid.set_uniform_velocity(0.05, 0.0, 0.0)  # a moving block, so the penalty force is not zero
    d.step_once()  # a hydrodynamic field for the reaction to sample
    torch.cuda.synchronize()
    return d, mode


def _coupler(d, mode):
    return d.mb_coupler if mode == 2 else d.penalty_coupler


def _run(hip, d, mode, noalloc):
    """One MPM substep with the reaction from the driver's current state,
    through the named entry point; returns its outputs and restores the state."""
    s, lbm = d.solid, d.lbm
    cp = _coupler(d, mode)
    stream = cp._bind(lbm)
    lbm._sync_hydro()
    saved = {k: getattr(s, k).clone() for k in SOLID_TENSORS}
    counters = cp._counters_buf
    refs = (ctypes.byref(lbm._params), ctypes.byref(lbm._state), ctypes.byref(s._params), ctypes.byref(s._state),
            ctypes.byref(cp.params), _native.ptr(counters), int(mode))
    torch.cuda.synchronize()
    if noalloc:
        hip.call("fsi_mpm_substep_coupled_noalloc", *refs, stream)
    else:
        ws = torch.zeros(_native.coupled_workspace_doubles(s.n_particles), dtype=torch.float64, device=lbm.device)
        hip.call("fsi_mpm_substep_coupled", *refs, _native.ptr(ws), stream)
    torch.cuda.synchronize()
    out = {k: getattr(s, k).cpu().numpy().copy() for k in SOLID_TENSORS}
    out["counters"] = cp._counters()
    out["inv_dx"] = float(s._params.inv_dx)
    for k, t in saved.items():
        getattr(s, k).copy_(t)
    torch.cuda.synchronize()
    return out


def _compare(a, b, label):
    for k in SOLID_TENSORS:
        # APIC C is a velocity gradient: its scale is |v| / dx even where it nearly vanishes (tests/parity.py)
        floor = float(np.abs(b["_v"]).max()) * b["inv_dx"] if k == "_C" else 0.0
        e = rel_max_err(a[k], b[k], floor=floor)
        print(label, k, f"{e:.2e}", flush=True)
        assert e <= 1e-5, f"{label}: {k} differs by {e:.2e}"
    ca, cb = a["counters"], b["counters"]
    for k in COUNTER_INTS:
        assert getattr(ca, k) == getattr(cb, k), f"{label}: {k}"
    for k in COUNTER_FLOATS:
        fa = np.uint32(getattr(ca, k)).view(np.float32)
        fb = np.uint32(getattr(cb, k)).view(np.float32)
        assert abs(float(fa) - float(fb)) <= 1e-5 * abs(float(fb)), f"{label}: {k} {fa} vs {fb}"
    for k in COUNTER_VECS:
        va, vb = np.array(getattr(ca, k)), np.array(getattr(cb, k))
        assert np.abs(va - vb).max() <= 1e-5 * max(np.abs(vb).max(), 1e-30), f"{label}: {k} {va} vs {vb}"
        assert np.abs(vb).max() > 0, f"{label}: {k} is zero: the reaction did nothing"


def test_noalloc_matches_workspace_entry_point(hip, tmp_path):
    big, mode_big = _driver("moving_boundary", tmp_path)
    small, mode_small = _driver("penalty", tmp_path)
    assert small.solid.n_particles < big.solid.n_particles
    print("particles", big.solid.n_particles, small.solid.n_particles, flush=True)
    ref_big = _run(hip, big, mode_big, noalloc=False)
    ref_small = _run(hip, small, mode_small, noalloc=False)
    # two calls in a row on the shared scratch, the second smaller
    got_big = _run(hip, big, mode_big, noalloc=True)
    got_small = _run(hip, small, mode_small, noalloc=True)
    _compare(got_big, ref_big, "moving_boundary")
    _compare(got_small, ref_small, "penalty")


def test_noalloc_rejects_too_many_particles(hip, tmp_path):
    d, mode = _driver("penalty", tmp_path)
    cp = _coupler(d, mode)
    params = _native.MpmParams.from_buffer_copy(d.solid._params)
    params.n_particles = _native.NOALLOC_MAX_PARTICLES + 1
    rc = hip.lib.fsi_mpm_substep_coupled_noalloc(
        ctypes.byref(d.lbm._params), ctypes.byref(d.lbm._state), ctypes.byref(params), ctypes.byref(d.solid._state),
        ctypes.byref(cp.params), _native.ptr(cp._counters_buf), int(mode), cp._bind(d.lbm))
    assert rc == -1
    assert b"scratch limit" in hip.lib.fsi_last_error()
