"""The resident moving-wall subcycle (fsi_lbm_subcycle_resident: the lattice
stays on chip, pull form) against the launch-per-substep push form
(fsi_lbm_substep), both through the serial CPU emulation of the kernels.

The two forms evaluate the same expressions (collision, bounce-back, face
reset, wall sums in the same order), so every field must agree bit for bit
after every FSI step: populations, moments, flags, the fp64 wall-force
accumulator and what the finalize makes of it, the last substep's link
buffers and statistics, and the particles.  The emulation runs the resident
form block by block as the GPU does (per-block value planes, halo planes at
the sender's index), so its block and halo addressing is checked here too."""
import numpy as np
import pytest

from mpm_lbm import _native, workloads
from mpm_lbm.sim.drivers.fsi_driver import FSIDriver3D
from tests.parity import driver_fields


def _driver(w, out, emu, resident):
    d = FSIDriver3D(w.config, str(out), particle_cloud=w.cloud, native=emu, device="cpu", write_geometry_file=False)
    d.initialize()
    d.lbm.resident_subcycle = resident
    return d


def _state(d):
    lbm = d.lbm
    f = driver_fields(d)
    f["accum"] = lbm._accum.numpy().copy()
    f["cellinfo"] = lbm._cellinfo.numpy().copy()
    if lbm._last_link_parity is not None:
        f["link_last"] = lbm._link[lbm._last_link_parity].numpy().copy()
    f["link_corr"] = lbm._link_corr.numpy().copy()
    f["mbs"] = lbm.get_moving_boundary_stats()
    f["acc"] = lbm.get_moving_boundary_accumulation_stats()
    return f


def _assert_same(a, b, where):
    assert a.keys() == b.keys()
    for k in a:
        if isinstance(a[k], dict):
            assert a[k] == b[k], f"{where}: {k} differs: {a[k]} vs {b[k]}"
        else:
            x, y = np.asarray(a[k]), np.asarray(b[k])
            same = np.array_equal(x.view(np.uint8), y.view(np.uint8)) if x.dtype.kind == "f" else np.array_equal(x, y)
            if not same:
                diff = np.argwhere(x != y)
                raise AssertionError(f"{where}: {k} differs at {len(diff)} entries, first {diff[:3].tolist()}")


def _resident_ran(d, substeps_total):
    lbm = d.lbm
    assert lbm._res_sync is not None and not lbm._res_declined
    words = lbm._res_sync.numpy()
    nblk = _native.res_blocks(lbm.nx, lbm.ny)
    assert np.all(words[0:64 * nblk:64] == substeps_total)
    assert words[64 * nblk] == 0  # no neighbour wait timed out


@pytest.mark.parametrize("substeps", [6, 5])
def test_resident_matches_push_form(emu, tmp_path, substeps):
    w = workloads.moving_boundary_flap(shape=(16, 12, 64), steps=3, substeps=substeps, flap=(2, 6, 6),
                                       mpm_substeps=2, out_dir=str(tmp_path))
    a = _driver(w, tmp_path / "push", emu, resident=False)
    b = _driver(w, tmp_path / "res", emu, resident=True)
    for step in range(3):
        a.step_once()
        b.step_once()
        _assert_same(_state(a), _state(b), f"FSI step {step}")
    _resident_ran(b, 3 * substeps)
    assert a.lbm._res_sync is None
    rows = [d.collect_diagnostics(3) for d in (a, b)]
    for r in rows:
        r.pop("elapsed_seconds")
    assert rows[0] == rows[1]


def test_resident_cell_force_pressure_inlet_no_accumulate(emu, tmp_path):
    """FSI_SUB_CELL_FORCE, a pressure face on both sides (the left one keeps
    its own velocity, the right one that of the cell behind it), an odd
    subcycle without the accumulator (its links stay pending)."""
    w = workloads.moving_boundary_flap(shape=(12, 16, 64), steps=2, substeps=4, flap=(2, 8, 6), mpm_substeps=2,
                                       out_dir=str(tmp_path))
    runs = []
    for resident in (False, True):
        d = _driver(w, tmp_path / f"r{int(resident)}", emu, resident=resident)
        d.step_once()
        lbm = d.lbm
        lbm.set_uniform_cell_force(2.0e-5, -1.0e-5, 0.5e-5)
        lbm.bc_x_left = _native.FSI_BC_PRESSURE
        lbm._build_params()
        lbm.run_moving_boundary_subcycle(7, d.config.dynamic_solid_threshold, accumulate=False)
        runs.append(d)
    assert runs[1].lbm._cell_force_active
    _resident_ran(runs[1], 4 + 7)
    _assert_same(_state(runs[0]), _state(runs[1]), "cell-force subcycle")


def test_resident_declines_other_shapes(emu, tmp_path):
    """A lattice the resident form does not cover (nz != 64) runs the push form."""
    w = workloads.moving_boundary_flap(n=16, steps=1, substeps=4, flap=(2, 6, 6), mpm_substeps=1,
                                       out_dir=str(tmp_path))
    d = _driver(w, tmp_path / "d", emu, resident=True)
    d.step_once()
    assert d.lbm._res_sync is None


def test_resident_entry_point_rejects(emu):
    """fsi_lbm_subcycle_resident returns FSI_E_UNSUPPORTED (not a crash) for
    what it does not cover."""
    import ctypes

    p = _native.LbmParams()
    p.nx, p.ny, p.nz = 16, 16, 32
    st = _native.LbmState()
    rc = emu.lib.fsi_lbm_subcycle_resident(ctypes.byref(p), ctypes.byref(st), 0, 4, _native.SUB_MOVING_WALL, None,
                                           None, None)
    assert rc == _native.E_UNSUPPORTED
