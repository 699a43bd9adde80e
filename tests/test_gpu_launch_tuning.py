"""The fused substep's launch shapes (tile of 128/256/512 cells, register
budget, workgroup order) give bit-identical results, and the autotuner that
times them leaves the lattice's state exactly as it found it."""
import ctypes

import numpy as np
import pytest

from mpm_lbm import _native, tuning, workloads
from mpm_lbm.sim.drivers.fsi_driver import FSIDriver3D

pytestmark = pytest.mark.gpu


def _driver(tmp_path, hip):
    w = workloads.moving_boundary_flap(n=40, steps=2, substeps=8, out_dir=str(tmp_path / "w"))
    d = FSIDriver3D(w.config, str(tmp_path / "run"), particle_cloud=w.cloud, native=hip, write_geometry_file=False,
                    use_graph=False)
    d.initialize()
    d.step_once()  # moving walls, links and the accumulator in use
    return d


def _state(lbm):
    import torch

    torch.cuda.synchronize()
    out = [lbm._F[0].cpu().numpy().copy(), lbm._F[1].cpu().numpy().copy(), lbm._counters.cpu().numpy().copy()]
    out += [t.cpu().numpy().copy() for t in (lbm._link[0], lbm._link[1], lbm._link_corr, lbm._accum)]
    return out


def _restore(lbm, st):
    import torch

    for t, a in zip([lbm._F[0], lbm._F[1], lbm._counters, lbm._link[0], lbm._link[1], lbm._link_corr, lbm._accum], st):
        t.copy_(torch.from_numpy(a).to(t.device))
    torch.cuda.synchronize()


def _run(lbm, shape, n=4):
    p = lbm._params
    p.launch_block, p.launch_minw, p.launch_linear = shape
    cur, parity = lbm.cur, 0
    flags = _native.SUB_MOVING_WALL | _native.SUB_GATHER_PREV
    for _ in range(n):
        lbm.native.call("fsi_lbm_substep", ctypes.byref(p), ctypes.byref(lbm._state), cur, parity, flags,
                        lbm._stream())
        cur, parity = 1 - cur, 1 - parity
    return _state(lbm)


def test_gpu_launch_shapes_agree(hip, tmp_path):
    """Within one register-budget variant (one code path) every tile size and
    workgroup order gives bit-identical results; the two variants (wall cells
    queued to one wave, or run in place) compile the same arithmetic apart
    from fma contraction, so they agree to rounding."""
    d = _driver(tmp_path, hip)
    lbm = d.lbm
    start = _state(lbm)
    refs = {}
    for minw in (1, 5):
        _restore(lbm, start)
        refs[minw] = _run(lbm, (256, minw, 0))
    for k, (a, b) in enumerate(zip(refs[1], refs[5])):
        scale = max(float(np.abs(b).max()), 1e-30)
        assert float(np.abs(a.astype(np.float64) - b).max()) <= 1e-6 * scale, k
    shapes = tuning.candidates(lbm._params.storage)
    assert len(shapes) == 12
    for shape in shapes + [(0, 0, 0)]:
        _restore(lbm, start)
        got = _run(lbm, shape)
        minw = shape[1] or (1 if lbm._params.nx * lbm._params.ny * lbm._params.nz <= 64 ** 3 else 5)
        for k, (a, b) in enumerate(zip(got, refs[minw])):
            assert np.array_equal(a, b), (shape, k)


def test_gpu_tuner_restores_state(hip, tmp_path):
    d = _driver(tmp_path, hip)
    before = _state(d.lbm)
    rec = d.tune_launch(cache_path=str(tmp_path / "cache.json"), launches=6)
    after = _state(d.lbm)
    print("launch tuning", rec, flush=True)
    for k, (a, b) in enumerate(zip(before, after)):
        assert np.array_equal(a, b), k
    assert len(rec["us"]) == 12 and not rec["cached"]
    assert list(rec["choice"]) == [0, 0, 0] or tuple(rec["choice"]) in tuning.candidates(_native.STORE_F32)
    assert d.lbm._params.launch_block == rec["choice"][0]
    # the cache answers a second lattice of the same signature
    d2 = _driver(tmp_path / "b", hip)
    tuning._CACHE.clear()
    rec2 = d2.tune_launch(cache_path=str(tmp_path / "cache.json"))
    assert rec2["cached"] and rec2["choice"] == rec["choice"]
    d.step_once()  # the run goes on with the chosen shape
