"""The resident moving-wall subcycle on the GPU (k_lbm_resident, the lattice
on chip for the whole subcycle) against the launch-per-substep push form
(k_lbm_substep) of the same library.

With mpm_transfer_order="particle" the runs are deterministic and use the
library built without fma fusion, where both forms evaluate the same
rounded expressions: every field must agree bit for bit after every FSI step,
at 64^3 with 120 substeps (the second step replayed from a captured graph)
and on a small lattice with an odd subcycle.  The default build's resident
path is held to the oracle by the 64^3 subcycled parity tests
(test_gpu_parity.py), which run through it."""
import dataclasses

import numpy as np
import pytest

from mpm_lbm import _native, workloads
from mpm_lbm.sim.drivers.fsi_config import FSIDriverConfig
from mpm_lbm.sim.drivers.fsi_driver import FSIDriver3D
from tests.parity import driver_fields

pytestmark = pytest.mark.gpu


def _driver(w, out, resident, order="particle"):
    cfg = FSIDriverConfig(**{**w.config.to_dict(), "mpm_transfer_order": order})
    d = FSIDriver3D(cfg, str(out), particle_cloud=w.cloud, device="cuda", write_geometry_file=False)
    d.initialize()
    d.lbm.resident_subcycle = resident
    return d


def _state(d):
    lbm = d.lbm
    f = driver_fields(d)
    f["accum"] = lbm._accum.cpu().numpy()
    f["cellinfo"] = lbm._cellinfo.cpu().numpy()
    f["link_last"] = lbm._link[lbm._last_link_parity].cpu().numpy()
    f["link_corr"] = lbm._link_corr.cpu().numpy()
    f["mbs"] = lbm.get_moving_boundary_stats()
    f["acc"] = lbm.get_moving_boundary_accumulation_stats()
    return f


def _differences(a, b):
    out = {}
    for k in a:
        if isinstance(a[k], dict):
            if a[k] != b[k]:
                out[k] = (a[k], b[k])
            continue
        x, y = np.asarray(a[k]), np.asarray(b[k])
        if x.dtype.kind == "f":
            if not np.array_equal(x.view(np.uint8), y.view(np.uint8)):
                scale = float(np.max(np.abs(y))) or 1.0
                out[k] = (int(np.count_nonzero(x != y)), float(np.max(np.abs(x - y))) / scale)
        elif not np.array_equal(x, y):
            out[k] = int(np.count_nonzero(x != y))
    return out


def _resident_ran(d, substeps_total):
    lbm = d.lbm
    assert lbm._res_sync is not None and not lbm._res_declined
    lbm.check_resident_waits()
    words = lbm._res_sync.cpu().numpy()
    nblk = _native.res_blocks(lbm.nx, lbm.ny)
    assert np.all(words[0:64 * nblk:64] == substeps_total)


@pytest.mark.parametrize("shape,substeps,flap,mpm", [((64, 64, 64), 120, (4, 32, 24), 10),
                                                     ((16, 12, 64), 7, (2, 6, 6), 2)])
def test_gpu_resident_bit_exact(tmp_path, shape, substeps, flap, mpm):
    print(f"[resident] {shape} x {substeps} substeps, particle order", flush=True)
    w = workloads.moving_boundary_flap(shape=shape, steps=2, substeps=substeps, flap=flap, mpm_substeps=mpm,
                                       out_dir=str(tmp_path))
    a = _driver(w, tmp_path / "push", resident=False)
    b = _driver(w, tmp_path / "res", resident=True)
    for step in range(2):
        a.step_once()
        b.step_once()
        diff = _differences(_state(a), _state(b))
        print(f"[resident] FSI step {step}: differing fields {diff}", flush=True)
        assert not diff, f"FSI step {step}: {diff}"
    _resident_ran(b, 2 * substeps)
    assert a.lbm._res_sync is None
    if substeps % 2 == 0:
        assert b._fsi_graph is not None  # the second step replayed the captured graph


def test_gpu_resident_default_build_runs(tmp_path):
    """The default build takes the resident subcycle at 64^3: its counters
    advance, no wait timed out, and the flow stays finite and close to the
    push form (the oracle comparison is test_gpu_parity's)."""
    w = workloads.moving_boundary_flap(n=64, steps=1, substeps=120, out_dir=str(tmp_path))
    a = _driver(w, tmp_path / "push", resident=False, order="atomic")
    b = _driver(w, tmp_path / "res", resident=True, order="atomic")
    a.step_once()
    b.step_once()
    _resident_ran(b, 120)
    fa, fb = driver_fields(a), driver_fields(b)
    fluid = fa["solid"] == 0
    assert np.array_equal(fa["solid"], fb["solid"])
    for k in ("F", "rho", "v"):
        x, y = fa[k][fluid], fb[k][fluid]
        assert np.all(np.isfinite(y))
        err = float(np.max(np.abs(x - y))) / float(np.max(np.abs(x)))
        print(f"[resident] default build {k}: {err:.3e}", flush=True)
        assert err < 1.0e-5
