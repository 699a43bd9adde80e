"""Per-cell bounce-back link classification (cellinfo, written by the GPU's
flag rebuild) against the oracle's, bit for bit, after every FSI step: the
Step-10 moving box (its flags change every step) and the 64^3 flap
workload.  tests/links.py states the bits."""
import numpy as np
import pytest

from mpm_lbm import workloads
from tests import links, step10
from tests.parity import make_pair

pytestmark = pytest.mark.gpu


def _run(d, o, steps, label):
    for step in range(steps):
        d.step_once()
        o.step_once()
        sol = o.fields()["solid"]
        assert np.array_equal(d.lbm.solid.to_numpy(), sol), (label, step)
        bad = links.cellinfo_problems(d, sol)
        print(label, step, "solid cells", int(sol.sum()), "problems", bad, flush=True)
        assert not bad, (label, step, bad)


def test_gpu_links_step10_moving_box(hip, tmp_path):
    from mpm_lbm.sim.drivers.fsi_driver import FSIDriver3D
    from oracle.oracle import OracleFSI

    cfg = step10.config("moving_boundary")
    cloud = step10.box_cloud(cfg)
    d = FSIDriver3D(cfg, str(tmp_path / "out"), particle_cloud=cloud, native=hip, write_geometry_file=False)
    d.initialize()
    o = OracleFSI(cfg, d.lbm.solid.to_numpy(), cloud, step10.MATERIAL)
    _run(d, o, 10, "step10")


def test_gpu_links_flap_64(hip, tmp_path):
    w = workloads.moving_boundary_flap(n=64, steps=1, substeps=120, out_dir=str(tmp_path))
    d, o = make_pair(w, tmp_path / "run", hip, None)
    _run(d, o, 1, "flap64")
