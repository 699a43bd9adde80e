"""Per-cell bounce-back link classification (cellinfo) on the kernels' own
code (serial emulation build) against the oracle's, bit for bit, after every
FSI step: the Step-10 moving box (solid flags change every step), the
subcycled flap, and a penalty block (static flags).  tests/links.py states
the bits; the GPU runs the same checks in tests/test_gpu_link_classification.py."""
import numpy as np
import pytest

from mpm_lbm import workloads
from tests import links, step10
from tests.parity import make_pair


def _run(d, o, steps):
    for step in range(steps):
        d.step_once()
        o.step_once()
        sol = o.fields()["solid"]
        assert np.array_equal(d.lbm.solid.to_numpy(), sol), step
        bad = links.cellinfo_problems(d, sol)
        assert not bad, (step, bad)


def test_links_step10_moving_box(emu, tmp_path):
    from mpm_lbm.sim.drivers.fsi_driver import FSIDriver3D
    from oracle.oracle import OracleFSI

    cfg = step10.config("moving_boundary")
    cloud = step10.box_cloud(cfg)
    d = FSIDriver3D(cfg, str(tmp_path / "out"), particle_cloud=cloud, native=emu, device="cpu",
                    write_geometry_file=False)
    d.initialize()
    o = OracleFSI(cfg, d.lbm.solid.to_numpy(), cloud, step10.MATERIAL)
    _run(d, o, 6)
    n_links = int(np.unpackbits(links.oracle_link_masks(o.fields()["solid"]).view(np.uint8)).sum())
    assert n_links == o.mbstats.link_count  # the classification the streaming counted


@pytest.mark.parametrize("kind", ["flap_subcycled", "penalty"])
def test_links_workloads(emu, tmp_path, kind):
    if kind == "penalty":
        w = workloads.penalty_block(n=16, steps=2, block=(4, 4, 2), out_dir=str(tmp_path))
    else:
        w = workloads.moving_boundary_flap(n=16, steps=2, substeps=6, flap=(2, 6, 6), mpm_substeps=3,
                                           out_dir=str(tmp_path))
    d, o = make_pair(w, tmp_path / "run", emu, "cpu")
    _run(d, o, 2)


def test_link_masks_pin_on_a_hand_case():
    """One solid cell in a periodic 4^3 box: each of its 18 fluid neighbours
    has exactly the one link pointing at it."""
    solid = np.zeros((4, 4, 4), np.int8)
    solid[1, 2, 3] = 1
    m = links.oracle_link_masks(solid)
    from oracle import oracle as orc

    for s in range(1, 19):
        e = orc.E[s]
        x = ((1 - e[0]) % 4, (2 - e[1]) % 4, (3 - e[2]) % 4)
        assert m[x] == np.uint32(1 << s), s
    assert int(np.count_nonzero(m)) == 18 and m[1, 2, 3] == 0
