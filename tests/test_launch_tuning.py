"""Host side of the substep launch autotuner: the candidate shapes per
storage, the driver's steady-substep flags, and that the tuner refuses a
CPU lattice (it times GPU launches)."""
import pytest

from mpm_lbm import _native, tuning, workloads
from mpm_lbm.sim.drivers.fsi_driver import FSIDriver3D


def test_candidate_shapes():
    f32 = tuning.candidates(_native.STORE_F32)
    assert len(f32) == 12 and len(set(f32)) == 12
    assert {b for b, _, _ in f32} == {128, 256, 512} and {w for _, w, _ in f32} == {1, 5}
    assert tuning.candidates(_native.STORE_F16_DEV) == [(256, 1, 0), (256, 1, 1), (256, 5, 0), (256, 5, 1)]


def test_driver_flags_and_cpu_refusal(emu, tmp_path):
    w = workloads.moving_boundary_flap(n=12, steps=1, substeps=4, out_dir=str(tmp_path))
    d = FSIDriver3D(w.config, str(tmp_path / "run"), particle_cloud=w.cloud, native=emu, device="cpu",
                    write_geometry_file=False, use_graph=False)
    d.initialize()
    assert tuning.driver_substep_flags(d) == _native.SUB_MOVING_WALL | _native.SUB_GATHER_PREV
    with pytest.raises(RuntimeError, match="GPU"):
        d.tune_launch()
    p = workloads.flap_case("penalty", 12, substeps=4, out_dir=str(tmp_path))
    dp = FSIDriver3D(p.config, str(tmp_path / "pen"), particle_cloud=p.cloud, native=emu, device="cpu",
                     write_geometry_file=False, use_graph=False)
    dp.initialize()
    assert tuning.driver_substep_flags(dp) == _native.SUB_STORED_MOMENTS
