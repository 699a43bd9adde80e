"""Precision form: FP16 / BF16 storage of F - w collided in FP32, and the
regularised collision, through the emulation build of the kernels.

Errors are stated against the FP64 oracle (max |a - b| / max |b| over fluid
cells after 3 FSI steps of 20 substeps, 16^3 flap case at tau = 0.55):
  FP16 storage: F <= 2e-4, u <= 2e-3;  BF16 storage: F <= 2e-3, u <= 2e-2.
The regularised collision in FP32 storage meets the parity bar of
tests/parity.py against the oracle's FP32 / FP64 restatements of it."""
import numpy as np
import pytest

from mpm_lbm import workloads
from mpm_lbm.sim.drivers.fsi_config import FSIDriverConfig
from tests.parity import check_history, make_pair, run_and_compare

BOUNDS = {"f16": {"F": 2e-4, "rho": 2e-4, "v": 2e-3}, "bf16": {"F": 2e-3, "rho": 2e-3, "v": 2e-2}}


def _workload(tmp_path, storage, collision):
    w = workloads.moving_boundary_flap(n=16, steps=3, substeps=20, flap=(2, 6, 6), mpm_substeps=3, tau=0.55,
                                       out_dir=str(tmp_path))
    cfg = FSIDriverConfig(**{**w.config.to_dict(), "lbm_population_storage": storage, "lbm_collision": collision})
    return workloads.Workload(w.name, cfg, w.cloud, w.geometry_json, w.material)


@pytest.mark.parametrize("collision", ["mrt", "regularized"])
@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_16bit_storage_error_vs_fp64(emu, tmp_path, storage, collision):
    d, o = make_pair(_workload(tmp_path, storage, collision), tmp_path / "run", emu, "cpu", dtype=np.float64)
    last = run_and_compare(d, o, 3)[-1]
    assert last["solid_equal"]
    for key, bound in BOUNDS[storage].items():
        assert last[key] <= bound, (storage, collision, key, last[key])


def test_regularized_fp32_meets_kernel_bar(emu, tmp_path):
    d, o = make_pair(_workload(tmp_path, "f32", "regularized"), tmp_path / "run", emu, "cpu", dtype=np.float32,
                     with_fp64=True)
    check_history(run_and_compare(d, o, 3), "regularized f32", 1e-5)


def test_16bit_round_trip_of_population_view(emu, tmp_path):
    """F.to_numpy() decodes w + d; from_numpy() re-encodes with round-to-nearest-even."""
    import torch

    for storage, half in (("f16", np.float16), ("bf16", None)):
        d, _ = make_pair(_workload(tmp_path, storage, "mrt"), tmp_path / f"rt_{storage}", emu, "cpu")
        F = d.lbm.F.to_numpy()
        w = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12, dtype=np.float32)
        assert np.allclose(F, w, atol=1e-7)  # rest state: deviations are exactly representable zeros
        rng = np.random.default_rng(0)
        G = (w + 1e-3 * rng.standard_normal(F.shape)).astype(np.float32)
        d.lbm.F.from_numpy(G)
        back = d.lbm.F.to_numpy()
        dev = torch.from_numpy(G - w)
        if half is not None:
            expect = dev.to(torch.float16).to(torch.float32).numpy() + w
        else:
            expect = dev.to(torch.bfloat16).to(torch.float32).numpy() + w
        assert np.array_equal(back, expect)
