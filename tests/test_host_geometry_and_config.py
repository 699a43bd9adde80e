"""Host-side layers: duct flags against the reference's own flag files,
lattice particle sampling, config validation messages, and the C ABI
library (loads, exports every declared symbol, struct layouts agree)."""
import json
import re
from pathlib import Path

import numpy as np
import pytest

from mpm_lbm.sim.drivers.fsi_config import FSIDriverConfig
from mpm_lbm.sim.geometry.config import GeometryConfig
from mpm_lbm.sim.geometry.duct_flap_proxy import duct_flap_proxy_component_masks, duct_flap_proxy_static_geometry
from mpm_lbm.sim.geometry.sampler import GeometrySampler3D

GOLDEN = Path(__file__).resolve().parent / "golden" / "duct_geometry"
ROOT = Path(__file__).resolve().parents[1]


@pytest.mark.parametrize("n", [48, 96])
def test_duct_flags_bit_exact_against_reference_files(n):
    z = np.load(GOLDEN / f"flags_{n}.npz")
    ref = np.unpackbits(z["packed"])[: n ** 3].reshape(n, n, n).astype(np.int8)
    gc = GeometryConfig.from_json(str(GOLDEN / f"geometry_{n}.json"))
    flags, report = duct_flap_proxy_static_geometry(n, gc)
    assert np.array_equal(flags, ref)
    ref_report = json.loads((GOLDEN / f"report_{n}.json").read_text())
    for key in ("duct_wall_cell_count", "fluid_cell_count", "inlet_fluid_cell_count",
                "pressure_outlet_fluid_cell_count", "solid_cell_count", "flap_static_cell_count"):
        assert report[key] == ref_report[key], key


def test_fortran_order_round_trip(tmp_path):
    """The driver writes flags in Fortran order and LBMFluid3D.init_geo reads them back (fluid.py:347-353)."""
    gc = GeometryConfig.from_json(str(GOLDEN / "geometry_48.json"))
    flags, _ = duct_flap_proxy_static_geometry(48, gc)
    path = tmp_path / "geo.dat"
    np.savetxt(path, flags.reshape(-1, order="F"), fmt="%d")
    back = np.loadtxt(path).reshape((48, 48, 48), order="F").astype(np.int8)
    assert np.array_equal(back, flags)


def test_lattice_sampling_masks_and_volume():
    gc = GeometryConfig.from_json(str(GOLDEN / "geometry_96.json"))
    cloud = GeometrySampler3D(gc).sample_particles()
    x = cloud["x"].astype(np.float64)
    assert x.shape == (8192, 3) and cloud["x"].dtype == np.float32
    masks = duct_flap_proxy_component_masks(x, gc)
    assert masks["flap"].all()  # every sample lies inside a flap
    assert np.array_equal(cloud["fixed_base_mask"].astype(bool), masks["fixed_base"])
    st = cloud["sampling_stats"]
    # inside candidates on the sampling lattice estimate the two flaps' volume 2 * 0.03 * 0.1 * 0.1
    assert st["geometry_volume"] == pytest.approx(2 * 0.03 * 0.1 * 0.1, rel=0.1)
    assert np.allclose(cloud["vol0"].sum(dtype=np.float64), st["geometry_volume"], rtol=1e-5)
    # evenly spaced subset: deterministic
    again = GeometrySampler3D(gc).sample_particles()
    assert np.array_equal(again["x"], cloud["x"])


@pytest.mark.parametrize("kwargs, message", [
    (dict(coupling_mode="bogus"), "coupling_mode must be one of"),
    (dict(lbm_open_boundary_semantics="zou_he_reconstruct_unknowns"), "is not implemented"),
    (dict(fsi_exchange_mode="lbm_subcycled_per_fsi_step", coupling_mode="penalty", lbm_substeps_per_fsi_step=4),
     "lbm_subcycled_per_fsi_step currently requires coupling_mode='moving_boundary'"),
    (dict(fsi_exchange_mode="lbm_subcycled_per_fsi_step", coupling_mode="moving_boundary", lbm_substeps_per_fsi_step=4),
     "subcycled FSI exchange requires"),
    (dict(lbm_substeps_per_fsi_step=2), "lbm_substeps_per_fsi_step must be 1"),
    (dict(lbm_tau_stability_policy="strict", lbm_viscosity_semantics="physical_nu_mapping", lbm_min_tau_margin=0.01),
     "LBM tau stability margin failed strict policy"),
    (dict(reaction_transfer_mode="interface_traction_conservative"), "planned but not implemented"),
    (dict(n_grid=0), "n_grid must be positive"),
])
def test_config_validation_messages(kwargs, message):
    with pytest.raises(ValueError, match=re.escape(message)):
        FSIDriverConfig(**kwargs)


def test_legacy_tau_and_physical_mapping():
    legacy = FSIDriverConfig().lbm_viscosity_mapping_report()
    assert legacy["tau"] == pytest.approx(0.1 / 3 + 0.5)
    c = FSIDriverConfig(lbm_viscosity_semantics="physical_nu_mapping", n_grid=32, mpm_substeps_per_lbm_step=1,
                        mpm_dt=1e-4, fluid_kinematic_viscosity_m2_s=(0.1 / 3.0) * (1 / 32) ** 2 / 1e-4)
    assert c.lbm_viscosity_mapping_report()["tau"] == pytest.approx(0.6, rel=1e-12)
    assert FSIDriverConfig.from_json  # JSON round trip
    d = c.to_dict()
    assert FSIDriverConfig(**d) == c


def test_driver_constructor_is_lazy(tmp_path):
    from mpm_lbm.sim.drivers.fsi_driver import TIMING_KEYS, FSIDriver3D

    out = tmp_path / "never_created"
    d = FSIDriver3D(FSIDriverConfig(), str(out))
    assert not out.exists() and not d.initialized and d.lbm is None
    assert tuple(d.timing.keys()) == TIMING_KEYS


def _declared_symbols():
    text = (ROOT / "include" / "fsi_hip.h").read_text()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(fsi_\w+)\(", text, flags=re.M)))


def test_hip_library_exports_every_declared_symbol():
    import ctypes

    from mpm_lbm import _native

    declared = _declared_symbols()
    assert set(declared) == set(_native.EXPORTED_SYMBOLS)
    for path in (_native.HIP_LIB_PATH, _native.HIP_EXACT_LIB_PATH):
        assert path.is_file(), f"{path.name} not built"
        lib = ctypes.CDLL(str(path))
        for name in declared:
            assert hasattr(lib, name), (path.name, name)
        # struct layouts: the library reports its sizeofs without touching a GPU
        _native.Native(path, "hip").check_abi()


def test_product_path_refuses_missing_library(monkeypatch, tmp_path):
    from mpm_lbm import _native

    monkeypatch.setattr(_native, "HIP_LIB_PATH", tmp_path / "missing.so")
    monkeypatch.setattr(_native, "HIP_EXACT_LIB_PATH", tmp_path / "missing_exact.so")
    monkeypatch.setattr(_native, "_HIP", {})
    with pytest.raises(_native.FsiError, match="not built"):
        _native.load_hip()
    with pytest.raises(_native.FsiError, match="not built"):
        _native.load_hip(exact=True)
