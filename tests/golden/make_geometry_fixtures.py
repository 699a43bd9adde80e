"""Pack the reference's committed duct-flap flag files into small fixtures.

Sources (data written by the reference FSIDriver3D, read as text):
  outputs/step104_driver_runs/fluent_duct_flap_setup_repair_48_5step_smoke/geo_duct_flap_proxy_48.dat
  outputs/step113_full_fsi_mirrored_duct_flap_96/geo_duct_flap_proxy_96.dat
with the geometry JSON each run used and its duct_static_geometry_report.json.
The .dat files hold one 0/1 flag per line in Fortran order; the fixture
stores the (n,n,n) int8 array bit-packed.
"""
import json
import shutil
from pathlib import Path

import numpy as np

REF = Path("/root/reference")
DST = Path(__file__).resolve().parent / "duct_geometry"
CASES = {
    48: ("outputs/step104_driver_runs/fluent_duct_flap_setup_repair_48_5step_smoke",
         "configs/step104_fluent_duct_flap_geometry_1024.json"),
    96: ("outputs/step113_full_fsi_mirrored_duct_flap_96",
         "outputs/step113_full_fsi_mirrored_duct_flap_96/mirrored_duct_flap_geometry_96_8192.json"),
}
DST.mkdir(parents=True, exist_ok=True)
for n, (run, geo) in CASES.items():
    flags = np.loadtxt(REF / run / f"geo_duct_flap_proxy_{n}.dat").astype(np.int8)
    flags = flags.reshape((n, n, n), order="F")
    np.savez_compressed(DST / f"flags_{n}.npz", packed=np.packbits(flags.reshape(-1)), shape=np.array(flags.shape))
    shutil.copyfile(REF / geo, DST / f"geometry_{n}.json")
    shutil.copyfile(REF / run / "duct_static_geometry_report.json", DST / f"report_{n}.json")
    print(n, int(flags.sum()), "solid cells")
