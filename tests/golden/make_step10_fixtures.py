"""Copy the reference's committed Step-10 mode-matrix outputs into fixtures.

Source (data files of the reference, read in place, never executed):
  outputs/step10_mode_matrix/{none,penalty,moving_boundary}/diagnostics_timeseries.csv
  outputs/step10_mode_matrix/{mode}/driver_config.json
  outputs/step10_mode_matrix/moving_boundary/particles_{x,J}.npy
They were produced by the reference FSIDriver3D on a GTX 1660 (Taichi CUDA):
32^3 periodic box, 4096 box particles, 20 LBM steps x 10 MPM substeps.
"""
import shutil
from pathlib import Path

SRC = Path("/root/reference/outputs/step10_mode_matrix")
DST = Path(__file__).resolve().parent / "step10_mode_matrix"

for mode in ("none", "penalty", "moving_boundary"):
    (DST / mode).mkdir(parents=True, exist_ok=True)
    for name in ("diagnostics_timeseries.csv", "driver_config.json"):
        shutil.copyfile(SRC / mode / name, DST / mode / name)
for name in ("particles_x.npy", "particles_J.npy"):
    shutil.copyfile(SRC / "moving_boundary" / name, DST / "moving_boundary" / name)
print("wrote", DST)
