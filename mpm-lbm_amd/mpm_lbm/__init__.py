"""MI355X-native FSI substep with the interface of lizhuoh9/MPM-LBM's FSIDriver3D.

Import paths mirror the reference package (``mpm_lbm.sim.drivers.fsi_driver``
and so on); the solver kernels run as hand-written gfx950 HIP code in
``_lib/libfsi_hip.so`` reached through ``mpm_lbm._native``.
"""
