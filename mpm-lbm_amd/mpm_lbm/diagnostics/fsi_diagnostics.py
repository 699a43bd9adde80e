"""FSIDiagnostics3D (reference: diagnostics/fsi_diagnostics.py:6-128).

The reference copies whole fields to NumPy for every row; here the same
reductions come from one device pass (`LBMFluid3D.reduce_stats`,
`MPMSolid3D.reduce_stats`) and only scalars cross to the host.
"""
import numpy as np


class FSIDiagnostics3D:
    @staticmethod
    def lbm_fluid_stats(lbm, stats=None):
        r = stats if stats is not None else lbm.reduce_stats()
        n = r["fluid_count"]
        mean = (r["fluid_sum_vx"] / n, r["fluid_sum_vy"] / n, r["fluid_sum_vz"] / n) if n > 0 else (0.0, 0.0, 0.0)
        base = lbm.get_stats() if stats is None else {
            "max_v": r["max_v"], "rho_min": r["rho_min"], "rho_max": r["rho_max"], "mass_total": r["mass_total"],
            "force_norm_max": r["force_norm_max"]}
        return {**base, "fluid_cell_count": int(n), "fluid_mean_velocity": tuple(float(v) for v in mean)}

    @staticmethod
    def projection_zone_fluid_mean_velocity(lbm, phi_threshold=1.0e-6, stats=None):
        """phi-weighted mean velocity over cells with phi > 1e-6 (solid cells included, as in the reference)."""
        if phi_threshold != 1.0e-6:
            raise NotImplementedError("the device reduction uses phi_threshold = 1e-6")
        r = stats if stats is not None else lbm.reduce_stats()
        w = r["phi_weight_sum"]
        if w <= 0.0:
            return np.zeros(3, dtype=np.float64)
        return np.array([r["phi_weighted_vx"], r["phi_weighted_vy"], r["phi_weighted_vz"]]) / w

    @staticmethod
    def far_field_fluid_mean_velocity(lbm, phi_threshold=1.0e-6, stats=None):
        if phi_threshold != 1.0e-6:
            raise NotImplementedError("the device reduction uses phi_threshold = 1e-6")
        r = stats if stats is not None else lbm.reduce_stats()
        n = r["far_count"]
        if n <= 0:
            return np.zeros(3, dtype=np.float64)
        return np.array([r["far_sum_vx"], r["far_sum_vy"], r["far_sum_vz"]]) / n

    @staticmethod
    def force_stats(lbm, stats=None):
        r = stats if stats is not None else lbm.reduce_stats()
        net_cell = (r["net_cell_force_x"], r["net_cell_force_y"], r["net_cell_force_z"])
        net_hydro = (r["net_hydro_force_x"], r["net_hydro_force_y"], r["net_hydro_force_z"])
        balance = np.asarray(net_cell) + np.asarray(net_hydro)
        return {
            "max_cell_force_norm": r["max_cell_force_norm_all"],
            "max_hydro_force_norm": r["max_hydro_force_norm_all"],
            "net_cell_force": net_cell,
            "net_hydro_force": net_hydro,
            "force_balance_error": float(np.linalg.norm(balance)),
        }

    @staticmethod
    def solid_mean_velocity_norm(solid, stats=None):
        r = stats if stats is not None else solid.reduce_stats()
        if r["total_mass"] <= 0.0:
            return np.zeros(3, dtype=np.float64)
        return np.array([r["mv_x"], r["mv_y"], r["mv_z"]]) / r["total_mass"]
