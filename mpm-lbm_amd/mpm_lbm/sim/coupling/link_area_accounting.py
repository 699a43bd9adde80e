"""Link-area proxy accounting of the moving-wall bounce-back links
(reference: sim/coupling/link_area_accounting.py:1-166).  Diagnostic only:
host reductions over the per-direction link statistics of the last
moving-wall substep, which LBMFluid3D reduces on the device
(get_moving_boundary_directional_stats)."""
from __future__ import annotations

import os

import numpy as np

VALID_AREA_POLICIES = ("uniform", "inverse_length", "length")


def _velocity_set(lbm):
    e = lbm.e.to_numpy() if hasattr(lbm.e, "to_numpy") else lbm.e
    return np.asarray(e, dtype=np.float64)


class LinkAreaMomentumAccounting3D:
    """Per-direction link tables, their area weighting and the summary
    record of link_area_accounting.py:86-145."""

    @staticmethod
    def direction_metadata(lbm):
        """Length and class (rest / axis / face diagonal) of each direction."""
        e = _velocity_set(lbm)
        manhattan = np.abs(e).sum(axis=1)
        if np.any(manhattan > 2):
            raise RuntimeError("unexpected D3Q19 direction with more than two nonzero components")
        cls = np.array(["rest", "axis", "face_diagonal"], dtype=object)[manhattan.astype(np.int64)]
        return {
            "e": e,
            "length": np.linalg.norm(e, axis=1),
            "direction_class": cls,
            "is_rest": cls == "rest",
            "is_axis": cls == "axis",
            "is_face_diagonal": cls == "face_diagonal",
        }

    @staticmethod
    def area_weights(lbm, policy="inverse_length"):
        """Weight per direction: 1, 1/|e| or |e| for the moving directions, 0 at rest."""
        if policy not in VALID_AREA_POLICIES:
            raise ValueError(f"policy must be one of {VALID_AREA_POLICIES}")
        meta = LinkAreaMomentumAccounting3D.direction_metadata(lbm)
        length = meta["length"]
        moving = ~meta["is_rest"]
        w = np.zeros_like(length)
        w[moving] = {"uniform": np.ones_like(length), "inverse_length": 1.0 / np.where(moving, length, 1.0),
                     "length": length}[policy][moving]
        return w

    @staticmethod
    def read_directional_stats(lbm):
        stats = lbm.get_moving_boundary_directional_stats()
        return {
            "link_count_by_dir": np.asarray(stats["link_count_by_dir"], dtype=np.int64),
            "fluid_impulse_by_dir": np.asarray(stats["fluid_impulse_by_dir"], dtype=np.float64),
            "solid_force_by_dir": np.asarray(stats["solid_force_by_dir"], dtype=np.float64),
            "correction_abs_sum_by_dir": np.asarray(stats["correction_abs_sum_by_dir"], dtype=np.float64),
            "correction_abs_max_by_dir": np.asarray(stats["correction_abs_max_by_dir"], dtype=np.float64),
        }

    @staticmethod
    def _weighted(directional, weights, policy):
        wf = np.sum(directional["fluid_impulse_by_dir"] * weights[:, None], axis=0, dtype=np.float64)
        ws = np.sum(directional["solid_force_by_dir"] * weights[:, None], axis=0, dtype=np.float64)
        return {"policy": policy, "area_weighted_fluid_impulse": wf, "area_weighted_solid_force": ws,
                "area_weighted_balance": wf + ws, "area_weights": weights}

    @staticmethod
    def area_weighted_impulse(lbm, policy="inverse_length"):
        directional = LinkAreaMomentumAccounting3D.read_directional_stats(lbm)
        weights = LinkAreaMomentumAccounting3D.area_weights(lbm, policy=policy)
        return LinkAreaMomentumAccounting3D._weighted(directional, weights, policy)

    @staticmethod
    def summarize_link_accounting(lbm, policy="inverse_length"):
        meta = LinkAreaMomentumAccounting3D.direction_metadata(lbm)
        directional = LinkAreaMomentumAccounting3D.read_directional_stats(lbm)
        weights = LinkAreaMomentumAccounting3D.area_weights(lbm, policy=policy)
        weighted = LinkAreaMomentumAccounting3D._weighted(directional, weights, policy)
        scalar = lbm.get_moving_boundary_stats()
        count = directional["link_count_by_dir"]
        dir_fluid = np.sum(directional["fluid_impulse_by_dir"], axis=0, dtype=np.float64)
        dir_solid = np.sum(directional["solid_force_by_dir"], axis=0, dtype=np.float64)
        sc_fluid = np.asarray(scalar["bb_net_fluid_impulse"], dtype=np.float64)
        sc_solid = np.asarray(scalar["bb_net_solid_force"], dtype=np.float64)
        hydro_sum = np.sum(lbm.hydro_force.to_numpy().reshape(-1, 3), axis=0, dtype=np.float64)
        err_fluid, err_solid, err_hydro = sc_fluid - dir_fluid, sc_solid - dir_solid, hydro_sum - dir_solid
        out = {
            "policy": policy,
            "total_link_count": int(count.sum()),
            "axis_link_count": int(count[meta["is_axis"]].sum()),
            "face_diagonal_link_count": int(count[meta["is_face_diagonal"]].sum()),
            "area_proxy_total": float(np.sum(count * weights)),
        }
        for name, vec in (("bb_net_fluid_impulse", sc_fluid), ("bb_net_solid_force", sc_solid),
                          ("directional_fluid_impulse", dir_fluid), ("directional_solid_force", dir_solid),
                          ("area_weighted_fluid_impulse", weighted["area_weighted_fluid_impulse"]),
                          ("area_weighted_solid_force", weighted["area_weighted_solid_force"]),
                          ("area_weighted_balance_error", weighted["area_weighted_balance"])):
            for axis, v in zip("xyz", vec):
                out[f"{name}_{axis}"] = float(v)
        for axis, v in zip("xyz", err_fluid):
            out[f"scalar_vs_directional_impulse_error_{axis}"] = float(abs(v))
        out["scalar_vs_directional_impulse_error_norm"] = float(np.linalg.norm(err_fluid))
        out["scalar_vs_directional_solid_error_x"] = float(abs(err_solid[0]))
        out["scalar_vs_directional_solid_error_norm"] = float(np.linalg.norm(err_solid))
        for axis, v in zip("xyz", hydro_sum):
            out[f"hydro_force_sum_{axis}"] = float(v)
        out["hydro_vs_directional_solid_error_x"] = float(abs(err_hydro[0]))
        out["hydro_vs_directional_solid_error_norm"] = float(np.linalg.norm(err_hydro))
        out["correction_abs_sum"] = float(np.sum(directional["correction_abs_sum_by_dir"]))
        out["correction_abs_max"] = float(np.max(directional["correction_abs_max_by_dir"]))
        return out

    @staticmethod
    def save_directional_npz(lbm, path):
        directional = LinkAreaMomentumAccounting3D.read_directional_stats(lbm)
        meta = LinkAreaMomentumAccounting3D.direction_metadata(lbm)
        os.makedirs(os.path.dirname(path), exist_ok=True)
        np.savez(path, e=meta["e"], length=meta["length"], direction_class=meta["direction_class"].astype(str),
                 **directional)
