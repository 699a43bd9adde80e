"""ctypes binding of libfsi_hip.so (C ABI declared in include/fsi_hip.h).

The structs below mirror the header field for field; `Native.check_abi()`
compares their sizes with the library's own `fsi_abi_sizes()` so a layout
drift fails loudly instead of corrupting memory.

Only `load_hip()` is used by the product path.  It raises if the gfx950
library has not been built: there is no CPU fallback.  `load_hip(exact=True)`
loads libfsi_hip_exact.so, the same kernels built with -ffp-contract=off
(every float operation rounded on its own, as the serial CPU build and the
oracle evaluate them), which the particle-ordered transfer mode uses.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

_HERE = Path(__file__).resolve().parent
LIB_DIR = _HERE / "_lib"
HIP_LIB_PATH = LIB_DIR / "libfsi_hip.so"
HIP_EXACT_LIB_PATH = LIB_DIR / "libfsi_hip_exact.so"

f32 = ctypes.c_float
i32 = ctypes.c_int32
u32 = ctypes.c_uint32
f64 = ctypes.c_double
vp = ctypes.c_void_p

FSI_BC_PERIODIC = 0
FSI_BC_PRESSURE = 1
FSI_BC_VELOCITY = 2

STORE_F32 = 0
STORE_F16_DEV = 1
STORE_BF16_DEV = 2
COLLIDE_MRT = 0
COLLIDE_REGULARIZED = 1
OPEN_BC_EQUILIBRIUM = 0
OPEN_BC_REGULARIZED = 1
OPEN_BC_REGULARIZED_LIMITED = 2
OPEN_BC_CONVECTIVE = 3

SUB_STORED_MOMENTS = 1
SUB_MOVING_WALL = 2
SUB_GATHER_PREV = 4
SUB_DIAG = 8
SUB_CELL_FORCE = 16

E_INVALID = -1  # FSI_E_INVALID
E_UNSUPPORTED = -3  # FSI_E_UNSUPPORTED

REDUCE_BLOCKS = 256


class LbmParams(ctypes.Structure):
    _fields_ = [
        ("nx", i32), ("ny", i32), ("nz", i32),
        ("bc_x_left", i32), ("bc_x_right", i32),
        ("s_v", f32), ("s_other", f32), ("rho0", f32),
        ("force", f32 * 3),
        ("rho_bc_x_left", f32), ("rho_bc_x_right", f32),
        ("vel_bc_x_left", f32 * 3), ("vel_bc_x_right", f32 * 3),
        ("dynamic_solid_threshold", f32),
        ("storage", i32), ("collision", i32),
        ("open_bc", i32), ("ob_limiter", i32), ("ob_floor_enabled", i32),
        ("ob_rho_min", f32), ("ob_rho_max", f32), ("ob_u_max", f32), ("ob_noneq_cap", f32), ("ob_floor", f32),
        ("launch_block", i32), ("launch_minw", i32), ("launch_linear", i32), ("launch_pad", i32),
    ]


class LbmCounters(ctypes.Structure):
    _fields_ = [
        ("bb_link_count", i32),
        ("bb_max_correction_bits", u32),
        ("mb_sample_count", i32),
        ("accum_norm_max_bits", u32),
        ("mean_norm_max_bits", u32),
        ("pad0", i32),
        ("bb_net_solid_force", f64 * 3),
        ("ob_step", u32 * 5), ("pad1", u32),
        ("ob_run", ctypes.c_uint64 * 5),
    ]


class LbmState(ctypes.Structure):
    _fields_ = [
        ("F", vp * 2),
        ("rho", vp), ("vel", vp),
        ("solid", vp), ("static_solid", vp), ("old_solid", vp),
        ("cellinfo", vp),
        ("solid_phi", vp), ("solid_mass", vp), ("solid_vel", vp),
        ("cell_force", vp), ("hydro_force", vp),
        ("link_c", vp * 2), ("link_corr", vp),
        ("hydro_accum", vp), ("reduce_ws", vp),
        ("counters", vp),
    ]


LBM_STATS_FIELDS = [
    "fluid_count", "rho_min", "rho_max", "mass_total", "max_v", "force_norm_max",
    "fluid_sum_vx", "fluid_sum_vy", "fluid_sum_vz",
    "phi_weight_sum", "phi_weighted_vx", "phi_weighted_vy", "phi_weighted_vz",
    "far_count", "far_sum_vx", "far_sum_vy", "far_sum_vz",
    "max_cell_force_norm_all", "max_hydro_force_norm_all",
    "net_cell_force_x", "net_cell_force_y", "net_cell_force_z",
    "net_hydro_force_x", "net_hydro_force_y", "net_hydro_force_z",
    "mass_rho_all",
]


class LbmStats(ctypes.Structure):
    _fields_ = [(name, f64) for name in LBM_STATS_FIELDS]


class MpmParams(ctypes.Structure):
    _fields_ = [
        ("gnx", i32), ("gny", i32), ("gnz", i32),
        ("n_particles", i32),
        ("bound", i32),
        ("damping_enabled", i32), ("lock_z_enabled", i32),
        ("dx", f32), ("inv_dx", f32), ("dt", f32),
        ("mu", f32), ("la", f32),
        ("gravity", f32 * 3),
        ("velocity_damping", f32),
        ("domain_max", f32 * 3),
    ]


class MpmState(ctypes.Structure):
    _fields_ = [
        ("x", vp), ("v", vp), ("C", vp), ("F", vp), ("Jp", vp),
        ("mass", vp), ("vol0", vp), ("fixed_mask", vp), ("fixed_x", vp),
        ("grid_v", vp), ("grid_m", vp), ("grid_f", vp),
    ]


class CouplingParams(ctypes.Structure):
    _fields_ = [
        ("inv_dx_norm", f32), ("cell_volume_norm", f32), ("vel_scale_norm_to_lbm", f32),
        ("eps_mass", f32),
        ("beta_lbm", f32), ("penalty_phi_min", f32), ("force_cap_lbm", f32),
        ("force_density_scale", f32), ("reaction_scale", f32),
        ("mb_phi_min", f32), ("mb_force_cap_norm", f32),
        ("area_scale", f32),
    ]

    def __init__(self, *args, **kw):
        super().__init__(*args, **kw)
        if "area_scale" not in kw:
            self.area_scale = 1.0  # the engineering transfer's reaction is unscaled


class CouplingCounters(ctypes.Structure):
    _fields_ = [
        ("projected_mass", f64), ("projected_volume_raw", f64), ("projected_volume_clamped", f64),
        ("max_phi_raw_bits", u32),
        ("active_cell_count", i32),
        ("active_reaction_particle_count", i32),
        ("active_force_cell_count", i32),
        ("max_particle_reaction_bits", u32),
        ("max_grid_reaction_bits", u32),
        ("max_cell_force_bits", u32), ("max_hydro_force_bits", u32),
        ("net_particle_reaction", f64 * 3),
        ("net_grid_reaction", f64 * 3),
        ("net_cell_force", f64 * 3), ("net_hydro_force", f64 * 3),
    ]


MPM_STATS_FIELDS = [
    "min_x0", "min_x1", "min_x2", "max_x0", "max_x1", "max_x2",
    "max_speed", "min_J", "max_J", "total_mass", "mv_x", "mv_y", "mv_z",
]


class MpmStats(ctypes.Structure):
    _fields_ = [(name, f64) for name in MPM_STATS_FIELDS]


BATCH_MODE_NONE = 0
BATCH_MODE_PENALTY = 1
BATCH_MODE_MOVING_BOUNDARY = 2
BATCH_DOMAIN_CELLS = 0
BATCH_DOMAIN_NODES = 1
BATCH_DOMAIN_PARTICLES = 2
BATCH_DOMAIN_PLANES = 3
BATCH_DOMAIN_INSTANCES = 4
BATCH_DOMAIN_BOX = 5
BATCH_BOX_BLOCKS = 64  # FSI_BATCH_BOX_BLOCKS
(BATCH_OP_LBM_SUBSTEP, BATCH_OP_LBM_REBUILD, BATCH_OP_LBM_MOMENTS, BATCH_OP_LBM_MB_BEGIN, BATCH_OP_LBM_MB_FINALIZE,
 BATCH_OP_PROJ_CLEAR, BATCH_OP_PROJ_PARTICLES, BATCH_OP_PROJ_NORMALIZE, BATCH_OP_PENALTY, BATCH_OP_MPM_CLEAR,
 BATCH_OP_MPM_P2G, BATCH_OP_MPM_REACTION, BATCH_OP_MPM_GRID_UPDATE, BATCH_OP_MPM_G2P,
 BATCH_OP_COUNTERS_CLEAR, BATCH_OP_MPM_P2G_REACT, BATCH_OP_MPM_GRID_UPDATE_REACT, BATCH_OP_LBM_SUBSTEP_MB,
 BATCH_OP_LBM_SUBSTEP_PENALTY, BATCH_OP_LBM_SUBSTEP_NONE, BATCH_OP_LBM_MB_FINALIZE_FOLD,
 BATCH_OP_BOX_ROLL, BATCH_OP_LBM_RESIDENT) = range(23)
BATCH_BOX_INTS = 32  # FSI_BATCH_BOX_INTS

# domain of each batched operation (the block map a launch takes is built for it)
BATCH_OP_DOMAIN = {
    BATCH_OP_LBM_SUBSTEP: BATCH_DOMAIN_PLANES, BATCH_OP_LBM_SUBSTEP_MB: BATCH_DOMAIN_PLANES,
    BATCH_OP_LBM_SUBSTEP_PENALTY: BATCH_DOMAIN_PLANES, BATCH_OP_LBM_SUBSTEP_NONE: BATCH_DOMAIN_PLANES,
    BATCH_OP_LBM_REBUILD: BATCH_DOMAIN_CELLS, BATCH_OP_LBM_MOMENTS: BATCH_DOMAIN_CELLS,
    BATCH_OP_LBM_MB_BEGIN: BATCH_DOMAIN_CELLS, BATCH_OP_LBM_MB_FINALIZE: BATCH_DOMAIN_CELLS,
    BATCH_OP_PROJ_CLEAR: BATCH_DOMAIN_BOX, BATCH_OP_PROJ_NORMALIZE: BATCH_DOMAIN_BOX,
    BATCH_OP_PENALTY: BATCH_DOMAIN_BOX,
    BATCH_OP_MPM_CLEAR: BATCH_DOMAIN_BOX, BATCH_OP_MPM_GRID_UPDATE: BATCH_DOMAIN_BOX,
    BATCH_OP_MPM_GRID_UPDATE_REACT: BATCH_DOMAIN_BOX,
    BATCH_OP_PROJ_PARTICLES: BATCH_DOMAIN_PARTICLES, BATCH_OP_MPM_P2G: BATCH_DOMAIN_PARTICLES,
    BATCH_OP_MPM_REACTION: BATCH_DOMAIN_PARTICLES, BATCH_OP_MPM_P2G_REACT: BATCH_DOMAIN_PARTICLES,
    BATCH_OP_MPM_G2P: BATCH_DOMAIN_PARTICLES,
    BATCH_OP_LBM_MB_FINALIZE_FOLD: BATCH_DOMAIN_INSTANCES, BATCH_OP_BOX_ROLL: BATCH_DOMAIN_INSTANCES,
    BATCH_OP_COUNTERS_CLEAR: BATCH_DOMAIN_INSTANCES,
}


class BatchInstance(ctypes.Structure):
    _fields_ = [
        ("lp", LbmParams), ("ls", LbmState), ("mp", MpmParams), ("ms", MpmState),
        ("cp", CouplingParams), ("cc", vp),
        ("mode", i32), ("cur0", i32), ("n_substeps", i32), ("mpm_substeps", i32), ("active", i32),
        ("cell_force", i32), ("pad0", i32), ("react_ws", vp), ("boxes", vp),
        ("res_halo", vp), ("res_sync", vp), ("resident", i32), ("pad1", i32),
    ]


class Bins(ctypes.Structure):
    """FsiBins: a particle binning (key, start, order) and its work ints."""
    _fields_ = [("key", vp), ("start", vp), ("order", vp), ("work", vp)]


class BatchOp(ctypes.Structure):
    """FsiBatchOp: one operation of a batched exchange graph."""
    _fields_ = [("op", i32), ("t", i32), ("block_map", vp), ("n_blocks", i32), ("branch", i32)]


# FsiBatchOp.branch: graph chains of a batched exchange (side chains 1..BATCH_SIDE_CHAINS)
BATCH_BRANCH_MAIN, BATCH_SIDE_CHAINS, BATCH_BRANCH_JOIN = 0, 3, -1


ABI_STRUCTS = [LbmParams, LbmState, LbmCounters, LbmStats, MpmParams, MpmState, CouplingParams,
               CouplingCounters, MpmStats, BatchInstance, Bins, BatchOp]

EXPORTED_SYMBOLS = [
    "fsi_last_error", "fsi_version", "fsi_abi_sizes", "fsi_sync",
    "fsi_lbm_init", "fsi_lbm_rebuild", "fsi_lbm_substep", "fsi_lbm_open_bc", "fsi_lbm_link_stats", "fsi_lbm_moments",
    "fsi_lbm_mb_begin", "fsi_lbm_mb_finalize", "fsi_lbm_reduce_stats", "fsi_lbm_subcycle_resident",
    "fsi_project", "fsi_penalty_force", "fsi_mpm_clear_grid", "fsi_mpm_p2g", "fsi_mpm_reaction",
    "fsi_mpm_grid_update", "fsi_mpm_g2p", "fsi_mpm_substep_coupled", "fsi_mpm_substep_coupled_noalloc",
    "fsi_mpm_subcycle_coupled", "fsi_mpm_reduce_stats",
    "fsi_bin_particles", "fsi_mpm_substep_ordered", "fsi_project_ordered",
    "fsi_batch_block_map", "fsi_batch_launch",
    "fsi_batch_graph_create", "fsi_batch_graph_update", "fsi_batch_graph_launch", "fsi_batch_graph_destroy",
]

PARTICLE_BLOCK = 64  # FSI_PARTICLE_BLOCK
NOALLOC_MAX_PARTICLES = 1 << 21  # FSI_NOALLOC_MAX_PARTICLES
LINK_STATS_VALUES = 6  # FSI_LINK_STATS_VALUES
LINK_STATS_WS_DOUBLES = 18 * 64 * LINK_STATS_VALUES  # FSI_LINK_STATS_WS_DOUBLES


ORDERED_TERMS_FLOATS = 28  # FSI_ORDERED_TERMS_FLOATS


def bin_count(nx: int, ny: int, nz: int) -> int:
    """FSI_BIN_COUNT(nx, ny, nz): bins without the overflow bin."""
    return (int(nx) + 3) * (int(ny) + 3) * (int(nz) + 3)


def bin_work_ints(n_bins: int) -> int:
    """FSI_BIN_WS_INTS(n_bins)."""
    return int(n_bins) + 1 + (int(n_bins) + 1 + 1023) // 1024


def res_blocks(nx: int, ny: int) -> int:
    """FSI_RES_BLOCKS(nx, ny): workgroups of the resident subcycle."""
    return (int(nx) // 4) * (int(ny) // 4)


def res_sync_words(nx: int, ny: int) -> int:
    """FSI_RES_SYNC_WORDS(nx, ny)."""
    return 64 * (res_blocks(nx, ny) + 1)


def res_halo_floats(n_cells: int) -> int:
    """FSI_RES_HALO_FLOATS(n_cells)."""
    return 2 * 18 * int(n_cells)


MPM_RES_SYNC_WORDS = 640  # FSI_MPM_RES_SYNC_WORDS
MPM_RES_TIMEOUT_WORD = 576  # FSI_MPM_RES_TIMEOUT_WORD


def mpm_res_scratch_floats(grid_nodes: int) -> int:
    """FSI_MPM_RES_SCRATCH_FLOATS(grid_nodes)."""
    return 12 * int(grid_nodes)


def coupled_workspace_doubles(n_particles: int) -> int:
    """FSI_COUPLED_WORKSPACE_DOUBLES(n_particles)."""
    return 9 * ((int(n_particles) + PARTICLE_BLOCK - 1) // PARTICLE_BLOCK)

P = ctypes.POINTER


class FsiError(RuntimeError):
    """A libfsi_hip entry point returned a negative status."""


class Native:
    """Typed wrapper over one loaded copy of the C ABI."""

    def __init__(self, path: os.PathLike | str, kind: str):
        self.path = str(path)
        self.kind = kind  # "hip" or "emulation"
        self.lib = ctypes.CDLL(self.path)
        L = self.lib
        L.fsi_last_error.restype = ctypes.c_char_p
        L.fsi_last_error.argtypes = []
        L.fsi_version.restype = ctypes.c_int
        L.fsi_abi_sizes.restype = ctypes.c_int
        L.fsi_abi_sizes.argtypes = [P(i32), ctypes.c_int]
        L.fsi_sync.argtypes = [vp]
        L.fsi_lbm_init.argtypes = [P(LbmParams), P(LbmState), vp]
        L.fsi_lbm_rebuild.argtypes = [P(LbmParams), P(LbmState), ctypes.c_int, ctypes.c_int, vp]
        L.fsi_lbm_substep.argtypes = [P(LbmParams), P(LbmState), ctypes.c_int, ctypes.c_int, u32, vp]
        L.fsi_lbm_open_bc.argtypes = [P(LbmParams), P(LbmState), ctypes.c_int, u32, vp]
        L.fsi_lbm_link_stats.argtypes = [P(LbmParams), P(LbmState), ctypes.c_int, ctypes.c_int, vp, vp, vp]
        L.fsi_lbm_moments.argtypes = [P(LbmParams), P(LbmState), ctypes.c_int, u32, vp]
        L.fsi_lbm_mb_begin.argtypes = [P(LbmParams), P(LbmState), vp]
        L.fsi_lbm_mb_finalize.argtypes = [P(LbmParams), P(LbmState), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, vp]
        L.fsi_lbm_reduce_stats.argtypes = [P(LbmParams), P(LbmState), vp, vp, vp]
        L.fsi_lbm_subcycle_resident.argtypes = [P(LbmParams), P(LbmState), ctypes.c_int, ctypes.c_int, u32, vp, vp,
                                                vp]
        L.fsi_project.argtypes = [P(LbmParams), P(LbmState), P(MpmParams), P(MpmState), P(CouplingParams), vp, vp]
        L.fsi_penalty_force.argtypes = [P(LbmParams), P(LbmState), P(CouplingParams), vp, vp]
        L.fsi_mpm_clear_grid.argtypes = [P(MpmParams), P(MpmState), vp]
        L.fsi_mpm_p2g.argtypes = [P(MpmParams), P(MpmState), vp]
        L.fsi_mpm_reaction.argtypes = [P(LbmParams), P(LbmState), P(MpmParams), P(MpmState), P(CouplingParams),
                                       vp, ctypes.c_int, vp]
        L.fsi_mpm_grid_update.argtypes = [P(MpmParams), P(MpmState), vp]
        L.fsi_mpm_g2p.argtypes = [P(MpmParams), P(MpmState), vp]
        L.fsi_mpm_substep_coupled.argtypes = [P(LbmParams), P(LbmState), P(MpmParams), P(MpmState),
                                              P(CouplingParams), vp, ctypes.c_int, vp, vp]
        L.fsi_mpm_substep_coupled_noalloc.argtypes = [P(LbmParams), P(LbmState), P(MpmParams), P(MpmState),
                                                      P(CouplingParams), vp, ctypes.c_int, vp]
        L.fsi_mpm_subcycle_coupled.argtypes = [P(LbmParams), P(LbmState), P(MpmParams), P(MpmState),
                                               P(CouplingParams), vp, ctypes.c_int, ctypes.c_int, vp, vp, vp, vp]
        L.fsi_mpm_reduce_stats.argtypes = [P(MpmParams), P(MpmState), vp, vp, vp]
        L.fsi_bin_particles.argtypes = [vp, i32, ctypes.c_float, i32, i32, i32, P(Bins), vp]
        L.fsi_mpm_substep_ordered.argtypes = [P(LbmParams), P(LbmState), P(MpmParams), P(MpmState),
                                              P(CouplingParams), vp, ctypes.c_int, P(Bins), vp, vp, vp]
        L.fsi_project_ordered.argtypes = [P(LbmParams), P(LbmState), P(MpmParams), P(MpmState), P(CouplingParams),
                                          vp, P(Bins), vp]
        L.fsi_batch_block_map.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp]
        L.fsi_batch_graph_create.argtypes = [P(BatchOp), ctypes.c_int, vp, ctypes.c_int, P(vp)]
        L.fsi_batch_graph_update.argtypes = [vp, P(BatchOp), ctypes.c_int]
        L.fsi_batch_graph_launch.argtypes = [vp, vp]
        L.fsi_batch_graph_destroy.argtypes = [vp]
        L.fsi_batch_launch.argtypes = [ctypes.c_int, vp, ctypes.c_int, vp, ctypes.c_int, ctypes.c_int, vp]
        for name in EXPORTED_SYMBOLS:
            getattr(L, name).restype = ctypes.c_int if name != "fsi_last_error" else ctypes.c_char_p
        self.check_abi()

    def check_abi(self):
        out = (i32 * len(ABI_STRUCTS))()
        n = self.lib.fsi_abi_sizes(out, len(ABI_STRUCTS))
        if n != len(ABI_STRUCTS):
            raise FsiError(f"{self.path}: fsi_abi_sizes returned {n}")
        for cls, size in zip(ABI_STRUCTS, out):
            if ctypes.sizeof(cls) != size:
                raise FsiError(
                    f"{self.path}: struct {cls.__name__} is {size} bytes in C but {ctypes.sizeof(cls)} in ctypes"
                )

    def try_call(self, name: str, *args) -> bool:
        """Call an entry point that may decline its arguments: False when it
        returns FSI_E_UNSUPPORTED, True on success; raise on other errors."""
        rc = getattr(self.lib, name)(*args)
        if rc == E_UNSUPPORTED:
            return False
        if rc != 0:
            msg = self.lib.fsi_last_error().decode("utf-8", "replace")
            raise FsiError(f"{name} failed ({rc}): {msg}")
        return True

    def call(self, name: str, *args):
        rc = getattr(self.lib, name)(*args)
        if rc < 0 or (rc != 0 and name != "fsi_batch_block_map"):
            msg = self.lib.fsi_last_error().decode("utf-8", "replace")
            raise FsiError(f"{name} failed ({rc}): {msg}")
        return rc


_HIP: dict = {}


def load_hip(exact: bool = False) -> Native:
    """Load the gfx950 library built in-tree (exact=True: its build without
    fma fusion); raise if it is missing."""
    path = HIP_EXACT_LIB_PATH if exact else HIP_LIB_PATH
    if path not in _HIP:
        # torch first: its bundled libamdhip64.so.7 must be the one this
        # library binds to, so both share one HIP runtime and its streams.
        import torch  # noqa: F401

        if not path.is_file():
            raise FsiError(
                f"{path} is not built: run __graft_entry__.build() (hipcc --offload-arch=gfx950) first"
            )
        _HIP[path] = Native(path, "hip")
    return _HIP[path]


def ptr(t) -> int:
    """Device (or host) address of a contiguous torch tensor, or 0 for None."""
    if t is None:
        return 0
    if not t.is_contiguous():
        raise ValueError("buffers passed to libfsi_hip must be contiguous")
    return int(t.data_ptr())
