/*
 * fsi_hip.h - C ABI of libfsi_hip.so, the gfx950 implementation of the FSI
 * substep that FSIDriver3D runs (lizhuoh9/MPM-LBM).
 *
 * Conventions
 *  - Every entry point returns 0 on success and a negative FSI_E_* code on
 *    failure; fsi_last_error() returns a thread-local message for the last
 *    failure on the calling thread.
 *  - Every buffer is caller-owned device memory (the Python host allocates it
 *    as torch tensors).  The library never allocates or frees caller memory.
 *  - Calls are stream-ordered and asynchronous on the hipStream_t passed as
 *    `stream` (NULL = default stream).  Nothing synchronises except
 *    fsi_sync().
 *  - Lattice cell (i,j,k) has linear index c = (i*ny + j)*nz + k, the same
 *    row-major order as the reference's Taichi fields.  Populations are SoA:
 *    F[s*N + c].  Vector fields are SoA: v[d*N + c].
 *  - Particle arrays are SoA: x[d*P + p], C and F as 9 planes [(r*3+col)*P + p].
 *
 * The reference has no FFI: its operator interface is the duck-typed Python
 * methods of LBMFluid3D / MPMSolid3D / the couplers.  Each entry point below
 * names the reference method(s) it replaces.
 */
#ifndef FSI_HIP_H
#define FSI_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FSI_OK 0
#define FSI_E_INVALID (-1)  /* bad argument or shape */
#define FSI_E_HIP (-2)      /* HIP runtime error */
#define FSI_E_UNSUPPORTED (-3) /* valid arguments outside what the entry point covers (the caller falls back) */

/* Boundary-condition codes of LBMConfig.bc_x_* (lbm/config.py:61-66). */
#define FSI_BC_PERIODIC 0
#define FSI_BC_PRESSURE 1
#define FSI_BC_VELOCITY 2

/* Population storage (FsiLbmParams.storage).  The 16-bit forms store the
 * deviation F_s - w_s from the rest weights; collisions run in FP32. */
#define FSI_STORE_F32 0
#define FSI_STORE_F16_DEV 1
#define FSI_STORE_BF16_DEV 2

/* Collision operator (FsiLbmParams.collision). */
#define FSI_COLLIDE_MRT 0         /* the reference's MRT with Guo forcing (fluid.py:399-413) */
#define FSI_COLLIDE_REGULARIZED 1 /* regularised BGK, tau = 1/s_v (no reference counterpart) */

/* Open-boundary treatment of the x faces (FsiLbmParams.open_bc), the
 * reference's lbm_open_boundary_semantics (fluid.py:538-570). */
#define FSI_OPEN_BC_EQUILIBRIUM 0         /* equilibrium_all_population_reset: Boundary_condition_legacy,
                                             applied by the face cells inside fsi_lbm_substep */
#define FSI_OPEN_BC_REGULARIZED 1         /* regularized_velocity_pressure */
#define FSI_OPEN_BC_REGULARIZED_LIMITED 2 /* regularized_velocity_pressure_limited */
#define FSI_OPEN_BC_CONVECTIVE 3          /* convective_pressure_outlet_experimental */

/* Substep flags. */
#define FSI_SUB_STORED_MOMENTS 1u /* collide with stored rho/vel instead of moments of F_src */
#define FSI_SUB_MOVING_WALL 2u    /* moving-wall bounce-back (streaming_moving_bounceback) */
#define FSI_SUB_GATHER_PREV 4u    /* wall cells gather the previous substep's link forces */
#define FSI_SUB_DIAG 8u           /* record per-link corrections for the diagnostics */
#define FSI_SUB_CELL_FORCE 16u    /* add cell_force to LBMConfig.force */

typedef struct FsiLbmParams {
  int32_t nx, ny, nz;  /* at least 3 each, at most 2^30 cells in all */
  int32_t bc_x_left, bc_x_right;
  float s_v, s_other, rho0;
  float force[3];
  float rho_bc_x_left, rho_bc_x_right;
  float vel_bc_x_left[3], vel_bc_x_right[3];
  float dynamic_solid_threshold;
  int32_t storage;   /* FSI_STORE_*: population storage in F[0], F[1] */
  int32_t collision; /* FSI_COLLIDE_*: collision operator */
  int32_t open_bc;   /* FSI_OPEN_BC_*: x-face treatment; other than FSI_OPEN_BC_EQUILIBRIUM the face
                        cells stream like inner cells and fsi_lbm_open_bc rewrites their unknown
                        populations after every substep */
  int32_t ob_limiter;       /* open_boundary_limiter_enabled: the limited mode clips rho, |u| and the
                               non-equilibrium part and counts; the limited and convective modes
                               apply the population floor when it is set */
  int32_t ob_floor_enabled; /* open_boundary_population_floor is set */
  float ob_rho_min, ob_rho_max, ob_u_max, ob_noneq_cap, ob_floor;
  /* Launch shape of fsi_lbm_substep (results do not depend on it; an
   * autotuner picks it per lattice, all 0 = the built-in choice):
   * launch_block  cells per workgroup, a run of the (j,k) plane of one
   *               i-plane: 128, 256 or 512 (0 = 256; FP32 storage only
   *               takes other than 256);
   * launch_minw   waves per SIMD the kernel's register budget allows: 1 (no
   *               cap) or 5 (capped, populations loaded with the flags);
   *               0 = 1 when the grid fits the chip in one wave generation;
   * launch_linear 1 = workgroups take tiles in dispatch order instead of
   *               one contiguous slab of i-planes per XCD. */
  int32_t launch_block, launch_minw, launch_linear, launch_pad;
} FsiLbmParams;

/* Device-side counters of the LBM (all reductions are order-independent:
 * integer adds, max of non-negative floats, fp64 sums). */
typedef struct FsiLbmCounters {
  int32_t bb_link_count;
  uint32_t bb_max_correction_bits; /* float bits, non-negative */
  int32_t mb_sample_count;
  uint32_t accum_norm_max_bits;
  uint32_t mean_norm_max_bits;
  int32_t pad0;
  double bb_net_solid_force[3];
  /* open-boundary limiter counters, in the order rho clips, velocity clips,
   * non-equilibrium clips, population floors, reconstructed populations:
   * of the last fsi_lbm_open_bc, and summed over the run (cleared by the
   * caller) */
  uint32_t ob_step[5];
  uint32_t pad1;
  uint64_t ob_run[5];
} FsiLbmCounters;

typedef struct FsiLbmState {
  float* F[2];            /* 19*N each */
  float* rho;             /* N */
  float* vel;             /* 3*N */
  int8_t* solid;          /* N */
  int8_t* static_solid;   /* N */
  int8_t* old_solid;      /* N */
  uint32_t* cellinfo;     /* N: solid bit, link / incoming masks, moving-wall bit */
  float* solid_phi;       /* N */
  float* solid_mass;      /* N */
  float* solid_vel;       /* 3*N */
  float* cell_force;      /* 3*N */
  float* hydro_force;     /* 3*N */
  float* link_c[2];       /* 18*N each: per-link force scalar f_s + bounced of links that bounce off a
                             moving wall or leave a BC-face cell, double-buffered by substep parity
                             (links into walls at rest are read back from the populations) */
  float* link_corr;       /* 18*N: per-link moving-wall correction of the DIAG substep */
  double* hydro_accum;    /* 4*N, per cell: subcycle force accumulator x, y, z (fp64) and the running
                             max of its norm over the substeps */
  double* reduce_ws;      /* FSI_LBM_REDUCE_WS_DOUBLES(N): per-workgroup records of the finalize */
  FsiLbmCounters* counters;
} FsiLbmState;

/* LBM field reductions for one diagnostics row (fluid.py:1931-1968,
 * fsi_diagnostics.py:8-97). */
typedef struct FsiLbmStats {
  double fluid_count;
  double rho_min, rho_max, mass_total;
  double max_v, force_norm_max;
  double fluid_sum_v[3];
  double phi_weight_sum, phi_weighted_v[3];
  double far_count, far_sum_v[3];
  double max_cell_force_norm_all, max_hydro_force_norm_all;
  double net_cell_force[3], net_hydro_force[3];
  double mass_rho_all;
} FsiLbmStats;

typedef struct FsiMpmParams {
  int32_t gnx, gny, gnz;   /* MPM grid nodes per axis (n_grid each in the reference); at least 3 each */
  int32_t n_particles;
  int32_t bound;           /* sticky-wall width (mpm/config.py:21) */
  int32_t damping_enabled, lock_z_enabled;
  float dx, inv_dx, dt;
  float mu, la;
  float gravity[3];
  float velocity_damping;
  float domain_max[3];     /* particle clamp upper bound + dx (1.0 per axis in the reference) */
} FsiMpmParams;

typedef struct FsiMpmState {
  float* x;        /* 3*P */
  float* v;        /* 3*P */
  float* C;        /* 9*P */
  float* F;        /* 9*P */
  float* Jp;       /* P */
  float* mass;     /* P */
  float* vol0;     /* P */
  int8_t* fixed_mask; /* P */
  float* fixed_x;  /* 3*P */
  float* grid_v;   /* 3*G */
  float* grid_m;   /* G */
  float* grid_f;   /* 3*G */
} FsiMpmState;

/* Coupling constants (projection.py:9-15, penalty.py:27-33, moving_boundary.py:24-32). */
typedef struct FsiCouplingParams {
  float inv_dx_norm;
  float cell_volume_norm;
  float vel_scale_norm_to_lbm;
  float eps_mass;
  float beta_lbm, penalty_phi_min, force_cap_lbm;
  float force_density_scale;      /* dx_norm / lbm_dt_phys^2 */
  float reaction_scale;
  float mb_phi_min, mb_force_cap_norm;
  float area_scale;               /* moving-boundary reaction factor of the link-area transfer
                                     (link_area.py:150-157); 1 for the engineering transfer */
} FsiCouplingParams;

/* Device counters of projection and reaction transfer. */
typedef struct FsiCouplingCounters {
  double projected_mass, projected_volume_raw, projected_volume_clamped;
  uint32_t max_phi_raw_bits;
  int32_t active_cell_count;
  int32_t active_reaction_particle_count;
  int32_t active_force_cell_count;
  uint32_t max_particle_reaction_bits;
  uint32_t max_grid_reaction_bits;
  uint32_t max_cell_force_bits, max_hydro_force_bits;
  double net_particle_reaction[3];
  double net_grid_reaction[3];
  double net_cell_force[3], net_hydro_force[3];
} FsiCouplingCounters;

typedef struct FsiMpmStats {
  double min_x[3], max_x[3];
  double max_speed, min_J, max_J, total_mass;
  double mass_weighted_v[3];
} FsiMpmStats;

/* Records of the moving-boundary finalize: 8 doubles per 256-cell workgroup. */
#define FSI_LBM_REDUCE_WS_DOUBLES(n_cells) (8 * (((n_cells) + 255) / 256))

/* Workgroup size of the per-particle kernels. */
#define FSI_PARTICLE_BLOCK 64

/* ---------------------------------------------------------------------------
 * Batched form: many independent FSI instances (mixed sizes and modes) in one
 * launch per operation.  `inst` is a device array of descriptors; a launch
 * covers the workgroups listed in a device block map (two int32 per
 * workgroup: instance index, workgroup index within that instance's domain),
 * built by fsi_batch_block_map for the instances an operation applies to.
 * No reference counterpart: the reference runs one instance per process.
 * ------------------------------------------------------------------------- */
#define FSI_BATCH_MODE_NONE 0
#define FSI_BATCH_MODE_PENALTY 1
#define FSI_BATCH_MODE_MOVING_BOUNDARY 2

#define FSI_BATCH_DOMAIN_CELLS 0
#define FSI_BATCH_DOMAIN_NODES 1
#define FSI_BATCH_DOMAIN_PARTICLES 2 /* workgroups of FSI_PARTICLE_BLOCK particles */
#define FSI_BATCH_DOMAIN_PLANES 3    /* substep tiles: nx * ceil(ny*nz / 256) workgroups */
#define FSI_BATCH_DOMAIN_INSTANCES 4 /* one workgroup per instance */
#define FSI_BATCH_DOMAIN_BOX 5       /* min(FSI_BATCH_BOX_BLOCKS, ceil(cells / 256)) workgroups per instance
                                        that loop over the instance's tracked box (cells or grid nodes) */
#define FSI_BATCH_BOX_BLOCKS 64

typedef struct FsiBatchInstance {
  FsiLbmParams lp;
  FsiLbmState ls;
  FsiMpmParams mp;
  FsiMpmState ms;
  FsiCouplingParams cp;
  FsiCouplingCounters* cc;
  int32_t mode;       /* FSI_BATCH_MODE_* */
  int32_t cur0;       /* buffer holding F at substep 0 of the current exchange */
  int32_t n_substeps; /* LBM substeps per exchange (1 for penalty / none) */
  int32_t mpm_substeps;
  int32_t active;     /* 0 = slot unused */
  int32_t cell_force; /* 1 = add cell_force (penalty) */
  int32_t pad0;
  double* react_ws;   /* FSI_COUPLED_WORKSPACE_DOUBLES(n_particles): reaction records of MPM_P2G_REACT */
  int32_t* boxes;     /* FSI_BATCH_BOX_INTS, zeroed when the instance enters the batch (see below) */
  float* res_halo;    /* resident != 0: FSI_RES_HALO_FLOATS(cells) floats (fsi_lbm_subcycle_resident's halo) */
  uint32_t* res_sync; /* resident != 0: FSI_RES_SYNC_WORDS(nx, ny) words, zero before the first exchange, kept */
  int32_t resident;   /* 1: a moving-boundary instance whose subcycle runs as LBM_RESIDENT (its PLANES domain is
                         empty, so the LBM_SUBSTEP* launches leave it out); it must meet what
                         fsi_lbm_subcycle_resident covers, and the device must hold its blocks at once */
  int32_t pad1;
} FsiBatchInstance;

/* Bounding boxes that confine the per-cell and per-node passes of the
 * coupling to where the particles are.  Two boxes per instance, at
 * FSI_BATCH_BOX_PROJ (lattice cells: projection fields and penalty forces)
 * and FSI_BATCH_BOX_MPM (MPM grid nodes): int32 [0] valid, [1..3] lower and
 * [4..6] upper corner (inclusive) of `cur`, [7..9] and [10..12] of `prev`.
 * BOX_ROLL makes `cur` the new `prev` (the whole lattice / grid while not
 * valid, as after zeroing) and empties `cur`; the particle passes then grow
 * `cur` to their stencils (PROJ_PARTICLES, MPM_P2G, MPM_P2G_REACT); the
 * clears cover `prev` (all the last pass wrote), NORMALIZE and the grid
 * updates `cur`, PENALTY the box around both.  Cells and nodes outside hold
 * zero before and after, so the results are those of whole-lattice passes. */
#define FSI_BATCH_BOX_INTS 32
#define FSI_BATCH_BOX_PROJ 0
#define FSI_BATCH_BOX_MPM 16

/* Operations of fsi_batch_launch. */
#define FSI_BATCH_OP_LBM_SUBSTEP 0  /* one fused substep (t = substep index within the exchange) */
#define FSI_BATCH_OP_LBM_REBUILD 1  /* rebuild flags/links (dynamic for moving-boundary instances) */
#define FSI_BATCH_OP_LBM_MOMENTS 2
#define FSI_BATCH_OP_LBM_MB_BEGIN 3
#define FSI_BATCH_OP_LBM_MB_FINALIZE 4
#define FSI_BATCH_OP_PROJ_CLEAR 5
#define FSI_BATCH_OP_PROJ_PARTICLES 6
#define FSI_BATCH_OP_PROJ_NORMALIZE 7
#define FSI_BATCH_OP_PENALTY 8
#define FSI_BATCH_OP_MPM_CLEAR 9
#define FSI_BATCH_OP_MPM_P2G 10
#define FSI_BATCH_OP_MPM_REACTION 11
#define FSI_BATCH_OP_MPM_GRID_UPDATE 12
#define FSI_BATCH_OP_MPM_G2P 13
#define FSI_BATCH_OP_COUNTERS_CLEAR 14 /* zero counters `t` of the map's instances (INSTANCES domain) */
#define FSI_BATCH_OP_MPM_P2G_REACT 15  /* P2G + the instance's reaction in one particle pass (as
                                          fsi_mpm_substep_coupled), records into react_ws */
#define FSI_BATCH_OP_MPM_GRID_UPDATE_REACT 16 /* grid update; folds react_ws into the counters */
/* LBM_SUBSTEP for a map of one mode only (moving-boundary, penalty, none):
 * the mode's flags are compiled into the kernel. */
#define FSI_BATCH_OP_LBM_SUBSTEP_MB 17
#define FSI_BATCH_OP_LBM_SUBSTEP_PENALTY 18
#define FSI_BATCH_OP_LBM_SUBSTEP_NONE 19
#define FSI_BATCH_OP_LBM_MB_FINALIZE_FOLD 20 /* after MB_FINALIZE: fold its records into the counters */
#define FSI_BATCH_OP_BOX_ROLL 21 /* t = 0: roll the projection box, 1: the MPM box */
/* The whole moving-wall subcycle of ONE instance with `resident` set, in one
 * launch with its lattice on chip (fsi_lbm_subcycle_resident from cur0 over
 * n_substeps substeps, with the flags of the LBM_SUBSTEP_MB launches it
 * replaces).  Its block map lists (instance, b) for b = 0 ..
 * FSI_RES_BLOCKS(nx, ny) - 1 of that one instance, in order: a launch must
 * not hold the blocks of two instances, which could not all be resident. */
#define FSI_BATCH_OP_LBM_RESIDENT 22

/* Domain of each operation (the map passed to fsi_batch_launch must be built
 * for it): LBM_SUBSTEP, LBM_SUBSTEP_MB / _PENALTY / _NONE -> PLANES; REBUILD, MOMENTS, MB_BEGIN,
 * MB_FINALIZE -> CELLS; PROJ_CLEAR, PROJ_NORMALIZE, PENALTY, MPM_CLEAR, MPM_GRID_UPDATE,
 * MPM_GRID_UPDATE_REACT -> BOX; PROJ_PARTICLES, MPM_P2G, MPM_REACTION,
 * MPM_P2G_REACT, MPM_G2P -> PARTICLES; LBM_MB_FINALIZE_FOLD, BOX_ROLL -> INSTANCES. */

/* Block map for one domain from the host copy of the descriptors: instances
 * that are active, whose mode bit is in `mode_mask` (bit m = 1 <<
 * FSI_BATCH_MODE_m) and that have more than `min_mpm` MPM substeps per LBM
 * step.  Writes two int32 per workgroup into map_host (may be NULL to count)
 * and returns the workgroup count. */
int fsi_batch_block_map(const FsiBatchInstance* inst_host, int n_inst, int domain, int mode_mask, int min_mpm,
                        int32_t* map_host);


/* One batched operation over the workgroups of a device block map.  `t` is
 * the substep index within the exchange (LBM_SUBSTEP, MOMENTS) or, for
 * COUNTERS_CLEAR, which counters (0 projection, 1 penalty, 2 reaction,
 * 3 bounce-back). */
int fsi_batch_launch(int op, const FsiBatchInstance* inst_dev, int n_inst, const int32_t* block_map_dev,
                     int n_blocks, int t, void* stream);

/* A batched exchange as a HIP graph built from its launch list (no stream
 * capture): one kernel node per operation, on the chain its `branch` names.
 * An operation is what fsi_batch_launch takes (n_blocks > 0).  When
 * instances come and go, fsi_batch_graph_update rewrites the block maps and
 * workgroup counts of the instantiated graph in place
 * (hipGraphExecKernelNodeSetParams) for a launch list of the same operations,
 * substep indices and branches; it fails (FSI_E_INVALID) on any other list.  The
 * instance table `inst_dev` and its slot count are fixed for the graph's
 * life: rewrite the table's contents, not its address. */
typedef struct FsiBatchOp {
  int32_t op;
  int32_t t;
  const int32_t* block_map; /* device */
  int32_t n_blocks;
  int32_t branch; /* graph chain: FSI_BATCH_BRANCH_MAIN, a side chain 1..FSI_BATCH_SIDE_CHAINS, or
                     FSI_BATCH_BRANCH_JOIN; fsi_batch_launch and the serial emulation run the list in order */
} FsiBatchOp;

/* Graph dependencies: a MAIN node follows the previous MAIN (or JOIN) node; a
 * node of side chain k follows the previous node of chain k, the first one of
 * a side chain the main node before it; a JOIN node follows the last nodes of
 * the main and every open side chain and ends those side chains.  The chains
 * may run at the same time: the list puts on each chain only operations on
 * other instances than the other chains' until the join, and the list order is
 * a valid serial order. */
#define FSI_BATCH_BRANCH_MAIN 0
#define FSI_BATCH_SIDE_CHAINS 3
#define FSI_BATCH_BRANCH_JOIN (-1)

int fsi_batch_graph_create(const FsiBatchOp* ops, int n_ops, const FsiBatchInstance* inst_dev, int n_inst,
                           void** graph_out);
int fsi_batch_graph_update(void* graph, const FsiBatchOp* ops, int n_ops);
int fsi_batch_graph_launch(void* graph, void* stream);
int fsi_batch_graph_destroy(void* graph);

const char* fsi_last_error(void);
int fsi_version(void);
/* Writes sizeof of FsiLbmParams, FsiLbmState, FsiLbmCounters, FsiLbmStats,
 * FsiMpmParams, FsiMpmState, FsiCouplingParams, FsiCouplingCounters,
 * FsiMpmStats, FsiBatchInstance, FsiBins, FsiBatchOp (in that order, up to n entries) so bindings can verify their
 * struct layouts.  Returns the number of entries written. */
int fsi_abi_sizes(int32_t* out, int n);
int fsi_sync(void* stream);

/* LBMFluid3D.init / static_init (fluid.py:329-344, :357-373):
 * F[0] = F[1] = feq(rho0, 0), rho = rho0, vel = 0 on every cell. */
int fsi_lbm_init(const FsiLbmParams* p, const FsiLbmState* st, void* stream);

/* update_dynamic_solid + reinitialize_new_fluid_cells (fluid.py:1848-1877),
 * plus the per-cell link masks that the reference recomputes inside every
 * streaming pass (fluid.py:432-440, :489-519).  `dynamic` = 0 keeps `solid`
 * as-is (penalty / none modes); `cur` is the buffer holding the current F. */
int fsi_lbm_rebuild(const FsiLbmParams* p, const FsiLbmState* st, int dynamic, int cur, void* stream);

/* One fused LBM substep: colission (fluid.py:399-413), streaming1 /
 * streaming_moving_bounceback (:432-440, :489-519), Boundary_condition_legacy
 * (:927-1024, x faces), reading F[cur] and writing F[1-cur].  With
 * FSI_SUB_MOVING_WALL the wall cells accumulate hydro forces
 * (accumulate_moving_boundary_force_sample, :469-473) one substep late,
 * deterministically.  `parity` selects link_c[parity] for writing. */
int fsi_lbm_substep(const FsiLbmParams* p, const FsiLbmState* st, int cur, int parity,
                    uint32_t flags, void* stream);

/* The moving-wall subcycle of FSIDriver3D (fsi_driver.py:533-550: n_substeps
 * x step_moving_bounceback, fluid.py:489-519, with the per-substep
 * accumulate_moving_boundary_force_sample, :469-473) in one launch that
 * keeps the lattice on chip: the same results as n calls of
 * fsi_lbm_substep(cur + t, parity t % 2, flags_t) for t = 0 .. n-1, where
 * flags_t = the given flags with FSI_SUB_STORED_MOMENTS on t = 0 only,
 * FSI_SUB_GATHER_PREV on t > 0 only and FSI_SUB_DIAG on t = n-1 only.  Reads
 * F[cur], writes F[(cur + n) % 2], and leaves the last substep's links for
 * fsi_lbm_mb_finalize(cur + n, parity (n-1) % 2).  `flags` must hold
 * FSI_SUB_MOVING_WALL.
 *
 * Covers FSI_STORE_F32, FSI_OPEN_BC_EQUILIBRIUM, nz = 64, nx and ny multiples
 * of 4, and at most as many blocks of 4 x 4 x 64 cells as the device holds
 * workgroups of 1024 work-items at once; returns FSI_E_UNSUPPORTED otherwise
 * (use fsi_lbm_substep).  halo_dev: FSI_RES_HALO_FLOATS(N) floats of scratch.
 * sync_dev: FSI_RES_SYNC_WORDS(nx, ny) words, zero before the first call and
 * kept between calls: word 64 b counts the substeps block b has finished,
 * word 64 * FSI_RES_BLOCKS(nx, ny) counts waits for a neighbour block that
 * timed out (the blocks then go on without it: nonzero means the results of
 * that call are invalid, and every word must be cleared before the next). */
#define FSI_RES_BLOCKS(nx, ny) (((nx) / 4) * ((ny) / 4))
#define FSI_RES_SYNC_WORDS(nx, ny) (64 * (FSI_RES_BLOCKS(nx, ny) + 1))
#define FSI_RES_HALO_FLOATS(n_cells) (2 * 18 * (n_cells))
int fsi_lbm_subcycle_resident(const FsiLbmParams* p, const FsiLbmState* st, int cur, int n_substeps,
                              uint32_t flags, float* halo_dev, uint32_t* sync_dev, void* stream);

/* Boundary_condition on the x faces (fluid.py:538-570, :927-958,
 * :1027-1111, :1349-1493) for F[buf] in place: p->open_bc selects the legacy
 * reset of all populations or the regularized / limited / convective
 * reconstruction of the unknown populations of the fluid face cells from
 * their inner neighbours.  The neighbours' previous moments are the stored
 * rho, vel with FSI_SUB_STORED_MOMENTS, else the moments of F[1-buf] (with
 * cell_force under FSI_SUB_CELL_FORCE): after fsi_lbm_substep(cur, ...) call
 * it with buf = 1 - cur and the substep's flags.  Sets ob_step and adds to
 * ob_run of the counters. */
int fsi_lbm_open_bc(const FsiLbmParams* p, const FsiLbmState* st, int buf, uint32_t flags, void* stream);

/* streaming3 moments (fluid.py:1795-1814): rho, vel from F[cur]. */
int fsi_lbm_moments(const FsiLbmParams* p, const FsiLbmState* st, int cur, uint32_t flags,
                    void* stream);

/* clear_moving_boundary_force_accumulator (fluid.py:461-466) and the
 * clear of clear_moving_boundary_diagnostics (:444-458). */
int fsi_lbm_mb_begin(const FsiLbmParams* p, const FsiLbmState* st, void* stream);

/* finalize_moving_boundary_diagnostics + finalize_moving_boundary_force_accumulator
 * (fluid.py:522-536, :476-486): wall cells gather the link forces of the last
 * substep, which wrote its populations to F[cur] and its link scalars to
 * link_c[parity]; hydro_force = accum / n_samples when `subcycled`, else the
 * last substep's link force. */
int fsi_lbm_mb_finalize(const FsiLbmParams* p, const FsiLbmState* st, int cur, int parity,
                        int n_samples, int subcycled, void* stream);

/* Per-direction statistics of the links of the last moving-wall substep
 * (bb_*_by_dir of streaming_moving_bounceback, fluid.py:489-519; read by
 * get_moving_boundary_directional_stats, fluid.py:2080-2090), which wrote
 * its populations to F[cur] and its link scalars to link_c[parity] with
 * FSI_SUB_DIAG.  out_dev: 19 records of FSI_LINK_STATS_VALUES doubles (link
 * count, solid force x, y, z, sum and max of |correction|), direction 0 all
 * zero; the fluid impulse is the negated solid force.  workspace_dev holds
 * FSI_LINK_STATS_WS_DOUBLES doubles.  Summed per workgroup in cell order,
 * then in workgroup order: deterministic. */
#define FSI_LINK_STATS_VALUES 6
#define FSI_LINK_STATS_BLOCKS 64
#define FSI_LINK_STATS_WS_DOUBLES (18 * FSI_LINK_STATS_BLOCKS * FSI_LINK_STATS_VALUES)
int fsi_lbm_link_stats(const FsiLbmParams* p, const FsiLbmState* st, int cur, int parity, double* workspace_dev,
                       double* out_dev, void* stream);

/* Number of per-block partial records the two-pass reductions use; their
 * workspaces hold FSI_REDUCE_BLOCKS records of the reduced struct. */
#define FSI_REDUCE_BLOCKS 256

/* get_stats + FSIDiagnostics3D LBM reductions (fluid.py:1931-1968,
 * fsi_diagnostics.py:8-97) into a device FsiLbmStats.  Deterministic: per-block
 * partials in `workspace_dev` (FSI_REDUCE_BLOCKS * sizeof(FsiLbmStats) bytes),
 * then a fixed-order final pass. */
int fsi_lbm_reduce_stats(const FsiLbmParams* p, const FsiLbmState* st, FsiLbmStats* workspace_dev,
                         FsiLbmStats* out_dev, void* stream);

/* MPMToLBMProjector3D.project (projection.py:20-97). */
int fsi_project(const FsiLbmParams* lp, const FsiLbmState* ls, const FsiMpmParams* mp,
                const FsiMpmState* ms, const FsiCouplingParams* cp, FsiCouplingCounters* cc,
                void* stream);

/* PenaltyFSICoupler3D.clear_force_fields + build_penalty_force (penalty.py:42-88). */
int fsi_penalty_force(const FsiLbmParams* lp, const FsiLbmState* ls, const FsiCouplingParams* cp,
                      FsiCouplingCounters* cc, void* stream);

/* MPMSolid3D.clear_grid (solid.py:213-217): zero grid_v, grid_m, grid_f. */
int fsi_mpm_clear_grid(const FsiMpmParams* mp, const FsiMpmState* ms, void* stream);

/* MPMSolid3D.p2g (solid.py:220-249); writes Jp = det F. */
int fsi_mpm_p2g(const FsiMpmParams* mp, const FsiMpmState* ms, void* stream);

/* Reaction transfer onto grid_f: mode 1 = PenaltyFSICoupler3D.add_lbm_reaction_to_mpm_grid
 * (penalty.py:90-129), mode 2 = MovingBoundaryFSICoupler3D.add_moving_boundary_reaction_to_mpm_grid
 * (moving_boundary.py:58-111). */
int fsi_mpm_reaction(const FsiLbmParams* lp, const FsiLbmState* ls, const FsiMpmParams* mp,
                     const FsiMpmState* ms, const FsiCouplingParams* cp, FsiCouplingCounters* cc,
                     int mode, void* stream);

/* MPMSolid3D.grid_update (solid.py:252-259) and g2p (:262-315). */
int fsi_mpm_grid_update(const FsiMpmParams* mp, const FsiMpmState* ms, void* stream);
int fsi_mpm_g2p(const FsiMpmParams* mp, const FsiMpmState* ms, void* stream);

/* Workspace of fsi_mpm_substep_coupled: 9 doubles per particle workgroup. */
#define FSI_COUPLED_WORKSPACE_DOUBLES(n_particles) \
  (9 * (((n_particles) + FSI_PARTICLE_BLOCK - 1) / FSI_PARTICLE_BLOCK))

/* One MPM substep with its coupling reaction, as the driver sequences issue it
 * (fsi_driver.py:497-517 and :552-575): clear_grid, then p2g together with
 * the reaction transfer in one pass over the particles, then grid_update
 * (which also folds the reaction counters, deterministically), then g2p.
 * mode 0: plain MPMSolid3D.substep (lp, ls, cp, cc and workspace may be
 * NULL); 1: with add_lbm_reaction_to_mpm_grid; 2: with
 * add_moving_boundary_reaction_to_mpm_grid.  Results equal those of the
 * separate entry points called in that order, up to the summation order of
 * the float scatter.  workspace: FSI_COUPLED_WORKSPACE_DOUBLES(n_particles). */
int fsi_mpm_substep_coupled(const FsiLbmParams* lp, const FsiLbmState* ls, const FsiMpmParams* mp,
                            const FsiMpmState* ms, const FsiCouplingParams* cp, FsiCouplingCounters* cc,
                            int mode, double* workspace, void* stream);

/* fsi_mpm_substep_coupled without the caller's workspace, for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  No counterpart in the reference.
 * The reaction records live in a fixed scratch array of the library, one per
 * GPU, sized for FSI_NOALLOC_MAX_PARTICLES particles (2.25 MiB); larger
 * particle counts are rejected with FSI_E_INVALID.  Calls no hipMalloc or
 * hipFree.  Every call on one GPU shares that GPU's scratch: calls that
 * overlap on one GPU (other streams, other host threads) race. */
#define FSI_NOALLOC_MAX_PARTICLES (1 << 21)
int fsi_mpm_substep_coupled_noalloc(const FsiLbmParams* lp, const FsiLbmState* ls, const FsiMpmParams* mp,
                                    const FsiMpmState* ms, const FsiCouplingParams* cp, FsiCouplingCounters* cc,
                                    int mode, void* stream);

/* n calls of fsi_mpm_substep_coupled (the MPM substeps of one FSI step,
 * fsi_driver.py:497-517 and :552-575: the lattice is not changed in between)
 * as one launch that keeps the particles on chip: one FSI_PARTICLE_BLOCK
 * workgroup per particle block, all resident at once, each holding its
 * particles' x, v, C and F in registers, with one grid-wide barrier per
 * substep (three grid buffers in `scratch`) and the grid update applied where
 * g2p gathers.  The state after the call is that of the n calls: particles,
 * counters (the last substep's reaction), grid_v / grid_m (the last
 * substep's updated grid), grid_f zero.
 * Returns FSI_E_UNSUPPORTED (nothing launched) when the workgroups would not
 * all be resident on the device, when there are no particles, or when the
 * reaction samples the lattice at another spacing than the MPM grid
 * (cp->inv_dx_norm != mp->inv_dx): the caller then runs the n calls.
 * workspace: FSI_COUPLED_WORKSPACE_DOUBLES(n_particles); scratch:
 * FSI_MPM_RES_SCRATCH_FLOATS(grid nodes) floats and sync:
 * FSI_MPM_RES_SYNC_WORDS words, both zeroed before the first call; each call
 * leaves its scratch and barrier words zero again.  Word
 * FSI_MPM_RES_TIMEOUT_WORD of sync counts barrier waits that gave up (never,
 * while every workgroup runs): after one the results are invalid and the
 * caller zeroes scratch and sync.  Calls sharing `sync` or `scratch` must not
 * overlap. */
#define FSI_MPM_RES_SCRATCH_FLOATS(grid_nodes) (12 * (long long)(grid_nodes))
#define FSI_MPM_RES_SYNC_WORDS 640
#define FSI_MPM_RES_TIMEOUT_WORD 576
int fsi_mpm_subcycle_coupled(const FsiLbmParams* lp, const FsiLbmState* ls, const FsiMpmParams* mp,
                             const FsiMpmState* ms, const FsiCouplingParams* cp, FsiCouplingCounters* cc, int mode,
                             int n, double* workspace, float* scratch, uint32_t* sync, void* stream);

/* ---------------------------------------------------------------------------
 * Particle binning and the particle-ordered (deterministic) transfers.
 *
 * A particle's bin is its quadratic B-spline stencil base per axis,
 * base = trunc(x * inv_dx - 0.5) (solid.py:222-224), offset by 2 over an
 * (nx + 3) x (ny + 3) x (nz + 3) box that holds every base whose stencil
 * reaches the nx x ny x nz node grid; bin FSI_BIN_COUNT(...) (the last) is the
 * overflow bin of any other particle.  start[b] .. start[b + 1] index order[],
 * which lists each bin's particles in increasing particle index: a counting
 * sort, deterministic whatever the thread schedule.
 * ------------------------------------------------------------------------- */
#define FSI_BIN_COUNT(nx, ny, nz) ((long long)((nx) + 3) * ((ny) + 3) * ((nz) + 3))
/* ints of FsiBins.work for `nbins` = FSI_BIN_COUNT(...) bins */
#define FSI_BIN_WS_INTS(nbins) ((nbins) + 1 + ((nbins) + 1 + 1023) / 1024)
typedef struct FsiBins {
  int32_t* key;   /* n_particles: each particle's bin */
  int32_t* start; /* nbins + 2 */
  int32_t* order; /* n_particles */
  int32_t* work;  /* FSI_BIN_WS_INTS(nbins) */
} FsiBins;

/* Bins the particles at positions x (SoA, 3 x n_particles) on spacing
 * 1 / inv_dx over an nx x ny x nz node grid (the binning key of p2g and of the
 * projection's scatter, solid.py:222-224, projection.py:41-44). */
int fsi_bin_particles(const float* x, int32_t n_particles, float inv_dx, int32_t nx, int32_t ny, int32_t nz,
                      const FsiBins* bins, void* stream);

/* Floats per particle of fsi_mpm_substep_ordered's `terms` buffer. */
#define FSI_ORDERED_TERMS_FLOATS 28

/* fsi_mpm_substep_coupled with every node sum taken in particle order: the
 * particles are binned (`bins`, sized for the MPM grid), a particle pass
 * stores p2g's terms and the reaction force in `terms`, and each grid node
 * then sums its contributors in increasing particle index - the order of the
 * reference's kernels run sequentially - and applies grid_update; g2p
 * follows.  Deterministic: two runs give identical bits.  The reaction is
 * scattered into grid_f as in the reference (solid.py:213-315,
 * penalty.py:90-129, moving_boundary.py:58-111); the lattice spacing must be
 * the MPM grid's (cp->inv_dx_norm == mp->inv_dx).  workspace:
 * FSI_COUPLED_WORKSPACE_DOUBLES(n_particles); terms: FSI_ORDERED_TERMS_FLOATS
 * per particle. */
int fsi_mpm_substep_ordered(const FsiLbmParams* lp, const FsiLbmState* ls, const FsiMpmParams* mp,
                            const FsiMpmState* ms, const FsiCouplingParams* cp, FsiCouplingCounters* cc, int mode,
                            const FsiBins* bins, float* terms, double* workspace, void* stream);

/* MPMToLBMProjector3D.project (projection.py:20-97) with each lattice cell's
 * sums taken in particle order (`bins` sized for the lattice); deterministic. */
int fsi_project_ordered(const FsiLbmParams* lp, const FsiLbmState* ls, const FsiMpmParams* mp,
                        const FsiMpmState* ms, const FsiCouplingParams* cp, FsiCouplingCounters* cc,
                        const FsiBins* bins, void* stream);

/* MPMSolid3D.reduce_stats + FSIDiagnostics3D.solid_mean_velocity_norm (solid.py:342-357,
 * fsi_diagnostics.py:117-124). */
int fsi_mpm_reduce_stats(const FsiMpmParams* mp, const FsiMpmState* ms, FsiMpmStats* workspace_dev,
                         FsiMpmStats* out_dev, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* FSI_HIP_H */
