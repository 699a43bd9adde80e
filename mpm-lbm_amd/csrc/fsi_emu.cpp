// fsi_emu.cpp - serial CPU emulation of libfsi_hip.so, for tests only.
//
// Exports the same C ABI as fsi_kernels.hip (include/fsi_hip.h) and runs the
// very same per-cell / per-particle bodies of fsi_core.h in serial loops over
// host memory.  The GPU library is the product; this emulation lets the CPU test
// suite check the kernels' arithmetic, indexing and boundary handling against
// the oracle on a machine without a GPU.  Built into libfsi_emu.so, which no
// product code loads.
#include <string.h>
#include <vector>

#include <string>

#include "fsi_core.h"

using namespace fsi;

namespace {
thread_local std::string g_err;
int fail(int code, const char* m) {
  g_err = m;
  return code;
}
long long ncells(const FsiLbmParams* p) { return (long long)p->nx * p->ny * p->nz; }

template <typename F>
void by_storage(int code, F&& fn) {
  if (code == FSI_STORE_F16_DEV)
    fn(PopF16{});
  else if (code == FSI_STORE_BF16_DEV)
    fn(PopBF16{});
  else
    fn(PopF32{});
}
}  // namespace

extern "C" {

const char* fsi_last_error(void) { return g_err.c_str(); }
int fsi_version(void) { return 1; }

int fsi_abi_sizes(int32_t* out, int n) {
  const int32_t sz[12] = {(int32_t)sizeof(FsiLbmParams),      (int32_t)sizeof(FsiLbmState),
                         (int32_t)sizeof(FsiLbmCounters),    (int32_t)sizeof(FsiLbmStats),
                         (int32_t)sizeof(FsiMpmParams),      (int32_t)sizeof(FsiMpmState),
                         (int32_t)sizeof(FsiCouplingParams), (int32_t)sizeof(FsiCouplingCounters),
                         (int32_t)sizeof(FsiMpmStats),       (int32_t)sizeof(FsiBatchInstance),
                         (int32_t)sizeof(FsiBins),           (int32_t)sizeof(FsiBatchOp)};
  int k = 0;
  for (; k < n && k < 12; ++k) out[k] = sz[k];
  return k;
}
int fsi_sync(void*) { return FSI_OK; }

int fsi_lbm_init(const FsiLbmParams* p, const FsiLbmState* st, void*) {
  by_storage(p->storage, [&](auto tag) {
    using Pop = decltype(tag);
    for (long long c = 0; c < ncells(p); ++c) lbm_init_cell<Pop>(*p, *st, c);
  });
  return FSI_OK;
}

int fsi_lbm_rebuild(const FsiLbmParams* p, const FsiLbmState* st, int dynamic, int cur, void*) {
  if (p->nx < 3 || p->ny < 3 || p->nz < 3) return fail(FSI_E_INVALID, "lattice must be at least 3 cells along every axis");
  // The GPU evaluates every cell from the pre-pass flags; emulate that by
  // snapshotting `solid` (rebuild reads neighbours' new flags from
  // static_solid / solid_phi only, and its own previous flag).
  by_storage(p->storage, [&](auto tag) {
    using Pop = decltype(tag);
    for (long long c = 0; c < ncells(p); ++c) lbm_rebuild_cell<Pop>(*p, *st, dynamic, cur, c);
  });
  return FSI_OK;
}

int fsi_lbm_substep(const FsiLbmParams* p, const FsiLbmState* st, int cur, int parity, uint32_t flags, void*) {
  if (p->nx < 3 || p->ny < 3 || p->nz < 3) return fail(FSI_E_INVALID, "lattice must be at least 3 cells along every axis");
  by_storage(p->storage, [&](auto tag) {
    using Pop = decltype(tag);
    for (int i = 0; i < p->nx; ++i)
      for (int j = 0; j < p->ny; ++j)
        for (int k = 0; k < p->nz; ++k) lbm_substep_cell<Pop>(*p, *st, cur, parity, flags, i, j, k);
  });
  return FSI_OK;
}

// The resident subcycle as the GPU runs it: blocks of RES_THREADS cells, each
// with its own copy of the LDS slot planes, the halo planes in `halo`; every
// block finishes phase A before any block starts phase B (the GPU's
// neighbour waits).  The per-lane registers are arrays indexed by cell.
int fsi_lbm_subcycle_resident(const FsiLbmParams* p, const FsiLbmState* st, int cur, int n, uint32_t flags,
                              float* halo, uint32_t* sync, void*) {
  if (const char* why = res_unsupported(*p, *st, cur, n, flags, halo, sync)) return fail(FSI_E_UNSUPPORTED, why);
  const uint32_t N = (uint32_t)ncells(p);
  const int nblk = res_blocks(*p);
  const size_t T = RES_THREADS;
  std::vector<float> lds((size_t)nblk * NLINK * T), uw((size_t)nblk * 3 * RES_RING),
      ou((size_t)nblk * 3 * RES_BJ * RES_NZ);
  std::vector<float> R((size_t)N * Q), U((size_t)N * 3);
  const float* Fin = st->F[cur];
  for (int b = 0; b < nblk; ++b)
    for (int r = 0; r < RES_RING; ++r) {
      const uint32_t y = res_ring_cell(*p, (uint32_t)b, r);
      for (int d = 0; d < 3; ++d) uw[((size_t)b * 3 + d) * RES_RING + r] = st->solid_vel[(size_t)d * N + y];
    }
  for (uint32_t c = 0; c < N; ++c)
    if (!(st->cellinfo[c] & CI_SOLID))
      for (int s = 0; s < Q; ++s) R[(size_t)c * Q + s] = Fin[(size_t)s * N + c];
  for (int step = 0; step < n; ++step) {
    uint32_t fl = flags & (FSI_SUB_MOVING_WALL | FSI_SUB_CELL_FORCE | FSI_SUB_DIAG);
    if (step == 0) fl |= flags & FSI_SUB_STORED_MOMENTS;
    const bool last = step == n - 1;
    const bool gather = (flags & FSI_SUB_GATHER_PREV) && !last;
    float* hp = halo + (size_t)(step & 1) * NLINK * N;
    for (int b = 0; b < nblk; ++b) {
      float* lv = lds.data() + (size_t)b * NLINK * T;
      for (uint32_t t = 0; t < T; ++t) {
        const ResLane L = res_lane(*p, (uint32_t)b, t);
        const uint32_t info = st->cellinfo[L.c];
        float* o = ou.data() + (size_t)b * 3 * RES_BJ * RES_NZ + (size_t)(L.lj * RES_NZ + L.k) * 3;
        if (info & CI_SOLID) {
          if (L.i == p->nx - 2) o[0] = o[1] = o[2] = 0.0f;
          continue;
        }
        const Nbr nb = nbr_of(*p, L.i, L.j, L.k);
        auto dot = [&](int s) {
          const float* w = uw.data() + (size_t)b * 3 * RES_RING + res_ring_slot(L.li, L.lj, L.k, EX[s], EY[s], EZ[s]);
          return wall_dot(s, w[0], w[RES_RING], w[2 * RES_RING]);
        };
        auto put = [&](const float* out) {
          for (int s = 1; s < Q; ++s) {
            if (res_in_block(L.li, L.lj, EX[s], EY[s]))
              lv[(size_t)(s - 1) * T + t] = out[s];
            else
              hp[(size_t)(s - 1) * N + L.c] = out[s];
          }
        };
        auto bounce = [&](int s, float v) {
          if (res_in_block(L.li, L.lj, EX[s], EY[s]))
            lv[(size_t)(LR[s] - 1) * T + (size_t)res_slot(L.li, L.lj, L.k, EX[s], EY[s], EZ[s])] = v;
          else
            hp[(size_t)(LR[s] - 1) * N + (L.c + (uint32_t)nb.d(s))] = v;
        };
        float* u = U.data() + (size_t)L.c * 3;
        res_collide_cell(*p, *st, fl, step & 1, last, N, L.c, info, face_flags(*p, L.i).o, R.data() + (size_t)L.c * Q,
                         u[0], u[1], u[2], dot, put, bounce);
        if (L.i == p->nx - 2) {
          o[0] = u[0];
          o[1] = u[1];
          o[2] = u[2];
        }
      }
    }
    for (int b = 0; b < nblk; ++b) {
      const float* lv = lds.data() + (size_t)b * NLINK * T;
      for (uint32_t t = 0; t < T; ++t) {
        const ResLane L = res_lane(*p, (uint32_t)b, t);
        const uint32_t info = st->cellinfo[L.c];
        const bool fluid = !(info & CI_SOLID);
        if (!fluid && !(gather && (info & CI_MASK))) continue;
        const Nbr nb = nbr_of(*p, L.i, L.j, L.k);
        auto pull = [&](int q) {
          if (res_in_block(L.li, L.lj, -EX[q], -EY[q]))
            return lv[(size_t)(q - 1) * T + (size_t)res_slot(L.li, L.lj, L.k, -EX[q], -EY[q], -EZ[q])];
          return hp[(size_t)(q - 1) * N + (L.c + (uint32_t)nb.d(LR[q]))];
        };
        const float* o = ou.data() + (size_t)b * 3 * RES_BJ * RES_NZ + (size_t)(L.lj * RES_NZ + L.k) * 3;
        const float* u = U.data() + (size_t)L.c * 3;
        res_pull_cell(*p, *st, L, info, face_flags(*p, L.i).o, u[0], u[1], u[2], o[0], o[1], o[2],
                      R.data() + (size_t)L.c * Q, st->hydro_accum + 4 * (size_t)L.c, pull);
      }
    }
  }
  float* Fout = st->F[(cur + n) & 1];
  for (uint32_t c = 0; c < N; ++c)
    if (!(st->cellinfo[c] & CI_SOLID))
      for (int s = 0; s < Q; ++s) Fout[(size_t)s * N + c] = R[(size_t)c * Q + s];
  for (int b = 0; b < nblk; ++b) sync[64 * b] += (uint32_t)n;
  return FSI_OK;
}

int fsi_lbm_open_bc(const FsiLbmParams* p, const FsiLbmState* st, int buf, uint32_t flags, void*) {
  if (p->nx < 3 || p->ny < 3 || p->nz < 3) return fail(FSI_E_INVALID, "lattice must be at least 3 cells along every axis");
  if (p->open_bc < FSI_OPEN_BC_EQUILIBRIUM || p->open_bc > FSI_OPEN_BC_CONVECTIVE)
    return fail(FSI_E_INVALID, "open_bc must be one of FSI_OPEN_BC_*");
  if (buf != 0 && buf != 1) return fail(FSI_E_INVALID, "buf must be 0 or 1");
  if (!st->counters) return fail(FSI_E_INVALID, "open_bc needs counters");
  uint32_t cnt[5] = {0u, 0u, 0u, 0u, 0u};
  by_storage(p->storage, [&](auto tag) {
    using Pop = decltype(tag);
    for (int side = 0; side < 2; ++side)
      for (int j = 0; j < p->ny; ++j)
        for (int k = 0; k < p->nz; ++k) lbm_open_bc_cell<Pop>(*p, *st, buf, flags, side == 1, j, k, cnt);
  });
  for (int q = 0; q < 5; ++q) {
    st->counters->ob_step[q] = cnt[q];
    st->counters->ob_run[q] += cnt[q];
  }
  return FSI_OK;
}

int fsi_lbm_link_stats(const FsiLbmParams* p, const FsiLbmState* st, int cur, int parity, double* ws, double* out,
                       void*) {
  if ((cur != 0 && cur != 1) || (parity != 0 && parity != 1)) return fail(FSI_E_INVALID, "cur and parity must be 0 or 1");
  if (!st->link_c[0] || !st->link_c[1] || !st->link_corr || !ws || !out)
    return fail(FSI_E_INVALID, "link_stats needs link buffers, link_corr, workspace, out");
  for (int q = 0; q < 19 * FSI_LINK_STATS_VALUES; ++q) out[q] = 0.0;
  by_storage(p->storage, [&](auto tag) {
    using Pop = decltype(tag);
    for (int s = 1; s < 19; ++s)
      for (long long y = 0; y < ncells(p); ++y)
        link_stats_cell<Pop>(*p, *st, cur, parity, s, (uint32_t)y, out + s * FSI_LINK_STATS_VALUES);
  });
  return FSI_OK;
}

int fsi_lbm_moments(const FsiLbmParams* p, const FsiLbmState* st, int cur, uint32_t flags, void*) {
  by_storage(p->storage, [&](auto tag) {
    using Pop = decltype(tag);
    for (long long c = 0; c < ncells(p); ++c) lbm_moments_cell<Pop>(*p, *st, cur, flags, c);
  });
  return FSI_OK;
}

// the moving-boundary counters (the open-boundary ones run on)
static void clear_mb_counters(FsiLbmCounters* k) {
  k->bb_link_count = 0;
  k->bb_max_correction_bits = 0;
  k->mb_sample_count = 0;
  k->accum_norm_max_bits = 0;
  k->mean_norm_max_bits = 0;
  k->bb_net_solid_force[0] = k->bb_net_solid_force[1] = k->bb_net_solid_force[2] = 0.0;
}

int fsi_lbm_mb_begin(const FsiLbmParams* p, const FsiLbmState* st, void*) {
  clear_mb_counters(st->counters);
  for (long long c = 0; c < ncells(p); ++c) lbm_mb_begin_cell(*p, *st, c);
  return FSI_OK;
}

int fsi_lbm_mb_finalize(const FsiLbmParams* p, const FsiLbmState* st, int cur, int parity, int n_samples,
                        int subcycled, void*) {
  FsiLbmCounters* k = st->counters;
  k->bb_link_count = 0;
  k->bb_max_correction_bits = 0;
  for (int d = 0; d < 3; ++d) k->bb_net_solid_force[d] = 0.0;
  for (long long c = 0; c < ncells(p); ++c) {
    FinalizeOut o;
    by_storage(p->storage, [&](auto tag) {
      using Pop = decltype(tag);
      o = lbm_finalize_cell<Pop>(*p, *st, cur, parity, n_samples, subcycled, c);
    });
    k->bb_link_count += o.links;
    atomic_max_nonneg(&k->bb_max_correction_bits, o.max_corr);
    if (subcycled) {
      atomic_max_nonneg(&k->mean_norm_max_bits, o.mean_norm);
      atomic_max_nonneg(&k->accum_norm_max_bits, o.runmax);
    }
    k->bb_net_solid_force[0] += o.hx;
    k->bb_net_solid_force[1] += o.hy;
    k->bb_net_solid_force[2] += o.hz;
  }
  if (subcycled) k->mb_sample_count = n_samples;
  return FSI_OK;
}

int fsi_lbm_reduce_stats(const FsiLbmParams* p, const FsiLbmState* st, FsiLbmStats*, FsiLbmStats* out, void*) {
  double* a = reinterpret_cast<double*>(out);
  for (int f = 0; f < LBM_STATS_N; ++f) a[f] = red_identity(lbm_stats_op(f));
  for (long long c = 0; c < ncells(p); ++c) lbm_stats_cell(*p, *st, c, a);
  return FSI_OK;
}

int fsi_project(const FsiLbmParams* lp, const FsiLbmState* ls, const FsiMpmParams* mp, const FsiMpmState* ms,
                const FsiCouplingParams* cp, FsiCouplingCounters* cc, void*) {
  const long long N = ncells(lp);
  for (long long c = 0; c < N; ++c) {
    ls->solid_phi[c] = 0.0f;
    ls->solid_mass[c] = 0.0f;
    ls->solid_vel[c] = ls->solid_vel[N + c] = ls->solid_vel[2 * N + c] = 0.0f;
  }
  cc->projected_mass = cc->projected_volume_raw = cc->projected_volume_clamped = 0.0;
  cc->max_phi_raw_bits = 0;
  cc->active_cell_count = 0;
  const GlobalSink sink = proj_sink(*ls, ncells(lp));
  for (long long q = 0; q < mp->n_particles; ++q) project_particle(*lp, *ls, *mp, *ms, *cp, q, sink);
  for (long long c = 0; c < N; ++c) {
    ProjOut o = normalize_cell(*lp, *ls, *cp, c);
    atomic_max_nonneg(&cc->max_phi_raw_bits, o.phi_raw > 0.0f ? o.phi_raw : 0.0f);
    cc->active_cell_count += o.active;
    cc->projected_mass += o.mass;
    cc->projected_volume_raw += (double)o.phi_raw * cp->cell_volume_norm;
    cc->projected_volume_clamped += (double)o.phi_clamped * cp->cell_volume_norm;
  }
  return FSI_OK;
}

int fsi_penalty_force(const FsiLbmParams* lp, const FsiLbmState* ls, const FsiCouplingParams* cp,
                      FsiCouplingCounters* cc, void*) {
  cc->active_force_cell_count = 0;
  cc->max_cell_force_bits = cc->max_hydro_force_bits = 0;
  for (int d = 0; d < 3; ++d) cc->net_cell_force[d] = cc->net_hydro_force[d] = 0.0;
  for (long long c = 0; c < ncells(lp); ++c) {
    float f[3];
    int act;
    penalty_cell(*lp, *ls, *cp, c, f, act);
    cc->active_force_cell_count += act;
    const float n = norm3(f[0], f[1], f[2]);
    atomic_max_nonneg(&cc->max_cell_force_bits, n);
    atomic_max_nonneg(&cc->max_hydro_force_bits, n);
    for (int d = 0; d < 3; ++d) {
      cc->net_cell_force[d] += f[d];
      cc->net_hydro_force[d] -= f[d];
    }
  }
  return FSI_OK;
}

int fsi_mpm_clear_grid(const FsiMpmParams* mp, const FsiMpmState* ms, void*) {
  const long long G = (long long)mp->gnx * mp->gny * mp->gnz;
  memset(ms->grid_v, 0, sizeof(float) * 3 * G);
  memset(ms->grid_m, 0, sizeof(float) * G);
  memset(ms->grid_f, 0, sizeof(float) * 3 * G);
  return FSI_OK;
}

int fsi_mpm_p2g(const FsiMpmParams* mp, const FsiMpmState* ms, void*) {
  if (mp->gnx < 3 || mp->gny < 3 || mp->gnz < 3)
    return fail(FSI_E_INVALID, "MPM grid must be at least 3 nodes along every axis");
  const GlobalSink sink = p2g_sink(*ms, (long long)mp->gnx * mp->gny * mp->gnz);
  for (long long q = 0; q < mp->n_particles; ++q) mpm_p2g_particle(*mp, *ms, q, sink);
  return FSI_OK;
}

int fsi_mpm_reaction(const FsiLbmParams* lp, const FsiLbmState* ls, const FsiMpmParams* mp, const FsiMpmState* ms,
                     const FsiCouplingParams* cp, FsiCouplingCounters* cc, int mode, void*) {
  if (mode != 1 && mode != 2) return fail(FSI_E_INVALID, "reaction mode must be 1 (penalty) or 2 (moving boundary)");
  cc->active_reaction_particle_count = 0;
  cc->max_particle_reaction_bits = cc->max_grid_reaction_bits = 0;
  for (int d = 0; d < 3; ++d) cc->net_particle_reaction[d] = cc->net_grid_reaction[d] = 0.0;
  for (long long q = 0; q < mp->n_particles; ++q) {
    ReactOut o = reaction_particle(*lp, *ls, *mp, *ms, *cp, mode, q, reaction_sink(*ms, (long long)mp->gnx * mp->gny * mp->gnz));
    cc->active_reaction_particle_count += o.active;
    atomic_max_nonneg(&cc->max_particle_reaction_bits, o.particle_norm);
    atomic_max_nonneg(&cc->max_grid_reaction_bits, o.max_contrib);
    for (int d = 0; d < 3; ++d) {
      cc->net_particle_reaction[d] += o.f[d];
      cc->net_grid_reaction[d] += o.g[d];
    }
  }
  return FSI_OK;
}

int fsi_mpm_substep_coupled(const FsiLbmParams* lp, const FsiLbmState* ls, const FsiMpmParams* mp,
                            const FsiMpmState* ms, const FsiCouplingParams* cp, FsiCouplingCounters* cc, int mode,
                            double* workspace, void* stream) {
  if (mode < 0 || mode > 2) return fail(FSI_E_INVALID, "mode must be 0 (none), 1 (penalty) or 2 (moving boundary)");
  if (mode && (!lp || !ls || !cp || !cc || !workspace))
    return fail(FSI_E_INVALID, "the reaction needs coupling params, counters, workspace, hydro_force, solid_phi");
  int rc = fsi_mpm_clear_grid(mp, ms, stream);
  if (!rc) rc = fsi_mpm_p2g(mp, ms, stream);
  if (!rc && mode) rc = fsi_mpm_reaction(lp, ls, mp, ms, cp, cc, mode, stream);
  if (!rc) rc = fsi_mpm_grid_update(mp, ms, stream);
  if (!rc) rc = fsi_mpm_g2p(mp, ms, stream);
  return rc;
}

int fsi_mpm_substep_coupled_noalloc(const FsiLbmParams* lp, const FsiLbmState* ls, const FsiMpmParams* mp,
                                    const FsiMpmState* ms, const FsiCouplingParams* cp, FsiCouplingCounters* cc,
                                    int mode, void* stream) {
  if (!mp) return fail(FSI_E_INVALID, "null MPM params");
  if (mp->n_particles > FSI_NOALLOC_MAX_PARTICLES)
    return fail(FSI_E_INVALID, "fsi_mpm_substep_coupled_noalloc: particles exceed its scratch limit");
  static std::vector<double> scratch((size_t)FSI_COUPLED_WORKSPACE_DOUBLES(FSI_NOALLOC_MAX_PARTICLES));
  return fsi_mpm_substep_coupled(lp, ls, mp, ms, cp, cc, mode, mode ? scratch.data() : nullptr, stream);
}

// The resident MPM subcycle as the kernel runs it, phase by phase (every
// workgroup's scatter, the barrier, every workgroup's gather): substep s
// scatters into scratch buffer s mod 3, each particle block clears the nodes
// of its stencil box of substep s - 2 in buffer (s + 1) mod 3, and g2p reads
// node velocities computed from the buffer.  The reaction goes to grid_f (the
// emulation's p2g and reaction are separate passes).  At the end the state's
// grid takes the last buffer's updated velocities and masses, and the
// scratch is left zero (the kernel's workgroups clear their last two boxes).
int fsi_mpm_subcycle_coupled(const FsiLbmParams* lp, const FsiLbmState* ls, const FsiMpmParams* mp,
                             const FsiMpmState* ms, const FsiCouplingParams* cp, FsiCouplingCounters* cc, int mode,
                             int n, double* workspace, float* scratch, uint32_t* sync, void* stream) {
  if (!mp || !ms) return fail(FSI_E_INVALID, "null MPM params/state");
  if (mode < 0 || mode > 2) return fail(FSI_E_INVALID, "mode must be 0 (none), 1 (penalty) or 2 (moving boundary)");
  if (n < 1) return fail(FSI_E_INVALID, "n must be at least 1");
  if (!scratch || !sync) return fail(FSI_E_INVALID, "the resident MPM subcycle needs scratch and sync");
  if (mode && (!lp || !ls || !cp || !cc || !workspace))
    return fail(FSI_E_INVALID, "the reaction needs coupling params, counters, workspace, hydro_force, solid_phi");
  if (mode && cp->inv_dx_norm != mp->inv_dx)
    return fail(FSI_E_UNSUPPORTED, "the reaction samples the lattice at another spacing than the MPM grid");
  if (mp->n_particles <= 0) return fail(FSI_E_UNSUPPORTED, "no particles");
  if (mp->gnx < 3 || mp->gny < 3 || mp->gnz < 3)
    return fail(FSI_E_INVALID, "MPM grid must be at least 3 nodes along every axis");
  const long long G = (long long)mp->gnx * mp->gny * mp->gnz, P = mp->n_particles;
  const long long nb = (P + FSI_PARTICLE_BLOCK - 1) / FSI_PARTICLE_BLOCK;
  auto buf = [&](int k, FsiMpmState& b) {
    b = *ms;
    b.grid_v = scratch + (size_t)k * 4 * G;
    b.grid_m = b.grid_v + 3 * G;
  };
  for (long long i = 0; i < 12 * G; ++i)
    if (scratch[i] != 0.0f) return fail(FSI_E_INVALID, "the resident MPM scratch is not zero on entry");
  std::vector<int> box(2 * 6 * nb, 0);  // per block: [s - 1] lo, hi; [s - 2] lo, hi
  for (long long b = 0; b < nb; ++b)
    for (int k = 0; k < 2; ++k)
      for (int d = 0; d < 3; ++d) {
        box[(b * 2 + k) * 6 + d] = 0;
        box[(b * 2 + k) * 6 + 3 + d] = -1;
      }
  std::vector<float> vel(3 * (size_t)G);
  for (int s = 0; s < n; ++s) {
    FsiMpmState cur;
    buf(s % 3, cur);
    memset(ms->grid_f, 0, sizeof(float) * 3 * G);
    int rc = fsi_mpm_p2g(mp, &cur, stream);
    if (!rc && mode) rc = fsi_mpm_reaction(lp, ls, mp, &cur, cp, cc, mode, stream);
    if (rc) return rc;
    FsiMpmState old;
    buf((s + 1) % 3, old);
    for (long long b = 0; b < nb; ++b) {
      int* bx = &box[b * 12];
      if (s >= 2) {
        const int* lo = bx + 6;
        const int* hi = bx + 9;
        for (int I = lo[0] < 0 ? 0 : lo[0]; I <= hi[0] && I < mp->gnx; ++I)
          for (int J = lo[1] < 0 ? 0 : lo[1]; J <= hi[1] && J < mp->gny; ++J)
            for (int K = lo[2] < 0 ? 0 : lo[2]; K <= hi[2] && K < mp->gnz; ++K) {
              const long long nn = ((long long)I * mp->gny + J) * mp->gnz + K;
              for (int d = 0; d < 3; ++d) old.grid_v[d * G + nn] = 0.0f;
              old.grid_m[nn] = 0.0f;
            }
      }
      for (int e = 0; e < 6; ++e) bx[6 + e] = bx[e];
      int lo[3] = {0x7FFFFFFF, 0x7FFFFFFF, 0x7FFFFFFF}, hi[3] = {-0x7FFFFFFF, -0x7FFFFFFF, -0x7FFFFFFF};
      for (long long q = b * FSI_PARTICLE_BLOCK; q < P && q < (b + 1) * FSI_PARTICLE_BLOCK; ++q) {
        const Stencil sn = particle_stencil(ms->x, P, q, mp->inv_dx);
        for (int d = 0; d < 3; ++d) {
          lo[d] = sn.base[d] < lo[d] ? sn.base[d] : lo[d];
          hi[d] = sn.base[d] + 2 > hi[d] ? sn.base[d] + 2 : hi[d];
        }
      }
      for (int d = 0; d < 3; ++d) {
        bx[d] = lo[d];
        bx[3 + d] = hi[d];
      }
    }
    // the gather's node velocities
    for (long long nn = 0; nn < G; ++nn) {
      const float m = cur.grid_m[nn];
      const float mom[3] = {cur.grid_v[nn], cur.grid_v[G + nn], cur.grid_v[2 * G + nn]};
      float v[3] = {mom[0], mom[1], mom[2]};
      if (m > 0.0f) {
        const int K = (int)(nn % mp->gnz), J = (int)((nn / mp->gnz) % mp->gny),
                  I = (int)(nn / ((long long)mp->gnz * mp->gny));
        const float f[3] = {ms->grid_f[nn], ms->grid_f[G + nn], ms->grid_f[2 * G + nn]};
        grid_node_velocity(*mp, m, mom, f, I, J, K, v);
      }
      for (int d = 0; d < 3; ++d) vel[d * G + nn] = v[d];
    }
    FsiMpmState gat = *ms;
    gat.grid_v = vel.data();
    rc = fsi_mpm_g2p(mp, &gat, stream);
    if (rc) return rc;
  }
  // the state's grid: the last buffer's updated velocities and masses
  FsiMpmState last;
  buf((n - 1) % 3, last);
  memcpy(ms->grid_v, last.grid_v, sizeof(float) * 3 * G);
  memcpy(ms->grid_m, last.grid_m, sizeof(float) * G);
  for (long long nn = 0; nn < G; ++nn) mpm_grid_node(*mp, *ms, nn);
  memset(scratch, 0, sizeof(float) * 12 * G);
  return FSI_OK;
}

// Binning, serially: keys in particle order, counts, exclusive scan, and
// slots taken in particle order (so each bin's list is already increasing).
static void bin_serial(const float* x, int32_t P, float inv, const BinGrid& bg, const FsiBins& b) {
  const long long n = bg.nb + 1;
  int32_t* count = b.work;
  for (long long i = 0; i < n; ++i) count[i] = 0;
  for (int32_t p = 0; p < P; ++p) {
    b.key[p] = particle_bin(bg, x, P, p, inv);
    ++count[b.key[p]];
  }
  int32_t run = 0;
  for (long long i = 0; i < n; ++i) {
    b.start[i] = run;
    run += count[i];
  }
  b.start[n] = run;
  for (long long i = 0; i < n; ++i) count[i] = 0;
  for (int32_t p = 0; p < P; ++p) b.order[b.start[b.key[p]] + count[b.key[p]]++] = p;
  for (long long i = 0; i < n; ++i) count[i] = 0;
}

int fsi_bin_particles(const float* x, int32_t n_particles, float inv_dx, int32_t nx, int32_t ny, int32_t nz,
                      const FsiBins* bins, void*) {
  if (nx < 1 || ny < 1 || nz < 1 || n_particles < 0 || !bins)
    return fail(FSI_E_INVALID, "fsi_bin_particles: bad grid, particles or buffers");
  bin_serial(x, n_particles, inv_dx, make_bins(nx, ny, nz), *bins);
  return FSI_OK;
}

int fsi_mpm_substep_ordered(const FsiLbmParams* lp, const FsiLbmState* ls, const FsiMpmParams* mp,
                            const FsiMpmState* ms, const FsiCouplingParams* cp, FsiCouplingCounters* cc, int mode,
                            const FsiBins* bins, float* terms, double* workspace, void*) {
  if (mode < 0 || mode > 2) return fail(FSI_E_INVALID, "mode must be 0 (none), 1 (penalty) or 2 (moving boundary)");
  if (!bins || (mp->n_particles > 0 && !terms)) return fail(FSI_E_INVALID, "null binning or terms buffers");
  if (mode && (!lp || !ls || !cp || !cc || !workspace))
    return fail(FSI_E_INVALID, "the reaction needs coupling params, counters, workspace, hydro_force, solid_phi");
  if (mode && cp->inv_dx_norm != mp->inv_dx)
    return fail(FSI_E_INVALID, "fsi_mpm_substep_ordered: the lattice spacing must be the MPM grid's");
  FsiLbmParams lpv;
  FsiLbmState lsv;
  FsiCouplingParams cpv;
  memset(&lpv, 0, sizeof(lpv));
  memset(&lsv, 0, sizeof(lsv));
  memset(&cpv, 0, sizeof(cpv));
  if (mode) {
    lpv = *lp;
    lsv = *ls;
    cpv = *cp;
    cc->active_reaction_particle_count = 0;
    cc->max_particle_reaction_bits = cc->max_grid_reaction_bits = 0;
    for (int d = 0; d < 3; ++d) cc->net_particle_reaction[d] = cc->net_grid_reaction[d] = 0.0;
  }
  for (long long q = 0; q < mp->n_particles; ++q) {
    const ReactOut o = mpm_ordered_terms_particle(lpv, lsv, *mp, *ms, cpv, mode, q, terms);
    if (mode) {
      cc->active_reaction_particle_count += o.active;
      atomic_max_nonneg(&cc->max_particle_reaction_bits, o.particle_norm);
      atomic_max_nonneg(&cc->max_grid_reaction_bits, o.max_contrib);
      for (int d = 0; d < 3; ++d) {
        cc->net_particle_reaction[d] += o.f[d];
        cc->net_grid_reaction[d] += o.g[d];
      }
    }
  }
  const BinGrid bg = make_bins(mp->gnx, mp->gny, mp->gnz);
  bin_serial(ms->x, mp->n_particles, mp->inv_dx, bg, *bins);
  const long long G = (long long)mp->gnx * mp->gny * mp->gnz;
  for (long long n = 0; n < G; ++n)
    mpm_ordered_node(*mp, *ms, bg, bins->start, bins->order, terms, mode != 0 && mp->n_particles > 0, n);
  for (long long q = 0; q < mp->n_particles; ++q) mpm_g2p_particle(*mp, *ms, q);
  return FSI_OK;
}

int fsi_project_ordered(const FsiLbmParams* lp, const FsiLbmState* ls, const FsiMpmParams* mp,
                        const FsiMpmState* ms, const FsiCouplingParams* cp, FsiCouplingCounters* cc,
                        const FsiBins* bins, void*) {
  if (!bins) return fail(FSI_E_INVALID, "null binning buffers");
  const BinGrid bg = make_bins(lp->nx, lp->ny, lp->nz);
  bin_serial(ms->x, mp->n_particles, cp->inv_dx_norm, bg, *bins);
  const long long N = ncells(lp);
  for (long long c = 0; c < N; ++c) project_ordered_cell(*lp, *ls, *mp, *ms, *cp, bg, bins->start, bins->order, c);
  cc->projected_mass = cc->projected_volume_raw = cc->projected_volume_clamped = 0.0;
  cc->max_phi_raw_bits = 0;
  cc->active_cell_count = 0;
  for (long long c = 0; c < N; ++c) {
    ProjOut o = normalize_cell(*lp, *ls, *cp, c);
    atomic_max_nonneg(&cc->max_phi_raw_bits, o.phi_raw > 0.0f ? o.phi_raw : 0.0f);
    cc->active_cell_count += o.active;
    cc->projected_mass += o.mass;
    cc->projected_volume_raw += (double)o.phi_raw * cp->cell_volume_norm;
    cc->projected_volume_clamped += (double)o.phi_clamped * cp->cell_volume_norm;
  }
  return FSI_OK;
}

int fsi_mpm_grid_update(const FsiMpmParams* mp, const FsiMpmState* ms, void*) {
  const long long G = (long long)mp->gnx * mp->gny * mp->gnz;
  for (long long n = 0; n < G; ++n) mpm_grid_node(*mp, *ms, n);
  return FSI_OK;
}

int fsi_mpm_g2p(const FsiMpmParams* mp, const FsiMpmState* ms, void*) {
  if (mp->gnx < 3 || mp->gny < 3 || mp->gnz < 3)
    return fail(FSI_E_INVALID, "MPM grid must be at least 3 nodes along every axis");
  for (long long q = 0; q < mp->n_particles; ++q) mpm_g2p_particle(*mp, *ms, q);
  return FSI_OK;
}

int fsi_mpm_reduce_stats(const FsiMpmParams* mp, const FsiMpmState* ms, FsiMpmStats*, FsiMpmStats* out, void*) {
  double* a = reinterpret_cast<double*>(out);
  for (int f = 0; f < MPM_STATS_N; ++f) a[f] = red_identity(mpm_stats_op(f));
  for (long long q = 0; q < mp->n_particles; ++q) mpm_stats_particle(*mp, *ms, q, a);
  return FSI_OK;
}


int fsi_batch_block_map(const FsiBatchInstance* inst, int n_inst, int domain, int mode_mask, int min_mpm,
                        int32_t* map) {
  if (!inst || n_inst < 0) return fail(FSI_E_INVALID, "null batch arguments");
  if (domain < FSI_BATCH_DOMAIN_CELLS || domain > FSI_BATCH_DOMAIN_BOX) return fail(FSI_E_INVALID, "unknown batch domain");
  long long total = 0;
  for (int i = 0; i < n_inst; ++i) {
    if (inst[i].lp.open_bc != FSI_OPEN_BC_EQUILIBRIUM)
      return fail(FSI_E_INVALID, "the batched form runs the legacy x-face reset only (open_bc 0)");
    if (inst[i].active)
      if (const char* why = batch_resident_invalid(inst[i])) return fail(FSI_E_INVALID, why);
    if (!inst[i].active || !(mode_mask & (1 << inst[i].mode)) || inst[i].mpm_substeps <= min_mpm) continue;
    const long long nb = batch_domain_blocks(inst[i], domain);
    if (map)
      for (long long b = 0; b < nb; ++b) {
        map[2 * (total + b)] = i;
        map[2 * (total + b) + 1] = (int32_t)b;
      }
    total += nb;
  }
  return (int)total;
}

// Box tests of the batched passes (FSI_BATCH_BOX_*), as the kernels make them.
static bool cell_in(const int32_t* r, long long e, int ny, int nz) {
  const long long plane = (long long)ny * nz, i = e / plane, rem = e - i * plane;
  return in_box(r, (int)i, (int)(rem / nz), (int)(rem % nz));
}

// Grows box record b's `cur` over particle q's stencil at spacing 1 / inv.
static void grow_box(int32_t* b, const float* x, long long P, long long q, float inv) {
  for (int d = 0; d < 3; ++d) {
    int base;
    float fx, w[3];
    bspline(x[d * P + q] * inv, base, fx, w);
    b[BOX_CUR + d] = base < b[BOX_CUR + d] ? base : b[BOX_CUR + d];
    b[BOX_CUR + 3 + d] = base + 2 > b[BOX_CUR + 3 + d] ? base + 2 : b[BOX_CUR + 3 + d];
  }
}

// Serial emulation of fsi_batch_launch: instance by instance (the instances
// named in the block map, in index order), work item by work item.
int fsi_batch_launch(int op, const FsiBatchInstance* inst, int n_inst, const int32_t* map, int n_blocks, int t, void*) {
  if (!map) return fail(FSI_E_INVALID, "null block map");
  if (op == FSI_BATCH_OP_COUNTERS_CLEAR) {  // the instances of the map (INSTANCES domain)
    for (int b = 0; b < n_blocks; ++b) batch_clear_counters(inst[map[2 * b]], t);
    return FSI_OK;
  }
  if (op == FSI_BATCH_OP_LBM_RESIDENT) {  // the one instance of the map, as fsi_lbm_subcycle_resident
    if (n_blocks <= 0) return FSI_OK;
    const FsiBatchInstance& I = inst[map[0]];
    return fsi_lbm_subcycle_resident(&I.lp, &I.ls, I.cur0, I.n_substeps, BATCH_RESIDENT_FLAGS, I.res_halo, I.res_sync,
                                     nullptr);
  }
  std::vector<char> sel((size_t)n_inst, 0);
  for (int b = 0; b < n_blocks; ++b) sel[(size_t)map[2 * b]] = 1;
  for (int i = 0; i < n_inst; ++i) {
    if (!sel[(size_t)i]) continue;
    const FsiBatchInstance& I = inst[i];
    const long long cells = batch_domain_size(I, FSI_BATCH_DOMAIN_CELLS);
    const long long nodes = batch_domain_size(I, FSI_BATCH_DOMAIN_NODES);
    const long long parts = batch_domain_size(I, FSI_BATCH_DOMAIN_PARTICLES);
    switch (op) {
      case FSI_BATCH_OP_LBM_SUBSTEP:
      case FSI_BATCH_OP_LBM_SUBSTEP_MB:
      case FSI_BATCH_OP_LBM_SUBSTEP_PENALTY:
      case FSI_BATCH_OP_LBM_SUBSTEP_NONE:
        by_storage(I.lp.storage, [&](auto tag) {
          using Pop = decltype(tag);
          (void)cells;
          for (int a = 0; a < I.lp.nx; ++a)
            for (int b = 0; b < I.lp.ny; ++b)
              for (int k = 0; k < I.lp.nz; ++k)
                {
                  uint32_t fl = batch_substep_flags(I, t), ofl;
                  batch_moment_source(I, t, fl, ofl, a, b, k);
                  lbm_substep_cell<Pop>(I.lp, I.ls, batch_cur(I, t), t & 1, fl, a, b, k, ofl);
                }
        });
        break;
      case FSI_BATCH_OP_LBM_REBUILD:
        by_storage(I.lp.storage, [&](auto tag) {
          using Pop = decltype(tag);
          for (long long c = 0; c < cells; ++c)
            lbm_rebuild_cell<Pop>(I.lp, I.ls, I.mode == FSI_BATCH_MODE_MOVING_BOUNDARY ? 1 : 0, I.cur0, c);
        });
        break;
      case FSI_BATCH_OP_LBM_MOMENTS:
        by_storage(I.lp.storage, [&](auto tag) {
          using Pop = decltype(tag);
          for (long long c = 0; c < cells; ++c)
            lbm_moments_cell<Pop>(I.lp, I.ls, batch_cur(I, t + 1), I.cell_force ? FSI_SUB_CELL_FORCE : 0u, c);
        });
        break;
      case FSI_BATCH_OP_LBM_MB_BEGIN:
        clear_mb_counters(I.ls.counters);
        for (long long c = 0; c < cells; ++c) lbm_mb_begin_cell(I.lp, I.ls, c);
        break;
      case FSI_BATCH_OP_LBM_MB_FINALIZE: {
        FsiLbmCounters* k = I.ls.counters;
        for (long long c = 0; c < cells; ++c) {
          FinalizeOut o;
          by_storage(I.lp.storage, [&](auto tag) {
            using Pop = decltype(tag);
            o = lbm_finalize_cell<Pop>(I.lp, I.ls, batch_cur(I, I.n_substeps), (I.n_substeps - 1) & 1, I.n_substeps,
                                       1, c);
          });
          k->bb_link_count += o.links;
          atomic_max_nonneg(&k->bb_max_correction_bits, o.max_corr);
          atomic_max_nonneg(&k->mean_norm_max_bits, o.mean_norm);
          atomic_max_nonneg(&k->accum_norm_max_bits, o.runmax);
          k->bb_net_solid_force[0] += o.hx;
          k->bb_net_solid_force[1] += o.hy;
          k->bb_net_solid_force[2] += o.hz;
        }
        k->mb_sample_count = I.n_substeps;
        break;
      }
      case FSI_BATCH_OP_LBM_MB_FINALIZE_FOLD:  // the serial finalize above sets the counters directly
        break;
      case FSI_BATCH_OP_BOX_ROLL:
        if (t == 0)
          box_roll(I.boxes + FSI_BATCH_BOX_PROJ, I.lp.nx, I.lp.ny, I.lp.nz);
        else
          box_roll(I.boxes + FSI_BATCH_BOX_MPM, I.mp.gnx, I.mp.gny, I.mp.gnz);
        break;
      case FSI_BATCH_OP_PROJ_CLEAR:
        for (long long c = 0; c < cells; ++c) {
          if (!cell_in(I.boxes + FSI_BATCH_BOX_PROJ + BOX_PREV, c, I.lp.ny, I.lp.nz)) continue;
          I.ls.solid_phi[c] = 0.0f;
          I.ls.solid_mass[c] = 0.0f;
          I.ls.solid_vel[c] = I.ls.solid_vel[cells + c] = I.ls.solid_vel[2 * cells + c] = 0.0f;
        }
        break;
      case FSI_BATCH_OP_PROJ_PARTICLES:
        for (long long q = 0; q < parts; ++q) {
          grow_box(I.boxes + FSI_BATCH_BOX_PROJ, I.ms.x, parts, q, I.cp.inv_dx_norm);
          project_particle(I.lp, I.ls, I.mp, I.ms, I.cp, q, proj_sink(I.ls, cells));
        }
        break;
      case FSI_BATCH_OP_PROJ_NORMALIZE:
        for (long long c = 0; c < cells; ++c) {
          if (!cell_in(I.boxes + FSI_BATCH_BOX_PROJ + BOX_CUR, c, I.lp.ny, I.lp.nz)) continue;
          ProjOut o = normalize_cell(I.lp, I.ls, I.cp, c);
          atomic_max_nonneg(&I.cc->max_phi_raw_bits, o.phi_raw > 0.0f ? o.phi_raw : 0.0f);
          I.cc->active_cell_count += o.active;
          I.cc->projected_mass += o.mass;
          I.cc->projected_volume_raw += (double)o.phi_raw * I.cp.cell_volume_norm;
          I.cc->projected_volume_clamped += (double)o.phi_clamped * I.cp.cell_volume_norm;
        }
        break;
      case FSI_BATCH_OP_PENALTY: {
        int32_t both[6];
        box_hull(I.boxes + FSI_BATCH_BOX_PROJ + BOX_CUR, I.boxes + FSI_BATCH_BOX_PROJ + BOX_PREV, both);
        if (t > 0 && I.lp.bc_x_right == FSI_BC_PRESSURE) {  // the plane the pressure outlet reads
          const long long plane = (long long)I.lp.ny * I.lp.nz, i0 = (long long)(I.lp.nx - 2) * plane;
          for (long long b = 0; b < plane; ++b)
            if (!batch_penalty_stored(I, I.lp.nx - 2, (int)(b / I.lp.nz), (int)(b % I.lp.nz)))
              by_storage(I.lp.storage, [&](auto tag) {
                using Pop = decltype(tag);
                lbm_moments_cell<Pop>(I.lp, I.ls, batch_cur(I, t), FSI_SUB_CELL_FORCE, i0 + b);
              });
        }
        for (long long c = 0; c < cells; ++c) {
          if (!cell_in(both, c, I.lp.ny, I.lp.nz)) continue;
          if (t > 0)
            by_storage(I.lp.storage, [&](auto tag) {
              using Pop = decltype(tag);
              lbm_moments_cell<Pop>(I.lp, I.ls, batch_cur(I, t), FSI_SUB_CELL_FORCE, c);
            });
          float f[3];
          int act;
          penalty_cell(I.lp, I.ls, I.cp, c, f, act);
          I.cc->active_force_cell_count += act;
          const float n = norm3(f[0], f[1], f[2]);
          atomic_max_nonneg(&I.cc->max_cell_force_bits, n);
          atomic_max_nonneg(&I.cc->max_hydro_force_bits, n);
          for (int d = 0; d < 3; ++d) {
            I.cc->net_cell_force[d] += f[d];
            I.cc->net_hydro_force[d] -= f[d];
          }
        }
        break;
      }
      case FSI_BATCH_OP_MPM_CLEAR:
        for (long long n = 0; n < nodes; ++n) {
          if (!cell_in(I.boxes + FSI_BATCH_BOX_MPM + BOX_PREV, n, I.mp.gny, I.mp.gnz)) continue;
          I.ms.grid_v[n] = I.ms.grid_v[nodes + n] = I.ms.grid_v[2 * nodes + n] = 0.0f;
          I.ms.grid_m[n] = 0.0f;
          I.ms.grid_f[n] = I.ms.grid_f[nodes + n] = I.ms.grid_f[2 * nodes + n] = 0.0f;
        }
        break;
      case FSI_BATCH_OP_MPM_P2G:
        for (long long q = 0; q < parts; ++q) {
          grow_box(I.boxes + FSI_BATCH_BOX_MPM, I.ms.x, parts, q, I.mp.inv_dx);
          mpm_p2g_particle(I.mp, I.ms, q, p2g_sink(I.ms, nodes));
        }
        break;
      case FSI_BATCH_OP_MPM_REACTION:
        if (I.mode == FSI_BATCH_MODE_NONE) break;
        for (long long q = 0; q < parts; ++q) {
          grow_box(I.boxes + FSI_BATCH_BOX_MPM, I.ms.x, parts, q, I.cp.inv_dx_norm);
          ReactOut o = reaction_particle(I.lp, I.ls, I.mp, I.ms, I.cp, I.mode == FSI_BATCH_MODE_MOVING_BOUNDARY ? 2 : 1, q,
                                          reaction_sink(I.ms, nodes));
          I.cc->active_reaction_particle_count += o.active;
          atomic_max_nonneg(&I.cc->max_particle_reaction_bits, o.particle_norm);
          atomic_max_nonneg(&I.cc->max_grid_reaction_bits, o.max_contrib);
          for (int d = 0; d < 3; ++d) {
            I.cc->net_particle_reaction[d] += o.f[d];
            I.cc->net_grid_reaction[d] += o.g[d];
          }
        }
        break;
      case FSI_BATCH_OP_MPM_P2G_REACT:
        for (long long q = 0; q < parts; ++q) {
          grow_box(I.boxes + FSI_BATCH_BOX_MPM, I.ms.x, parts, q, I.mp.inv_dx);
          if (I.mode != FSI_BATCH_MODE_NONE && I.cp.inv_dx_norm != I.mp.inv_dx)
            grow_box(I.boxes + FSI_BATCH_BOX_MPM, I.ms.x, parts, q, I.cp.inv_dx_norm);
          mpm_p2g_particle(I.mp, I.ms, q, p2g_sink(I.ms, nodes));
        }
        if (I.mode == FSI_BATCH_MODE_NONE) break;
        I.cc->active_reaction_particle_count = 0;
        I.cc->max_particle_reaction_bits = I.cc->max_grid_reaction_bits = 0;
        for (int d = 0; d < 3; ++d) I.cc->net_particle_reaction[d] = I.cc->net_grid_reaction[d] = 0.0;
        for (long long q = 0; q < parts; ++q) {
          ReactOut o = reaction_particle(I.lp, I.ls, I.mp, I.ms, I.cp, I.mode == FSI_BATCH_MODE_MOVING_BOUNDARY ? 2 : 1, q,
                                          reaction_sink(I.ms, nodes));
          I.cc->active_reaction_particle_count += o.active;
          atomic_max_nonneg(&I.cc->max_particle_reaction_bits, o.particle_norm);
          atomic_max_nonneg(&I.cc->max_grid_reaction_bits, o.max_contrib);
          for (int d = 0; d < 3; ++d) {
            I.cc->net_particle_reaction[d] += o.f[d];
            I.cc->net_grid_reaction[d] += o.g[d];
          }
        }
        break;
      case FSI_BATCH_OP_MPM_GRID_UPDATE:
      case FSI_BATCH_OP_MPM_GRID_UPDATE_REACT:
        for (long long n = 0; n < nodes; ++n)
          if (cell_in(I.boxes + FSI_BATCH_BOX_MPM + BOX_CUR, n, I.mp.gny, I.mp.gnz)) mpm_grid_node(I.mp, I.ms, n);
        break;
      case FSI_BATCH_OP_MPM_G2P:
        for (long long q = 0; q < parts; ++q) mpm_g2p_particle(I.mp, I.ms, q);
        break;
      default:
        return fail(FSI_E_INVALID, "unknown batch op");
    }
  }
  return FSI_OK;
}

// The batched exchange graph of the emulation: its launch list, run in order.
namespace {
struct EmuBatchGraph {
  const FsiBatchInstance* inst;
  int n_inst;
  std::vector<FsiBatchOp> ops;
};

int check_ops(const FsiBatchOp* ops, int n_ops, const EmuBatchGraph* g) {
  if (g && (size_t)n_ops != g->ops.size()) return fail(FSI_E_INVALID, "batch graph update with another number of operations");
  for (int i = 0; i < n_ops; ++i) {
    if (ops[i].op < FSI_BATCH_OP_LBM_SUBSTEP || ops[i].op > FSI_BATCH_OP_LBM_RESIDENT)
      return fail(FSI_E_INVALID, "unknown batch operation");
    if (!ops[i].block_map || ops[i].n_blocks <= 0)
      return fail(FSI_E_INVALID, "batch graph operation without a block map or workgroups");
    if (ops[i].branch < FSI_BATCH_BRANCH_JOIN || ops[i].branch > FSI_BATCH_SIDE_CHAINS)
      return fail(FSI_E_INVALID, "unknown batch graph branch");
    if (g && (ops[i].op != g->ops[(size_t)i].op || ops[i].t != g->ops[(size_t)i].t ||
              ops[i].branch != g->ops[(size_t)i].branch))
      return fail(FSI_E_INVALID, "batch graph update with other operations");
  }
  return FSI_OK;
}
}  // namespace

int fsi_batch_graph_create(const FsiBatchOp* ops, int n_ops, const FsiBatchInstance* inst, int n_inst,
                           void** graph_out) {
  if (!graph_out) return fail(FSI_E_INVALID, "null graph handle");
  *graph_out = nullptr;
  if (!ops || n_ops <= 0 || !inst || n_inst <= 0) return fail(FSI_E_INVALID, "empty batch graph");
  const int rc = check_ops(ops, n_ops, nullptr);
  if (rc != FSI_OK) return rc;
  *graph_out = new EmuBatchGraph{inst, n_inst, std::vector<FsiBatchOp>(ops, ops + n_ops)};
  return FSI_OK;
}

int fsi_batch_graph_update(void* graph, const FsiBatchOp* ops, int n_ops) {
  EmuBatchGraph* g = static_cast<EmuBatchGraph*>(graph);
  if (!g || !ops) return fail(FSI_E_INVALID, "null batch graph");
  const int rc = check_ops(ops, n_ops, g);
  if (rc != FSI_OK) return rc;
  g->ops.assign(ops, ops + n_ops);
  return FSI_OK;
}

int fsi_batch_graph_launch(void* graph, void* stream) {
  EmuBatchGraph* g = static_cast<EmuBatchGraph*>(graph);
  if (!g) return fail(FSI_E_INVALID, "null batch graph");
  for (const FsiBatchOp& o : g->ops) {
    const int rc = fsi_batch_launch(o.op, g->inst, g->n_inst, o.block_map, o.n_blocks, o.t, stream);
    if (rc != FSI_OK) return rc;
  }
  return FSI_OK;
}

int fsi_batch_graph_destroy(void* graph) {
  delete static_cast<EmuBatchGraph*>(graph);
  return FSI_OK;
}

// Test-only entries of the emulation build: the per-point bodies the kernels
// inline, exposed so the closed-form checks can call them directly.
int fsi_emu_bspline(int n, const float* X, int32_t* base, float* fx, float* w) {
  for (int i = 0; i < n; ++i) {
    int b;
    bspline(X[i], b, fx[i], w + 3 * i);
    base[i] = b;
  }
  return FSI_OK;
}

int fsi_emu_regularized_collide(int n, const float* F, const float* rho, const float* u, const float* g, float sv,
                                float* f) {
  for (int c = 0; c < n; ++c) {
    const bool has_force = g[3 * c] != 0.0f || g[3 * c + 1] != 0.0f || g[3 * c + 2] != 0.0f;
    regularized_collide(F + Q * c, rho[c], u[3 * c], u[3 * c + 1], u[3 * c + 2], has_force, g[3 * c], g[3 * c + 1],
                        g[3 * c + 2], sv, f + Q * c);
  }
  return FSI_OK;
}

}  // extern "C"
