// fsi_kernels.hip - gfx950 kernels and the extern "C" entry points of
// libfsi_hip.so (declared in include/fsi_hip.h).
//
// Launch shapes: one work-item per lattice cell / MPM node / particle, 256-wide
// workgroups (4 wave64s).  The fused LBM substep keeps a cell's 19 populations
// in VGPRs between the coalesced SoA load of F[cur] and the pushes into
// F[1-cur]; consecutive work-items own consecutive k, so each of the 19 push
// streams is a contiguous (shifted) run of addresses.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>
#include <mutex>
#include <string>
#include <vector>

#include "fsi_core.h"

using namespace fsi;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(FSI_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return FSI_OK;
}

constexpr int TPB = 256;

inline unsigned blocks_for(long long n) { return (unsigned)((n + TPB - 1) / TPB); }

// SIMDs of the current device (4 per CU), queried once per device.
inline long long simds_on_device() {
  static thread_local int dev_cached = -1;
  static thread_local long long simds = 0;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 1024;
  if (dev != dev_cached) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
    simds = 4LL * cus;
    dev_cached = dev;
  }
  return simds;
}

// Workgroup size of the per-particle kernels.  Particle counts are small next
// to lattice sizes (tens of thousands), so small workgroups spread them over
// all 256 CUs instead of a few dozen.
constexpr int PTPB = FSI_PARTICLE_BLOCK;
// The particle workgroups are one wave: their tile boxes and reductions are
// wave reductions, and the fused scatter's read-modify-write relies on it.
static_assert(PTPB == 64, "particle workgroups are one wave64");
inline unsigned pblocks_for(long long n) { return (unsigned)((n + PTPB - 1) / PTPB); }

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Run STMT with POP bound to the population accessor of a storage code.
#define FSI_BY_STORAGE(code, STMT)          \
  do {                                      \
    if ((code) == FSI_STORE_F16_DEV) {      \
      using POP = PopF16;                   \
      STMT;                                 \
    } else if ((code) == FSI_STORE_BF16_DEV) { \
      using POP = PopBF16;                  \
      STMT;                                 \
    } else {                                \
      using POP = PopF32;                   \
      STMT;                                 \
    }                                       \
  } while (0)

int check_lbm(const FsiLbmParams* p, const FsiLbmState* st) {
  if (!p || !st) return fail(FSI_E_INVALID, "null params/state");
  if (p->nx < 3 || p->ny < 3 || p->nz < 3)
    return fail(FSI_E_INVALID, "lattice must be at least 3 cells along every axis");
  if ((long long)p->nx * p->ny * p->nz > (1LL << 30))
    return fail(FSI_E_INVALID, "lattice must hold at most 2^30 cells (32-bit cell offsets)");
  if (p->bc_x_left < 0 || p->bc_x_left > 2 || p->bc_x_right < 0 || p->bc_x_right > 2)
    return fail(FSI_E_INVALID, "bc_x_left/bc_x_right must be 0, 1 or 2");
  if (p->storage < FSI_STORE_F32 || p->storage > FSI_STORE_BF16_DEV)
    return fail(FSI_E_INVALID, "storage must be FSI_STORE_F32, FSI_STORE_F16_DEV or FSI_STORE_BF16_DEV");
  if (p->collision != FSI_COLLIDE_MRT && p->collision != FSI_COLLIDE_REGULARIZED)
    return fail(FSI_E_INVALID, "collision must be FSI_COLLIDE_MRT or FSI_COLLIDE_REGULARIZED");
  if (p->open_bc < FSI_OPEN_BC_EQUILIBRIUM || p->open_bc > FSI_OPEN_BC_CONVECTIVE)
    return fail(FSI_E_INVALID, "open_bc must be one of FSI_OPEN_BC_*");
  if (!st->F[0] || !st->F[1] || !st->rho || !st->vel || !st->cellinfo || !st->solid)
    return fail(FSI_E_INVALID, "null LBM buffer");
  return FSI_OK;
}

int check_mpm(const FsiMpmParams* mp, const FsiMpmState* ms) {
  if (!mp || !ms) return fail(FSI_E_INVALID, "null MPM params/state");
  if (mp->gnx < 3 || mp->gny < 3 || mp->gnz < 3)
    return fail(FSI_E_INVALID, "MPM grid must be at least 3 nodes along every axis");
  if (mp->n_particles < 0) return fail(FSI_E_INVALID, "negative particle count");
  if (!ms->x || !ms->v || !ms->C || !ms->F || !ms->Jp || !ms->mass || !ms->vol0 || !ms->grid_v ||
      !ms->grid_m || !ms->grid_f)
    return fail(FSI_E_INVALID, "null MPM buffer");
  return FSI_OK;
}

// Kinds of block_reduce_n beside RED_SUM / RED_MIN / RED_MAX: the maximum
// of floats, at least 0 (block_max_f), and the sum of ints (block_sum_i).
constexpr int RED_MAXF0 = 3, RED_SUMI = 4;

// ---------------------------------------------------------------------------
// Block reductions: a wave64 reduction through DPP lane moves (register to
// register: quad permutes, row mirrors, then the row broadcasts of lanes 15
// and 31; the full result lands in lane 63), then one LDS slot per wave.  A
// lane whose DPP source is masked off takes the identity.
// ---------------------------------------------------------------------------
template <typename T, typename Op>
__device__ __forceinline__ T wave_dpp_reduce(T v, T id, Op op) {
  v = op(v, __builtin_amdgcn_update_dpp(id, v, 0xB1, 0xF, 0xF, false));   // quad_perm [1,0,3,2]
  v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x4E, 0xF, 0xF, false));   // quad_perm [2,3,0,1]
  v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x141, 0xF, 0xF, false));  // row_half_mirror
  v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x140, 0xF, 0xF, false));  // row_mirror
  v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x142, 0xA, 0xF, false));  // row_bcast:15 into rows 1, 3
  v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x143, 0xC, 0xF, false));  // row_bcast:31 into rows 2, 3
  return v;
}

// Lane 63's value, in every lane (uniform).
__device__ __forceinline__ int lane63(int v) { return __builtin_amdgcn_readlane(v, 63); }
__device__ __forceinline__ float lane63(float v) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63)); }
__device__ __forceinline__ double lane63(double v) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b & 0xFFFFFFFFll), 63);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)((unsigned long long)b >> 32), 63);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Reduction over the wave, uniform in every lane.
__device__ __forceinline__ double wave_reduce(double v, int op) {
  if (op == RED_SUM) return lane63(wave_dpp_reduce(v, 0.0, [](double a, double b) { return a + b; }));
  if (op == RED_MIN) return lane63(wave_dpp_reduce(v, 1.0e300, [](double a, double b) { return b < a ? b : a; }));
  return lane63(wave_dpp_reduce(v, -1.0e300, [](double a, double b) { return b > a ? b : a; }));
}

// Result valid in thread 0 (in every thread of a one-wave workgroup, which
// skips the LDS stage and its barriers).
__device__ double block_reduce(double v, int op, double* smem) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  v = wave_reduce(v, op);
  if (blockDim.x == 64) return v;
  __syncthreads();
  if (lane == 0) smem[wave] = v;
  __syncthreads();
  double r = red_identity(op);
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x / 64); ++w) r = red_combine(op, r, smem[w]);
  return r;
}

__device__ __forceinline__ float block_max_f(float v, float* smem) {
  v = lane63(wave_dpp_reduce(v, -__builtin_inff(), [](float a, float b) { return b > a ? b : a; }));
  if (blockDim.x == 64) return v > 0.0f ? v : 0.0f;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) smem[wave] = v;
  __syncthreads();
  float r = 0.0f;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x / 64); ++w) r = smem[w] > r ? smem[w] : r;
  return r;
}

// N reductions of one workgroup at once, results valid in thread 0.  Each
// value goes through the wave reduction of its kind (block_reduce,
// block_max_f or block_sum_i); lane 0 of each wave stores the N wave results,
// and thread 0 combines the waves in wave order from the same starting value.
// The results therefore carry the same bits as N separate reductions, with
// 2 barriers instead of 2 N.  smem holds N * (waves per workgroup) doubles.
template <int N>
__device__ __forceinline__ void block_reduce_n(double (&v)[N], const int (&op)[N], double* smem) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, waves = (int)(blockDim.x / 64);
#pragma unroll
  for (int i = 0; i < N; ++i) {
    if (op[i] == RED_MAXF0)
      v[i] = (double)lane63(wave_dpp_reduce((float)v[i], -__builtin_inff(), [](float a, float b) { return b > a ? b : a; }));
    else if (op[i] == RED_SUMI)
      v[i] = (double)lane63(wave_dpp_reduce((int)v[i], 0, [](int a, int b) { return a + b; }));
    else
      v[i] = wave_reduce(v[i], op[i]);
  }
  if (blockDim.x == 64) {
#pragma unroll
    for (int i = 0; i < N; ++i)
      if (op[i] == RED_MAXF0) v[i] = v[i] > 0.0 ? v[i] : 0.0;
    return;
  }
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int i = 0; i < N; ++i) smem[i * waves + wave] = v[i];
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < N; ++i) {
      if (op[i] == RED_MAXF0) {
        float r = 0.0f;
        for (int w = 0; w < waves; ++w) r = (float)smem[i * waves + w] > r ? (float)smem[i * waves + w] : r;
        v[i] = (double)r;
      } else if (op[i] == RED_SUMI) {
        int r = 0;
        for (int w = 0; w < waves; ++w) r += (int)smem[i * waves + w];
        v[i] = (double)r;
      } else {
        double r = red_identity(op[i]);
        for (int w = 0; w < waves; ++w) r = red_combine(op[i], r, smem[i * waves + w]);
        v[i] = r;
      }
    }
  }
}

__device__ __forceinline__ int block_sum_i(int v, int* smem) {
  v = lane63(wave_dpp_reduce(v, 0, [](int a, int b) { return a + b; }));
  if (blockDim.x == 64) return v;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) smem[wave] = v;
  __syncthreads();
  int r = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x / 64); ++w) r += smem[w];
  return r;
}

// ---------------------------------------------------------------------------
// Workgroup-private scatter tiles for the particle-to-grid passes.  The nodes a
// workgroup's particles touch lie in the box spanned by their B-spline base
// nodes (+2 along each axis).  When that box holds at most TILE_NODES nodes the
// workgroup accumulates its scatter in LDS (ds_add_f32) and then adds each
// touched node to the grid once; otherwise it scatters with global atomics.
// Particles stored in spatial order (as lattice-sampled clouds are) give
// boxes of a few hundred nodes for 256 particles, replacing 27 global atomics
// per particle and channel by a handful.
// ---------------------------------------------------------------------------
// 512 nodes hold the stencil box of a 64-particle workgroup of a
// lattice-ordered cloud (8 cells of 8 particles, also across a row end) and
// keep the tiles small enough for several workgroups per CU.
constexpr int TILE_NODES = 512;

struct TileBox {
  int lx, ly, lz, ex, ey, ez, V;
};

__device__ __forceinline__ int wave_min_i(int v) {
  return lane63(wave_dpp_reduce(v, 0x7FFFFFFF, [](int a, int b) { return b < a ? b : a; }));
}
__device__ __forceinline__ int wave_max_i(int v) {
  return lane63(wave_dpp_reduce(v, -0x7FFFFFFF - 1, [](int a, int b) { return b > a ? b : a; }));
}

// Box of the workgroup's particles' stencils, [min base, max base + 2] per
// axis, from the base nodes the bodies then use: every node a valid particle
// of the workgroup scatters to lies in it.  V = 0: no valid particle, or the
// box holds more than `cap` nodes.  The particle workgroups are one wave, so
// the bounds are wave reductions (uniform in every lane; no LDS, no barrier).
__device__ __forceinline__ TileBox block_tile_box(const int* base, bool valid, int cap) {
  int lo[3], hi[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    lo[d] = wave_min_i(valid ? base[d] : 0x7FFFFFFF);
    hi[d] = wave_max_i(valid ? base[d] : -0x7FFFFFFF);
  }
  TileBox t;
  t.lx = lo[0];
  t.ly = lo[1];
  t.lz = lo[2];
  t.V = 0;
  t.ex = t.ey = t.ez = 0;
  if (lo[0] <= hi[0] && lo[1] <= hi[1] && lo[2] <= hi[2]) {
    const long long ex = (long long)hi[0] - lo[0] + 3, ey = (long long)hi[1] - lo[1] + 3,
                    ez = (long long)hi[2] - lo[2] + 3;
    if (ex * ey * ez <= cap) {
      t.ex = (int)ex;
      t.ey = (int)ey;
      t.ez = (int)ez;
      t.V = (int)(ex * ey * ez);
    }
  }
  return t;
}

// LDS scatter target for particles whose stencils the box was built from:
// every node is inside, so there is no test and no fallback path.
//
// The tile is kept in TILE_COPIES copies, lane parity choosing the copy, and
// the flush sums them.  In a lattice-ordered cloud (8 particles per cell,
// sub-cell index = 4 ox + 2 oy + oz) lane parity is oz, and the only two
// lanes of a wave that hit the same node in one instruction are an oz = 1
// particle of cell k and an oz = 0 particle of cell k + 1: in separate copies
// no two lanes of a copy share a stencil base, which the fused P2G scatter
// needs for its non-atomic read-modify-write (p2g_react_particle).  Any other
// particle order stays correct (the adds are atomic then), only slower.
constexpr int TILE_COPIES = 2;

struct LdsTileSink {
  float* tile;  // this lane's copy
  TileBox b;
  __device__ __forceinline__ void add(int c, long long, int I, int J, int K, float v) const {
    __hip_atomic_fetch_add(tile + c * b.V + ((I - b.lx) * b.ey + (J - b.ly)) * b.ez + (K - b.lz), v, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
  }
};

// this lane's copy of a tile of nch channels
__device__ __forceinline__ float* lane_copy(float* tile, const TileBox& b, int nch) {
  return tile + (threadIdx.x & (TILE_COPIES - 1)) * nch * b.V;
}

__device__ __forceinline__ Stencil stencil_if(bool valid, const float* x, long long P, long long q, float inv) {
  Stencil sn;
  if (valid) {
    sn = particle_stencil(x, P, q, inv);
  } else {
    sn.base[0] = sn.base[1] = sn.base[2] = 0;
  }
  return sn;
}

// LDS scatter target; a node outside the box goes to the global grid.
struct LdsSink {
  float* tile;
  TileBox b;
  GlobalSink g;
  __device__ __forceinline__ void add(int c, long long n, int I, int J, int K, float v) const {
    const unsigned a = (unsigned)(I - b.lx), e = (unsigned)(J - b.ly), f = (unsigned)(K - b.lz);
    // workgroup scope for the LDS add (ds_add_f32), device scope for the global one: the two stay
    // separate instructions instead of one flat atomic on a merged pointer
    if (a < (unsigned)b.ex && e < (unsigned)b.ey && f < (unsigned)b.ez)
      __hip_atomic_fetch_add(tile + c * b.V + ((int)a * b.ey + (int)e) * b.ez + (int)f, v, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_WORKGROUP);
    else
      atomic_add_f32(g.ch[c] + n, v);
  }
};

// Flush of a tile whose channels go to the grid arrays ch[0..nch): the
// copies are summed, then one global atomic per touched node and channel.
__device__ __forceinline__ void tile_flush_ch(const float* tile, const TileBox& b, float* const* ch, int nch,
                                              int copies, int gnx, int gny, int gnz) {
  __syncthreads();
  const int plane = b.ey * b.ez;
  for (int i = threadIdx.x; i < b.V; i += (int)blockDim.x) {
    const int a = i / plane, r = i - a * plane, e = r / b.ez, f = r - e * b.ez;
    const int I = b.lx + a, J = b.ly + e, K = b.lz + f;
    if (I < 0 || J < 0 || K < 0 || I >= gnx || J >= gny || K >= gnz) continue;
    const long long n = ((long long)I * gny + J) * gnz + K;
    for (int c = 0; c < nch; ++c) {
      float v = tile[c * b.V + i];
      for (int k = 1; k < copies; ++k) v += tile[(k * nch + c) * b.V + i];
      if (v != 0.0f) atomic_add_f32(ch[c] + n, v);
    }
  }
}

__device__ __forceinline__ void tile_clear_n(float* tile, int n) {
  for (int i = threadIdx.x; i < n; i += (int)blockDim.x) tile[i] = 0.0f;
  __syncthreads();
}

// Keeps the compiler from moving LDS accesses across this point; the wave
// itself issues its LDS instructions in program order.
__device__ __forceinline__ void wave_order() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The projection and fused P2G tiles update a node with LDS loads and stores
// (read-modify-write) instead of one float atomic per channel.  That is exact
// when the workgroup is one wave and no two of its lanes that share a tile
// copy have the same stencil base: in any one instruction the lanes of a copy
// then address distinct nodes, and two updates of one node by different lanes
// come from different iterations of the 27-node loop, which the wave issues in
// order.  In a lattice-ordered cloud (8 particles per cell, lane parity = the
// particle's z half of the cell, which is also its tile copy) the condition
// always holds.  Otherwise the lanes are split into rounds (scatter_rounds):
// within a round no two lanes of a copy share a base, and the rounds run one
// after the other, so the read-modify-write stays exact in any particle order.

// The read-modify-write rounds of the workgroup's lanes: returns this lane's
// round (-1 for an invalid lane) and in `nrounds` their number (uniform).
// Round r holds the lanes that win their (copy, base) slot once the lanes of
// rounds < r are taken out, so no two lanes of a round and copy share a base.
// Each pass settles at least one lane, so there are at most 64.  Scratch: the
// tile, before it is cleared.
__device__ __forceinline__ int scatter_rounds(const Stencil& sn, bool valid, const TileBox& b, float* tile,
                                              int& nrounds) {
  int* own = reinterpret_cast<int*>(tile);
  const int key = valid ? (int)(threadIdx.x & (TILE_COPIES - 1)) * b.V +
                              ((sn.base[0] - b.lx) * b.ey + (sn.base[1] - b.ly)) * b.ez + (sn.base[2] - b.lz)
                        : 0;
  int round = -1, r = 0;
  bool pending = valid;
  while (__syncthreads_or(pending)) {
    if (pending) own[key] = (int)threadIdx.x;
    __syncthreads();
    if (pending && own[key] == (int)threadIdx.x) {
      round = r;
      pending = false;
    }
    ++r;
  }
  nrounds = r;
  return round;
}

// Where the fused P2G scatter puts a node's 4 floats.
enum FusedScatter {
  SCATTER_TILE_RMW = 0,  // LDS tile, read-modify-write, in rounds of lanes with distinct bases per copy
  SCATTER_GLOBAL = 2     // grid arrays, float atomics (box too large for a tile)
};

__device__ __forceinline__ float pick3(int i, const float (&v)[3]) { return i == 0 ? v[0] : (i == 1 ? v[1] : v[2]); }

// Projection (projection.py:24-64) of the workgroup's particles onto the
// lattice: solid_phi, solid_mass and solid_vel.  A node's 5 floats sit in the
// tile as a float4 (phi, mass, momentum x, y) and a float (momentum z), and
// are updated by read-modify-write when the stencil bases are exclusive per
// tile copy (as in the fused P2G scatter), else with LDS atomics; a box too
// large for the tile scatters with global atomics.  The arithmetic per node
// is project_particle's.
constexpr int PROJ_TILE_NODES = TILE_NODES;

template <int KIND>
__device__ __forceinline__ void proj_scatter(const FsiLbmParams& lp, const FsiCouplingParams& cp, const Stencil& sn,
                                             float vol, float mass, float vx, float vy, float vz, bool checked,
                                             float4* t4, float* t1, const TileBox& b, const GlobalSink& g) {
  const int* base = sn.base;
  const float(&w)[3][3] = sn.w;
  const int nx = lp.nx, ny = lp.ny, nz = lp.nz;
  bool zin[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) zin[c] = !checked || (0 <= base[2] + c && base[2] + c < nz);
#pragma unroll 1
  for (int r = 0; r < 9; ++r) {
    const int a = r / 3, e = r - 3 * (r / 3);
    const float wae = pick3(a, w[0]) * pick3(e, w[1]);
    const int I = base[0] + a, J = base[1] + e;
    const bool row_in = !checked || (0 <= I && I < nx && 0 <= J && J < ny);
    const int row = ((I - b.lx) * b.ey + (J - b.ly)) * b.ez + (base[2] - b.lz);
    const long long crow = ((long long)I * ny + J) * nz + base[2];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float wt = wae * w[2][c];
      const float mc = wt * mass;
      const float vf = wt * vol / cp.cell_volume_norm;
      const float p2 = mc * vx, p3 = mc * vy, p4 = mc * vz;
      if (KIND == SCATTER_TILE_RMW) {
        float4 v = t4[row + c];
        float u = t1[row + c];
        v.x += vf;
        v.y += mc;
        v.z += p2;
        v.w += p3;
        u += p4;
        t4[row + c] = v;
        t1[row + c] = u;
        wave_order();
      } else if (row_in && zin[c]) {
        atomic_add_f32(g.ch[0] + crow + c, vf);
        atomic_add_f32(g.ch[1] + crow + c, mc);
        atomic_add_f32(g.ch[2] + crow + c, p2);
        atomic_add_f32(g.ch[3] + crow + c, p3);
        atomic_add_f32(g.ch[4] + crow + c, p4);
      }
    }
  }
}

__device__ __forceinline__ void proj_block(const FsiLbmParams& lp, const FsiLbmState& ls, const FsiMpmParams& mp,
                                           const FsiMpmState& ms, const FsiCouplingParams& cp, long long q, bool valid,
                                           float* tile) {
  const GlobalSink g = proj_sink(ls, (long long)lp.nx * lp.ny * lp.nz);
  const long long P = mp.n_particles;
  // the particle's data, all loaded at once and without a branch (an idle
  // lane reads particle 0)
  const long long qs = valid ? q : 0;
  Stencil sn = particle_stencil(ms.x, P, qs, cp.inv_dx_norm);
  float J = ms.Jp[qs];
  J = J > 0.0f ? J : 0.0f;
  const float vol = ms.vol0[qs] * J;
  const float mass = ms.mass[qs];
  const float vx = ms.v[qs] * cp.vel_scale_norm_to_lbm;
  const float vy = ms.v[P + qs] * cp.vel_scale_norm_to_lbm;
  const float vz = ms.v[2 * P + qs] * cp.vel_scale_norm_to_lbm;
  if (!valid) sn.base[0] = sn.base[1] = sn.base[2] = 0;
  const TileBox b = block_tile_box(sn.base, valid, PROJ_TILE_NODES);
  int kind = SCATTER_GLOBAL;
  bool checked = true;
  float4* t4 = reinterpret_cast<float4*>(tile);
  float* t1 = reinterpret_cast<float*>(t4 + TILE_COPIES * b.V);
  const int copy = (int)(threadIdx.x & (TILE_COPIES - 1));
  int round = -1, nrounds = 0;
  if (b.V > 0) {
    kind = SCATTER_TILE_RMW;
    round = scatter_rounds(sn, valid, b, tile, nrounds);
    tile_clear_n(tile, TILE_COPIES * 5 * b.V);
    checked = !(b.lx >= 0 && b.ly >= 0 && b.lz >= 0 && b.lx + b.ex <= lp.nx && b.ly + b.ey <= lp.ny &&
                b.lz + b.ez <= lp.nz);
  }
  kind = __builtin_amdgcn_readfirstlane(kind);
  checked = __builtin_amdgcn_readfirstlane((int)checked) != 0;
  if (valid) {
    float4* m4 = t4 + copy * b.V;
    float* m1 = t1 + copy * b.V;
    if (kind == SCATTER_TILE_RMW) {
      for (int r = 0; r < nrounds; ++r)
        if (round == r) proj_scatter<SCATTER_TILE_RMW>(lp, cp, sn, vol, mass, vx, vy, vz, checked, m4, m1, b, g);
    } else {
      proj_scatter<SCATTER_GLOBAL>(lp, cp, sn, vol, mass, vx, vy, vz, checked, m4, m1, b, g);
    }
  }
  if (b.V > 0) {
    // flush: the copies summed per node, one global atomic per touched node
    // and channel
    __syncthreads();
    const int plane = b.ey * b.ez;
    for (int i = threadIdx.x; i < b.V; i += (int)blockDim.x) {
      const int a = i / plane, r = i - a * plane, e = r / b.ez, f = r - e * b.ez;
      const int I = b.lx + a, Jn = b.ly + e, K = b.lz + f;
      if (I < 0 || Jn < 0 || K < 0 || I >= lp.nx || Jn >= lp.ny || K >= lp.nz) continue;
      const long long n = ((long long)I * lp.ny + Jn) * lp.nz + K;
      float4 v = t4[i];
      float u = t1[i];
      for (int k = 1; k < TILE_COPIES; ++k) {
        const float4 o = t4[k * b.V + i];
        v.x += o.x;
        v.y += o.y;
        v.z += o.z;
        v.w += o.w;
        u += t1[k * b.V + i];
      }
      if (v.x != 0.0f) atomic_add_f32(g.ch[0] + n, v.x);
      if (v.y != 0.0f) atomic_add_f32(g.ch[1] + n, v.y);
      if (v.z != 0.0f) atomic_add_f32(g.ch[2] + n, v.z);
      if (v.w != 0.0f) atomic_add_f32(g.ch[3] + n, v.w);
      if (u != 0.0f) atomic_add_f32(g.ch[4] + n, u);
    }
  }
}

// ---------------------------------------------------------------------------
// LBM kernels
// ---------------------------------------------------------------------------
template <typename Pop>
__global__ void __launch_bounds__(TPB) k_lbm_init(FsiLbmParams p, FsiLbmState st) {
  const long long c = (long long)blockIdx.x * TPB + threadIdx.x;
  if (c < (long long)p.nx * p.ny * p.nz) lbm_init_cell<Pop>(p, st, c);
}

template <typename Pop>
__global__ void __launch_bounds__(TPB) k_lbm_rebuild(FsiLbmParams p, FsiLbmState st, int dynamic, int cur) {
  const long long c = (long long)blockIdx.x * TPB + threadIdx.x;
  if (c < (long long)p.nx * p.ny * p.nz) lbm_rebuild_cell<Pop>(p, st, dynamic, cur, c);
}

// Substep launch geometry.  A workgroup covers TPB consecutive cells of one
// i-plane (the (j,k) plane is contiguous in memory), so i, the x-neighbour
// offsets and the x-face BC tests are uniform per workgroup.  With xcd_tile
// each XCD receives one contiguous run of tiles, i.e. a slab of i-planes, and
// keeps that slab (and its L2 lines) in every substep.

// MINW: the waves per SIMD the register budget must allow.  A grid that fits
// on the chip in one wave generation at 4 waves per SIMD runs without a cap
// (MINW 1: more registers, fewer waits); a larger grid keeps 5 waves per SIMD
// resident to hide memory latency and loads populations speculatively.
//
// Wall cells of a gathering moving-wall substep, in a one-generation grid:
// a wave's row of cells typically holds a wall cell at each end, and run in
// place each of them would cost its whole wave the wall work with the other
// lanes idle.  Instead the workgroup queues its wall cells in LDS and one
// wave runs them (one more wave per further 64), chosen by the workgroup
// index so that the four SIMDs share the wall work.  That wave issues the
// queued cells' link loads together with its own cells' population loads,
// so it still waits for memory once.
template <typename Pop, uint32_t FIXED, int MINW, int BS>
__global__ void __launch_bounds__(BS, MINW) k_lbm_substep(FsiLbmParams p, FsiLbmState st, int cur, int parity,
                                                          uint32_t flags, unsigned tiles_per_plane) {
  const unsigned L = p.launch_linear ? blockIdx.x : xcd_tile(blockIdx.x, gridDim.x);
  const unsigned i = L / tiles_per_plane;
  const unsigned jk = (L - i * tiles_per_plane) * BS + threadIdx.x;
  const unsigned plane = (unsigned)p.ny * (unsigned)p.nz;
  const uint32_t fl = (flags & ~(FSI_SUB_STORED_MOMENTS | FSI_SUB_MOVING_WALL | FSI_SUB_CELL_FORCE)) | FIXED;
  if (MINW > 1 || !(FIXED & FSI_SUB_MOVING_WALL) || !(fl & FSI_SUB_GATHER_PREV)) {
    if (jk >= plane) return;
    const unsigned j = jk / (unsigned)p.nz;
    const unsigned k = jk - j * (unsigned)p.nz;
    lbm_substep_cell<Pop, (MINW > 1)>(p, st, cur, parity, fl, (int)i, (int)j, (int)k);
    return;
  }
  constexpr unsigned WAVES = BS / 64;
  __shared__ uint32_t q_jk[BS], q_info[BS];
  __shared__ uint32_t q_n[WAVES];
  const uint32_t N = (uint32_t)p.nx * (uint32_t)p.ny * (uint32_t)p.nz;
  const bool in = jk < plane;
  const unsigned j = in ? jk / (unsigned)p.nz : 0u;
  const unsigned k = in ? jk - j * (unsigned)p.nz : 0u;
  const uint32_t c = ((uint32_t)i * (uint32_t)p.ny + j) * (uint32_t)p.nz + k;
  const uint32_t info = in ? cell_info(st, c) : CI_SOLID;  // outside: solid without links
  const bool wall = (info & CI_SOLID) && (info & CI_MASK);
  const uint64_t bal = __ballot(wall);
  const unsigned wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  if (wall) {
    const unsigned at = wave * 64 + (unsigned)__popcll(bal & ((1ull << lane) - 1ull));
    q_jk[at] = jk;
    q_info[at] = info;
  }
  if (lane == 0) q_n[wave] = (uint32_t)__popcll(bal);
  __syncthreads();
  uint32_t n[WAVES], total = 0;
#pragma unroll
  for (unsigned w = 0; w < WAVES; ++w) total += (n[w] = q_n[w]);
  // queue entries 64t .. 64t+63 are run by wave (blockIdx.x + t) % WAVES
  const unsigned trip = (wave + WAVES - blockIdx.x % WAVES) % WAVES;
  uint32_t e = trip * 64 + lane;
  const bool queued = e < total;
  const Pop Fs = pop_of<Pop>(st.F[cur]);
  const FaceFlags ff = face_flags(p, (int)i);
  const bool fluid = !(info & CI_SOLID);
  float F[Q];
  if (fluid) {
#pragma unroll
    for (int s = 0; s < Q; ++s) F[s] = Fs.ld(N, s, c);
  }
  float Fw[Q];
  double acc[4];
  uint32_t wc = 0, winfo = 0;
  if (queued) {
    unsigned w = 0;
#pragma unroll
    for (unsigned v = 0; v + 1 < WAVES; ++v)
      if (w == v && e >= n[v]) {
        e -= n[v];
        w = v + 1;
      }
    const uint32_t wjk = q_jk[w * 64 + e];
    winfo = q_info[w * 64 + e];
    const unsigned wj = wjk / (unsigned)p.nz, wk = wjk - wj * (unsigned)p.nz;
    wc = ((uint32_t)i * (uint32_t)p.ny + wj) * (uint32_t)p.nz + wk;
    wall_links_load<Pop>(st, Fs, parity, N, wc, winfo, nbr_of(p, (int)i, (int)wj, (int)wk), ff, Fw, acc, 0u);
  }
  if (fluid)
    lbm_fluid_cell<Pop>(p, st, Fs, cur, parity, fl, (int)i, (int)j, (int)k, N, c, info, nbr_of(p, (int)i, (int)j, (int)k),
                        ff, F, fl);
  if (queued) wall_links_accumulate<Pop>(st, wc, winfo, ff, Fw, acc);
}

// ---------------------------------------------------------------------------
// Resident moving-wall subcycle (fsi_lbm_subcycle_resident).
//
// One workgroup of RES_THREADS (1024) work-items per block of 4 x 4 x 64
// cells, every block resident at once (one per CU at 64^3): each cell's 19
// populations stay in VGPRs for the whole subcycle.  Per substep, phase A
// fills each fluid cell's slots: in the block's LDS planes when the cell that
// reads a slot is in the block, else in the halo planes in device memory;
// phase B reads them (res_collide_cell / res_pull_cell in fsi_core.h).
//
// Between the phases a block waits for its 8 neighbour blocks (periodic in i
// and j) to have published the substep.  Each block has a counter word in
// `sync`.  The halo values and the counters are accessed as agent-scope
// relaxed atomics, which are coherent across the XCDs' L2 caches, so no cache
// write-back or invalidation is needed per substep (as fences they measured
// +3.2 us per substep at 64^3).  A block's counter store is ordered after its
// halo stores: every wave waits until its stores have completed
// (s_waitcnt vmcnt(0), the wait a release fence ends with) before the barrier
// after which the counter is stored; a reader issues its halo loads only
// after its poll has returned the new count and a barrier.  Counters only grow
// (each counts every substep its block has run, over all calls), so a launch
// needs no reset and can be replayed from a graph.  Every wait is bounded: a
// neighbour not heard from within RES_MAX_POLLS polls is counted in the error
// word and the block goes on without it, so no launch can hang.
// ---------------------------------------------------------------------------
constexpr uint32_t RES_SYNC_STRIDE = 64;
constexpr int RES_MAX_POLLS = 1 << 16;

// Element c of a halo plane (uniform base, 32-bit offset, as at32) as an
// agent-scope relaxed atomic.
__device__ __forceinline__ void res_put(float* plane, uint32_t c, float v) {
#if defined(__HIP_DEVICE_COMPILE__)
  __hip_atomic_store(&at32(plane, c), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#endif
}
__device__ __forceinline__ float res_get(float* plane, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __hip_atomic_load(&at32(plane, c), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#else
  return 0.0f;
#endif
}

// The subcycle of block b of the lattice (p, st); the workgroup is that
// block's 1024 work-items.
template <bool FORCE>
__device__ __forceinline__ void res_subcycle_block(const FsiLbmParams& p, const FsiLbmState& st, int cur, int n,
                                                   uint32_t flags, float* __restrict__ halo,
                                                   uint32_t* __restrict__ sync, uint32_t b) {
  __shared__ float lv[NLINK * RES_THREADS];     // slot planes, directions 1..18
  __shared__ float uw[3 * RES_RING];            // solid velocities of the block and the ring around it
  __shared__ double wacc[4 * RES_THREADS];      // wall cells' accumulator records
  __shared__ float ou[3 * RES_BJ * RES_NZ];     // velocity of the cells at i = nx - 2
  __shared__ uint32_t s_base;
  __shared__ int bail;
  const uint32_t t = threadIdx.x;
  const ResLane L = res_lane(p, b, t);
  const uint32_t Nk = wave_uniform((uint32_t)p.nx * (uint32_t)p.ny * (uint32_t)RES_NZ);
  const uint32_t info = cell_info(st, L.c);
  const bool fluid = !(info & CI_SOLID);
  const bool wall = !fluid && (info & CI_MASK) != 0;
  const bool face_o = bc_face(p, L.i);
  const PopF32 Fin = pop_of<PopF32>(st.F[cur]), Fout = pop_of<PopF32>(st.F[(cur + n) & 1]);
  uint32_t* own = sync + RES_SYNC_STRIDE * b;
  if (t == 0) {
    s_base = __hip_atomic_load(own, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    bail = 0;
  }
  // lanes 1..8 of wave 0 each watch one neighbour block (the 3 x 3 around it)
  uint32_t* watch = nullptr;
  if (t >= 1 && t <= 8) {
    const int d = (int)t <= 4 ? (int)t - 1 : (int)t;
    const int nbi = p.nx / RES_BI, nbj = p.ny / RES_BJ;
    const int bi = (int)b / nbj, bj = (int)b - bi * nbj;
    const int wi = (bi + d / 3 - 1 + nbi) % nbi, wj = (bj + d % 3 - 1 + nbj) % nbj;
    watch = sync + RES_SYNC_STRIDE * (uint32_t)(wi * nbj + wj);
  }
  float F[Q];
  if (fluid) {
#pragma unroll
    for (int s = 0; s < Q; ++s) F[s] = Fin.ld(Nk, s, L.c);
  }
  // the wall velocities are fixed for the subcycle
  for (int r = (int)t; r < RES_RING; r += RES_THREADS) {
    const uint32_t y = res_ring_cell(p, b, r);
#pragma unroll
    for (int d = 0; d < 3; ++d) uw[d * RES_RING + r] = st.solid_vel[(size_t)d * Nk + y];
  }
  const bool accumulate = wall && (flags & FSI_SUB_GATHER_PREV) != 0;
  if (accumulate) {
#pragma unroll
    for (int d = 0; d < 4; ++d) wacc[4 * t + d] = st.hydro_accum[4 * (size_t)L.c + d];
  }
  __syncthreads();
  const uint32_t base = s_base;
  const uint32_t fixed = FSI_SUB_MOVING_WALL | (FORCE ? FSI_SUB_CELL_FORCE : 0u) | (flags & FSI_SUB_DIAG);
  float* ob = ou + (L.lj * RES_NZ + L.k) * 3;
  // the wave's row by edge class (a block's first, inner or last row in i and
  // j), for which the pulls are compiled with the LDS-or-halo choice of each
  // direction fixed: one straight run of loads per row
  const int row_class = res_row_class(L.li, RES_BI) * 3 + res_row_class(L.lj, RES_BJ);
  float ux = 0.0f, uy = 0.0f, uz = 0.0f;
  for (int step = 0; step < n; ++step) {
    const uint32_t fl = step == 0 ? fixed | (flags & FSI_SUB_STORED_MOMENTS) : fixed;
    const bool last = step == n - 1;
    const bool gather = accumulate && !last;
    // The lane's k, opaque to the optimiser: the cell index, neighbour offsets
    // and slot addresses derived from it are formed in every substep instead
    // of being held across the loop in registers (which spilled).
    int kk = L.k;
    asm volatile("" : "+v"(kk));
    // likewise the lattice size, from which every plane base is formed
    uint32_t N = Nk;
    asm volatile("" : "+s"(N));
    ResLane Lc = L;
    Lc.k = kk;
    Lc.c = L.c - (uint32_t)L.k + (uint32_t)kk;
    const uint32_t c = Lc.c;
    const Nbr nb = nbr_of(p, Lc.i, Lc.j, kk);
    float* hp = halo + (size_t)(step & 1) * NLINK * N;
    const uint32_t tt = (uint32_t)(Lc.li * RES_BJ + Lc.lj) * RES_NZ + (uint32_t)kk;
    if (fluid) {
      // slot (s, c) and, on a link, slot (LR[s], c + e_s): in LDS when the
      // cell reading it is in the block (a wave-uniform test), else in halo
      auto dot = [&](int s) {
        const float* w = uw + res_ring_slot(Lc.li, Lc.lj, kk, EX[s], EY[s], EZ[s]);
        return wall_dot(s, w[0], w[RES_RING], w[2 * RES_RING]);
      };
      // the values that leave the block first, so that their stores are under
      // way while the block's own go to LDS
      auto put = [&](const float* out) {
#pragma unroll
        for (int s = 1; s < Q; ++s)
          if (!res_in_block(Lc.li, Lc.lj, EX[s], EY[s])) res_put(hp + (size_t)(s - 1) * N, c, out[s]);
#pragma unroll
        for (int s = 1; s < Q; ++s)
          if (res_in_block(Lc.li, Lc.lj, EX[s], EY[s])) lv[(s - 1) * RES_THREADS + tt] = out[s];
      };
      auto bounce = [&](int s, float v) {
        if (res_in_block(Lc.li, Lc.lj, EX[s], EY[s]))
          lv[(LR[s] - 1) * RES_THREADS + res_slot(Lc.li, Lc.lj, kk, EX[s], EY[s], EZ[s])] = v;
        else
          res_put(hp + (size_t)(LR[s] - 1) * N, c + (uint32_t)nb.d(s), v);
      };
      res_collide_cell(p, st, fl, step & 1, last, N, c, info, face_o, F, ux, uy, uz, dot, put, bounce);
    }
    if (L.i == p.nx - 2) {
      ob[0] = fluid ? ux : 0.0f;
      ob[1] = fluid ? uy : 0.0f;
      ob[2] = fluid ? uz : 0.0f;
    }
    __builtin_amdgcn_s_waitcnt(0);  // this wave's halo stores (and loads) have completed
    __syncthreads();
    if (t == 0) __hip_atomic_store(own, base + (uint32_t)step + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (watch && !bail) {
      const uint32_t want = base + (uint32_t)step + 1u;
      int polls = 0;
      while ((int32_t)(__hip_atomic_load(watch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - want) < 0) {
        if (++polls > RES_MAX_POLLS) {
          atomicAdd(sync + RES_SYNC_STRIDE * (uint32_t)res_blocks(p), 1u);
          bail = 1;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    __syncthreads();
    if (fluid || gather) {
      float v[Q];
      switch (row_class) {
#define FSI_RES_PULL(CI, CJ)                                                                        \
  case (CI) * 3 + (CJ):                                                                             \
    res_pull_values(v, [&](int q) -> float {                                                        \
      if (res_class_in((CI), (CJ), -EX[q], -EY[q]))                                                 \
        return lv[(q - 1) * RES_THREADS + res_slot(Lc.li, Lc.lj, kk, -EX[q], -EY[q], -EZ[q])];     \
      return res_get(hp + (size_t)(q - 1) * N, c + (uint32_t)nb.d(LR[q]));                          \
    });                                                                                             \
    break;
        FSI_RES_PULL(0, 0)
        FSI_RES_PULL(0, 1)
        FSI_RES_PULL(0, 2)
        FSI_RES_PULL(1, 0)
        FSI_RES_PULL(1, 1)
        FSI_RES_PULL(1, 2)
        FSI_RES_PULL(2, 0)
        FSI_RES_PULL(2, 1)
        FSI_RES_PULL(2, 2)
#undef FSI_RES_PULL
      }
      res_take_values(p, st, Lc, info, face_o, ux, uy, uz, ob[0], ob[1], ob[2], F, wacc + 4 * tt, v);
    }
    __syncthreads();
  }
  if (fluid) {
#pragma unroll
    for (int s = 0; s < Q; ++s) Fout.st(Nk, s, L.c, F[s]);
  }
  if (accumulate) {
#pragma unroll
    for (int d = 0; d < 4; ++d) st.hydro_accum[4 * (size_t)L.c + d] = wacc[4 * t + d];
  }
}

template <bool FORCE>
__global__ void __launch_bounds__(RES_THREADS, 1) k_lbm_resident(FsiLbmParams p, FsiLbmState st, int cur, int n,
                                                                  uint32_t flags, float* __restrict__ halo,
                                                                  uint32_t* __restrict__ sync) {
  res_subcycle_block<FORCE>(p, st, cur, n, flags, halo, sync, blockIdx.x);
}

__global__ void __launch_bounds__(64) k_open_bc_clear(FsiLbmState st) {
  if (threadIdx.x < 5) st.counters->ob_step[threadIdx.x] = 0u;
}

// x-face boundary pass: one thread per (j, k) of a face, blockIdx.y = face;
// the limited mode's clip counters are summed per workgroup, then added
template <typename Pop>
__global__ void __launch_bounds__(TPB) k_lbm_open_bc(FsiLbmParams p, FsiLbmState st, int buf, uint32_t flags) {
  __shared__ int si[TPB / 64];
  const unsigned plane = (unsigned)p.ny * (unsigned)p.nz;
  const unsigned t = blockIdx.x * TPB + threadIdx.x;
  uint32_t cnt[5] = {0u, 0u, 0u, 0u, 0u};
  if (t < plane) {
    const unsigned j = t / (unsigned)p.nz;
    lbm_open_bc_cell<Pop>(p, st, buf, flags, blockIdx.y == 1, (int)j, (int)(t - j * (unsigned)p.nz), cnt);
  }
  if (p.open_bc != FSI_OPEN_BC_REGULARIZED_LIMITED || !p.ob_limiter) return;
#pragma unroll
  for (int q = 0; q < 5; ++q) {
    const int n = block_sum_i((int)cnt[q], si);
    if (threadIdx.x == 0 && n > 0) {
      atomicAdd(&st.counters->ob_step[q], (uint32_t)n);
      atomicAdd((unsigned long long*)&st.counters->ob_run[q], (unsigned long long)n);
    }
  }
}

// Per-direction link statistics: blockIdx.y + 1 is the direction; workgroup
// b sums its contiguous share of the cells, thread-strided, into one record.
template <typename Pop>
__global__ void __launch_bounds__(TPB) k_link_stats_partial(FsiLbmParams p, FsiLbmState st, int cur, int parity,
                                                            double* ws) {
  __shared__ double sd[TPB / 64];
  __shared__ float sf[TPB / 64];
  const int s = (int)blockIdx.y + 1;
  const uint32_t N = (uint32_t)p.nx * (uint32_t)p.ny * (uint32_t)p.nz;
  const uint32_t per = (N + gridDim.x - 1) / gridDim.x;
  const uint32_t c0 = blockIdx.x * per, c1 = c0 + per < N ? c0 + per : N;
  double acc[FSI_LINK_STATS_VALUES] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (uint32_t y = c0 + threadIdx.x; y < c1; y += TPB) link_stats_cell<Pop>(p, st, cur, parity, s, y, acc);
  double* rec = ws + ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * FSI_LINK_STATS_VALUES;
#pragma unroll
  for (int q = 0; q < 5; ++q) {
    const double v = block_reduce(acc[q], RED_SUM, sd);
    if (threadIdx.x == 0) rec[q] = v;
  }
  const float m = block_max_f((float)acc[5], sf);
  if (threadIdx.x == 0) rec[5] = (double)m;
}

// fold: thread s-1 sums direction s's records in workgroup order
__global__ void __launch_bounds__(64) k_link_stats_fold(const double* ws, int blocks, double* out) {
  const int t = threadIdx.x;
  if (t < FSI_LINK_STATS_VALUES) out[t] = 0.0;  // direction 0
  if (t >= 18) return;
  double a[FSI_LINK_STATS_VALUES] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int b = 0; b < blocks; ++b) {
    const double* r = ws + ((size_t)t * blocks + b) * FSI_LINK_STATS_VALUES;
    for (int q = 0; q < 5; ++q) a[q] += r[q];
    a[5] = r[5] > a[5] ? r[5] : a[5];
  }
  for (int q = 0; q < FSI_LINK_STATS_VALUES; ++q) out[(t + 1) * FSI_LINK_STATS_VALUES + q] = a[q];
}

template <typename Pop>
__global__ void __launch_bounds__(TPB) k_lbm_moments(FsiLbmParams p, FsiLbmState st, int cur, uint32_t flags) {
  const long long c = (long long)blockIdx.x * TPB + threadIdx.x;
  if (c < (long long)p.nx * p.ny * p.nz) lbm_moments_cell<Pop>(p, st, cur, flags, c);
}

__global__ void __launch_bounds__(TPB) k_lbm_mb_begin(FsiLbmParams p, FsiLbmState st) {
  const long long c = (long long)blockIdx.x * TPB + threadIdx.x;
  if (c == 0) {
    st.counters->bb_link_count = 0;
    st.counters->bb_max_correction_bits = 0;
    st.counters->mb_sample_count = 0;
    st.counters->accum_norm_max_bits = 0;
    st.counters->mean_norm_max_bits = 0;
    st.counters->bb_net_solid_force[0] = 0.0;
    st.counters->bb_net_solid_force[1] = 0.0;
    st.counters->bb_net_solid_force[2] = 0.0;
  }
  if (c < (long long)p.nx * p.ny * p.nz) lbm_mb_begin_cell(p, st, c);
}

// The finalize reductions, two-pass: every workgroup leaves one record of
// FIN_REC doubles in st.reduce_ws (links, max correction, max mean norm, max
// running norm, net force x/y/z); one workgroup folds them in workgroup order
// into the counters.  Deterministic, and free of the same-address atomics that
// a thousand workgroups would otherwise serialise on.
constexpr int FIN_REC = 8;

// sd: 7 doubles per wave of the workgroup
__device__ __forceinline__ void finalize_block_record(const FinalizeOut& o, double* rec, double* sd) {
  double v[7] = {(double)o.links, (double)o.max_corr, (double)o.mean_norm, (double)o.runmax, o.hx, o.hy, o.hz};
  constexpr int op[7] = {RED_SUMI, RED_MAXF0, RED_MAXF0, RED_MAXF0, RED_SUM, RED_SUM, RED_SUM};
  block_reduce_n(v, op, sd);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < 7; ++i) rec[i] = v[i];
    rec[7] = 0.0;
  }
}

// fold of n_rec finalize records into the counters, by one whole workgroup
__device__ __forceinline__ void finalize_fold(const double* ws, int n_rec, FsiLbmCounters* k, int subcycled,
                                              int n_samples, double* sd) {
  double a[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int r = threadIdx.x; r < n_rec; r += (int)blockDim.x) {
    const double* x = ws + (size_t)r * FIN_REC;
    a[0] += x[0];
    a[1] = x[1] > a[1] ? x[1] : a[1];
    a[2] = x[2] > a[2] ? x[2] : a[2];
    a[3] = x[3] > a[3] ? x[3] : a[3];
    a[4] += x[4];
    a[5] += x[5];
    a[6] += x[6];
  }
  const double links = block_reduce(a[0], RED_SUM, sd);
  const double mc = block_reduce(a[1], RED_MAX, sd);
  const double mn = block_reduce(a[2], RED_MAX, sd);
  const double rm = block_reduce(a[3], RED_MAX, sd);
  const double hx = block_reduce(a[4], RED_SUM, sd);
  const double hy = block_reduce(a[5], RED_SUM, sd);
  const double hz = block_reduce(a[6], RED_SUM, sd);
  if (threadIdx.x == 0) {
    k->bb_link_count = (int32_t)links;
    k->bb_max_correction_bits = __float_as_uint((float)mc);
    if (subcycled) {
      const uint32_t mb = __float_as_uint((float)mn), rb = __float_as_uint((float)rm);
      if (mb > k->mean_norm_max_bits) k->mean_norm_max_bits = mb;
      if (rb > k->accum_norm_max_bits) k->accum_norm_max_bits = rb;
      k->mb_sample_count = n_samples;
    }
    k->bb_net_solid_force[0] = hx;
    k->bb_net_solid_force[1] = hy;
    k->bb_net_solid_force[2] = hz;
  }
}

template <typename Pop>
__global__ void __launch_bounds__(TPB) k_lbm_finalize(FsiLbmParams p, FsiLbmState st, int cur, int parity,
                                                      int n_samples, int subcycled) {
  __shared__ double sd[7 * TPB / 64];
  const long long c = (long long)blockIdx.x * TPB + threadIdx.x;
  FinalizeOut o;
  o.links = 0;
  o.max_corr = 0.0f;
  o.mean_norm = 0.0f;
  o.runmax = 0.0f;
  o.hx = o.hy = o.hz = 0.0;
  if (c < (long long)p.nx * p.ny * p.nz) o = lbm_finalize_cell<Pop>(p, st, cur, parity, n_samples, subcycled, c);
  finalize_block_record(o, st.reduce_ws + (size_t)blockIdx.x * FIN_REC, sd);
}

__global__ void __launch_bounds__(TPB) k_lbm_finalize_fold(FsiLbmState st, int n_rec, int subcycled, int n_samples) {
  __shared__ double sd[TPB / 64];
  finalize_fold(st.reduce_ws, n_rec, st.counters, subcycled, n_samples, sd);
}

__global__ void __launch_bounds__(TPB) k_lbm_stats_partial(FsiLbmParams p, FsiLbmState st, FsiLbmStats* part) {
  __shared__ double sd[TPB / 64];
  double a[LBM_STATS_N];
#pragma unroll
  for (int f = 0; f < LBM_STATS_N; ++f) a[f] = red_identity(lbm_stats_op(f));
  const long long N = (long long)p.nx * p.ny * p.nz;
  for (long long c = (long long)blockIdx.x * TPB + threadIdx.x; c < N; c += (long long)gridDim.x * TPB)
    lbm_stats_cell(p, st, c, a);
  double* out = reinterpret_cast<double*>(part + blockIdx.x);
  for (int f = 0; f < LBM_STATS_N; ++f) {
    const double r = block_reduce(a[f], lbm_stats_op(f), sd);
    if (threadIdx.x == 0) out[f] = r;
  }
}

__global__ void __launch_bounds__(64) k_stats_final(const double* part, int nparts, int nfields, int which,
                                                    double* out) {
  // one wave per field: lane l folds partials l, l + 64, ..., then the wave
  // reduction combines the lanes, both in a fixed order: deterministic
  __shared__ double sd[1];
  const int f = (int)blockIdx.x;
  if (f >= nfields) return;
  const int op = which == 0 ? lbm_stats_op(f) : mpm_stats_op(f);
  double r = red_identity(op);
  for (int b = (int)threadIdx.x; b < nparts; b += 64) r = red_combine(op, r, part[(long long)b * nfields + f]);
  r = block_reduce(r, op, sd);
  if (threadIdx.x == 0) out[f] = r;
}

// ---------------------------------------------------------------------------
// Projection / penalty
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(TPB) k_proj_clear(FsiLbmParams p, FsiLbmState st, FsiCouplingCounters* cc) {
  const long long c = (long long)blockIdx.x * TPB + threadIdx.x;
  const long long N = (long long)p.nx * p.ny * p.nz;
  if (c == 0) {
    cc->projected_mass = 0.0;
    cc->projected_volume_raw = 0.0;
    cc->projected_volume_clamped = 0.0;
    cc->max_phi_raw_bits = 0;
    cc->active_cell_count = 0;
  }
  if (c < N) {
    st.solid_phi[c] = 0.0f;
    st.solid_mass[c] = 0.0f;
    st.solid_vel[c] = 0.0f;
    st.solid_vel[N + c] = 0.0f;
    st.solid_vel[2 * N + c] = 0.0f;
  }
}

__global__ void __launch_bounds__(PTPB) k_proj_particles(FsiLbmParams lp, FsiLbmState ls, FsiMpmParams mp,
                                                         FsiMpmState ms, FsiCouplingParams cp) {
  __shared__ __align__(16) float tile[TILE_COPIES * 5 * PROJ_TILE_NODES];
  const long long q = (long long)blockIdx.x * PTPB + threadIdx.x;
  proj_block(lp, ls, mp, ms, cp, q, q < mp.n_particles, tile);
}

__global__ void __launch_bounds__(TPB) k_proj_normalize(FsiLbmParams lp, FsiLbmState ls, FsiCouplingParams cp,
                                                        FsiCouplingCounters* cc) {
  __shared__ double sd[5 * TPB / 64];
  const long long c = (long long)blockIdx.x * TPB + threadIdx.x;
  ProjOut o;
  o.phi_raw = 0.0f;
  o.phi_clamped = 0.0f;
  o.mass = 0.0f;
  o.active = 0;
  if (c < (long long)lp.nx * lp.ny * lp.nz) o = normalize_cell(lp, ls, cp, c);
  double v[5] = {(double)(o.phi_raw > 0.0f ? o.phi_raw : 0.0f), (double)o.active, (double)o.mass,
                 (double)o.phi_raw * cp.cell_volume_norm, (double)o.phi_clamped * cp.cell_volume_norm};
  constexpr int op[5] = {RED_MAXF0, RED_SUMI, RED_SUM, RED_SUM, RED_SUM};
  block_reduce_n(v, op, sd);
  const float mp = (float)v[0];
  const int act = (int)v[1];
  const double m = v[2], vr = v[3], vc = v[4];
  if (threadIdx.x == 0) {
    // a zero maximum (most workgroups) cannot raise the counter: no atomic on
    // the shared word
    if (__float_as_uint(mp) != 0u) atomicMax(&cc->max_phi_raw_bits, __float_as_uint(mp));
    if (act) atomicAdd(&cc->active_cell_count, act);
    if (m != 0.0) atomicAdd(&cc->projected_mass, m);
    if (vr != 0.0) atomicAdd(&cc->projected_volume_raw, vr);
    if (vc != 0.0) atomicAdd(&cc->projected_volume_clamped, vc);
  }
}

__global__ void k_penalty_clear(FsiCouplingCounters* cc) {
  if (threadIdx.x == 0) {
    cc->active_force_cell_count = 0;
    cc->max_cell_force_bits = 0;
    cc->max_hydro_force_bits = 0;
    for (int d = 0; d < 3; ++d) {
      cc->net_cell_force[d] = 0.0;
      cc->net_hydro_force[d] = 0.0;
    }
  }
}

__global__ void __launch_bounds__(TPB) k_penalty(FsiLbmParams lp, FsiLbmState ls, FsiCouplingParams cp,
                                                 FsiCouplingCounters* cc) {
  __shared__ double sd[5 * TPB / 64];
  const long long c = (long long)blockIdx.x * TPB + threadIdx.x;
  float f[3] = {0.0f, 0.0f, 0.0f};
  int active = 0;
  if (c < (long long)lp.nx * lp.ny * lp.nz) penalty_cell(lp, ls, cp, c, f, active);
  double v[5] = {(double)norm3(f[0], f[1], f[2]), (double)active, (double)f[0], (double)f[1], (double)f[2]};
  constexpr int op[5] = {RED_MAXF0, RED_SUMI, RED_SUM, RED_SUM, RED_SUM};
  block_reduce_n(v, op, sd);
  const float mfn = (float)v[0];
  const int act = (int)v[1];
  const double net[3] = {v[2], v[3], v[4]};
  if (threadIdx.x == 0) {
    if (act) atomicAdd(&cc->active_force_cell_count, act);
    if (__float_as_uint(mfn) != 0u) {
      atomicMax(&cc->max_cell_force_bits, __float_as_uint(mfn));
      atomicMax(&cc->max_hydro_force_bits, __float_as_uint(mfn));
    }
    for (int d = 0; d < 3; ++d)
      if (net[d] != 0.0) {
        atomicAdd(&cc->net_cell_force[d], net[d]);
        atomicAdd(&cc->net_hydro_force[d], -net[d]);
      }
  }
}

// ---------------------------------------------------------------------------
// MPM kernels
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(TPB) k_mpm_clear(FsiMpmParams mp, FsiMpmState ms) {
  const long long n = (long long)blockIdx.x * TPB + threadIdx.x;
  const long long G = (long long)mp.gnx * mp.gny * mp.gnz;
  if (n < G) {
    ms.grid_v[n] = 0.0f;
    ms.grid_v[G + n] = 0.0f;
    ms.grid_v[2 * G + n] = 0.0f;
    ms.grid_m[n] = 0.0f;
    ms.grid_f[n] = 0.0f;
    ms.grid_f[G + n] = 0.0f;
    ms.grid_f[2 * G + n] = 0.0f;
  }
}

// P2G of particle q (valid) by a whole workgroup.
__device__ __forceinline__ void p2g_block(const FsiMpmParams& mp, const FsiMpmState& ms, long long q, bool valid,
                                          float* tile) {
  const GlobalSink g = p2g_sink(ms, (long long)mp.gnx * mp.gny * mp.gnz);
  const Stencil sn = stencil_if(valid, ms.x, mp.n_particles, q, mp.inv_dx);
  const TileBox b = block_tile_box(sn.base, valid, TILE_NODES);
  if (b.V > 0) {
    tile_clear_n(tile, TILE_COPIES * 4 * b.V);
    LdsTileSink k;
    k.tile = lane_copy(tile, b, 4);
    k.b = b;
    if (valid) mpm_p2g_particle(mp, ms, q, sn, k);
    tile_flush_ch(tile, b, g.ch, 4, TILE_COPIES, mp.gnx, mp.gny, mp.gnz);
  } else if (valid) {
    mpm_p2g_particle(mp, ms, q, sn, g);
  }
}

__global__ void __launch_bounds__(PTPB) k_mpm_p2g(FsiMpmParams mp, FsiMpmState ms) {
  __shared__ float tile[TILE_COPIES * 4 * TILE_NODES];
  const long long q = (long long)blockIdx.x * PTPB + threadIdx.x;
  p2g_block(mp, ms, q, q < mp.n_particles, tile);
}

// Reaction scatter of particle q (valid) by a whole workgroup.
__device__ __forceinline__ ReactOut reaction_block(const FsiLbmParams& lp, const FsiLbmState& ls,
                                                   const FsiMpmParams& mp, const FsiMpmState& ms,
                                                   const FsiCouplingParams& cp, int mode, long long q, bool valid,
                                                   float* tile) {
  ReactOut o;
  o.active = 0;
  o.particle_norm = 0.0f;
  o.max_contrib = 0.0f;
  o.f[0] = o.f[1] = o.f[2] = 0.0f;
  o.g[0] = o.g[1] = o.g[2] = 0.0f;
  const GlobalSink g = reaction_sink(ms, (long long)mp.gnx * mp.gny * mp.gnz);
  const Stencil sn = stencil_if(valid, ms.x, mp.n_particles, q, cp.inv_dx_norm);
  const TileBox b = block_tile_box(sn.base, valid, TILE_NODES);
  if (b.V > 0) {
    tile_clear_n(tile, TILE_COPIES * 3 * b.V);
    LdsTileSink k;
    k.tile = lane_copy(tile, b, 3);
    k.b = b;
    if (valid) o = reaction_particle(lp, ls, mp, ms, cp, mode, q, sn, k);
    tile_flush_ch(tile, b, g.ch, 3, TILE_COPIES, mp.gnx, mp.gny, mp.gnz);
  } else if (valid) {
    o = reaction_particle(lp, ls, mp, ms, cp, mode, q, sn, g);
  }
  return o;
}

__global__ void k_reaction_clear(FsiCouplingCounters* cc) {
  if (threadIdx.x == 0) {
    cc->active_reaction_particle_count = 0;
    cc->max_particle_reaction_bits = 0;
    cc->max_grid_reaction_bits = 0;
    for (int d = 0; d < 3; ++d) {
      cc->net_particle_reaction[d] = 0.0;
      cc->net_grid_reaction[d] = 0.0;
    }
  }
}

__global__ void __launch_bounds__(PTPB) k_mpm_reaction(FsiLbmParams lp, FsiLbmState ls, FsiMpmParams mp,
                                                      FsiMpmState ms, FsiCouplingParams cp,
                                                      FsiCouplingCounters* cc, int mode) {
  __shared__ double sd[TPB / 64];
  __shared__ float sf[TPB / 64];
  __shared__ int si[TPB / 64];
  __shared__ float tile[TILE_COPIES * 3 * TILE_NODES];
  const long long q = (long long)blockIdx.x * PTPB + threadIdx.x;
  const ReactOut o = reaction_block(lp, ls, mp, ms, cp, mode, q, q < mp.n_particles, tile);
  const int act = block_sum_i(o.active, si);
  const float mpn = block_max_f(o.particle_norm, sf);
  const float mgc = block_max_f(o.max_contrib, sf);
  double net[3], netg[3];
  for (int d = 0; d < 3; ++d) net[d] = block_reduce((double)o.f[d], RED_SUM, sd);
  for (int d = 0; d < 3; ++d) netg[d] = block_reduce((double)o.g[d], RED_SUM, sd);
  if (threadIdx.x == 0) {
    if (act) atomicAdd(&cc->active_reaction_particle_count, act);
    if (__float_as_uint(mpn) != 0u) atomicMax(&cc->max_particle_reaction_bits, __float_as_uint(mpn));
    if (__float_as_uint(mgc) != 0u) atomicMax(&cc->max_grid_reaction_bits, __float_as_uint(mgc));
    for (int d = 0; d < 3; ++d) {
      if (net[d] != 0.0) atomicAdd(&cc->net_particle_reaction[d], net[d]);
      if (netg[d] != 0.0) atomicAdd(&cc->net_grid_reaction[d], netg[d]);
    }
  }
}

// ---------------------------------------------------------------------------
// Fused MPM substep with the coupling reaction.  One pass over the particles
// scatters P2G and the reaction into an LDS tile; the reaction reads the J
// that P2G just stored for the particle, as the reference's p2g ->
// add_*_reaction_to_mpm_grid order does.  Each block leaves its reaction
// reductions in a workspace record; block 0 of the grid update folds the
// records in block order into the counters (deterministic, and no contended
// same-address atomics).
//
// The reaction's grid force enters the grid update only as dt * f / m next to
// the momentum (solid.py:252-259), so the fused pass adds dt * f into the
// particle's momentum before its scatter and leaves grid_f at zero: the update
// is the same in exact arithmetic, and each node receives 4 floats (momentum
// x, y, z and mass), kept together in the tile (array of float4).
// ---------------------------------------------------------------------------
constexpr int FUSED_TILE_NODES = TILE_NODES;
constexpr int REACT_REC = 9;  // active count, max particle norm, max grid contribution, net particle force x, y, z,
                              // net in-grid force x, y, z

// The 27-node scatter of p2g_react_particle for one scatter kind: node (a, e,
// c) receives wt * (m + aff (x_node - x_p)) and wt * mass.  Returns the
// largest weight of a node inside the grid.
template <int KIND>
__device__ __forceinline__ float fused_scatter(const FsiMpmParams& mp, const Stencil& sn, const P2gTerms& t,
                                               const float (&mv)[3], bool checked, float4* tile, const TileBox& b,
                                               const GlobalSink& g) {
  const int* base = sn.base;
  const float* fx = sn.fx;
  const float(&w)[3][3] = sn.w;
  const int gnx = mp.gnx, gny = mp.gny, gnz = mp.gnz;
  float wmax = 0.0f;
  float dpz[3];
  bool zin[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    dpz[c] = ((float)c - fx[2]) * mp.dx;
    zin[c] = !checked || (0 <= base[2] + c && base[2] + c < gnz);
  }
#pragma unroll 1
  for (int r = 0; r < 9; ++r) {
    const int a = r / 3, e = r - 3 * (r / 3);
    const float wae = pick3(a, w[0]) * pick3(e, w[1]);
    const float dpx = ((float)a - fx[0]) * mp.dx;
    const float dpy = ((float)e - fx[1]) * mp.dx;
    const float bx = mv[0] + t.aff[0] * dpx + t.aff[1] * dpy;
    const float by = mv[1] + t.aff[3] * dpx + t.aff[4] * dpy;
    const float bz = mv[2] + t.aff[6] * dpx + t.aff[7] * dpy;
    const int I = base[0] + a, J = base[1] + e;
    const bool row_in = !checked || (0 <= I && I < gnx && 0 <= J && J < gny);
    float4* row = tile + ((I - b.lx) * b.ey + (J - b.ly)) * b.ez + (base[2] - b.lz);
    const long long nrow = ((long long)I * gny + J) * gnz + base[2];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float wt = wae * w[2][c];
      const float ax = wt * (bx + t.aff[2] * dpz[c]);
      const float ay = wt * (by + t.aff[5] * dpz[c]);
      const float az = wt * (bz + t.aff[8] * dpz[c]);
      const float am = wt * t.mass;
      const bool in = row_in && zin[c];
      wmax = in && wt > wmax ? wt : wmax;
      if (KIND == SCATTER_TILE_RMW) {
        float4 v = row[c];
        v.x += ax;
        v.y += ay;
        v.z += az;
        v.w += am;
        row[c] = v;
        wave_order();
      } else if (in) {
        atomic_add_f32(g.ch[0] + nrow + c, ax);
        atomic_add_f32(g.ch[1] + nrow + c, ay);
        atomic_add_f32(g.ch[2] + nrow + c, az);
        atomic_add_f32(g.ch[3] + nrow + c, am);
      }
    }
  }
  return wmax;
}

// P2G of particle q (+ for MODE 1 penalty / 2 moving boundary with `fold`, the
// reaction force, added into the momentum) with one scatter of 4 floats per
// stencil node.  `checked` = false: every stencil node of the workgroup is
// inside the MPM grid and (fold) the lattice, so nothing is tested; otherwise
// the gather tests each lattice cell, and a node outside the grid lands in the
// tile's margin (which the flush skips) or, for the global scatter, is
// dropped.  `kind` and `checked` are uniform over the workgroup.
//
// The largest node contribution norm3(wt * f) over the nodes inside the grid
// is norm3(wmax * f) for the largest such weight wmax: each rounded step
// (product, squares, sum, square root) is monotonic in wt.
//
// The particle's state `in` was loaded at the start of the kernel, and the
// lattice gather is issued before the polar decomposition: both memory round
// trips overlap work instead of following it.
template <int MODE>
__device__ __forceinline__ ReactOut p2g_react_particle(const FsiLbmParams& lp, const FsiLbmState& ls,
                                                       const FsiMpmParams& mp, const FsiMpmState& ms,
                                                       const FsiCouplingParams& cp, long long q, const P2gIn& in,
                                                       const Stencil& sn, bool fold, bool checked, int kind,
                                                       int round, int nrounds, float4* tile, const TileBox& b,
                                                       const GlobalSink& g) {
  ReactOut o;
  o.active = 0;
  o.particle_norm = o.max_contrib = 0.0f;
  o.f[0] = o.f[1] = o.f[2] = 0.0f;
  o.g[0] = o.g[1] = o.g[2] = 0.0f;
  HydroGather G;
  if (MODE && fold) hydro_gather_load(lp, ls, MODE, sn, checked, G);
  const P2gTerms t = p2g_terms_of(mp, in);
  ms.Jp[q] = t.J;
  float mv[3] = {t.mv[0], t.mv[1], t.mv[2]};
  if (MODE && fold) {
    o = reaction_from_gather(cp, MODE, t.J, in.vol0, sn, G);
    const float win = inside_weight(sn, mp.gnx, mp.gny, mp.gnz);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      mv[d] += mp.dt * o.f[d];
      o.g[d] = o.f[d] * win;
    }
  }
  float wmax = 0.0f;
  if (kind == SCATTER_TILE_RMW) {
    for (int r = 0; r < nrounds; ++r)
      if (round == r) wmax = fused_scatter<SCATTER_TILE_RMW>(mp, sn, t, mv, checked, tile, b, g);
  } else {
    wmax = fused_scatter<SCATTER_GLOBAL>(mp, sn, t, mv, checked, tile, b, g);
  }
  if (MODE && fold) {
    const float cn = norm3(wmax * o.f[0], wmax * o.f[1], wmax * o.f[2]);
    o.max_contrib = cn > 0.0f ? cn : 0.0f;
  }
  return o;
}

// Flush of the fused tile: the copies are summed per node, then one global
// atomic per touched node and channel into grid_v (x, y, z) and grid_m.
__device__ __forceinline__ void tile_flush_fused(const float4* tile, const TileBox& b, const GlobalSink& g, int gnx,
                                                 int gny, int gnz) {
  __syncthreads();
  const int plane = b.ey * b.ez;
  for (int i = threadIdx.x; i < b.V; i += (int)blockDim.x) {
    const int a = i / plane, r = i - a * plane, e = r / b.ez, f = r - e * b.ez;
    const int I = b.lx + a, J = b.ly + e, K = b.lz + f;
    if (I < 0 || J < 0 || K < 0 || I >= gnx || J >= gny || K >= gnz) continue;
    const long long n = ((long long)I * gny + J) * gnz + K;
    float4 v = tile[i];
    for (int k = 1; k < TILE_COPIES; ++k) {
      const float4 u = tile[k * b.V + i];
      v.x += u.x;
      v.y += u.y;
      v.z += u.z;
      v.w += u.w;
    }
    if (v.x != 0.0f) atomic_add_f32(g.ch[0] + n, v.x);
    if (v.y != 0.0f) atomic_add_f32(g.ch[1] + n, v.y);
    if (v.z != 0.0f) atomic_add_f32(g.ch[2] + n, v.z);
    if (v.w != 0.0f) atomic_add_f32(g.ch[3] + n, v.w);
  }
}

// P2G (+ reaction, MODE 1 penalty / 2 moving boundary) of particle q (valid) by
// a whole workgroup; the block's reaction record goes to rec[0..REACT_REC).
template <int MODE>
__device__ __forceinline__ void p2g_react_block(const FsiLbmParams& lp, const FsiLbmState& ls, const FsiMpmParams& mp,
                                                const FsiMpmState& ms, const FsiCouplingParams& cp, double* rec,
                                                long long q, bool valid, float* tile, double* sd,
                                                float* sf, int* si) {
  const long long G = (long long)mp.gnx * mp.gny * mp.gnz;
  const GlobalSink gp = p2g_sink(ms, G), gr = reaction_sink(ms, G);
  // x, then the particle's p2g inputs, without a branch (an idle lane reads
  // particle 0), so the stencil waits for x alone and the rest stays in flight
  const long long qs = valid ? q : 0;
  Stencil sn = particle_stencil(ms.x, mp.n_particles, qs, mp.inv_dx);
  const P2gIn in = p2g_load(ms, mp.n_particles, qs);
  if (!valid) sn.base[0] = sn.base[1] = sn.base[2] = 0;
  const TileBox b = block_tile_box(sn.base, valid, FUSED_TILE_NODES);
  // the reaction samples the lattice with cp.inv_dx_norm; it rides on the P2G
  // stencil when that spacing is the MPM grid's (the drivers' setting), and
  // otherwise scatters on its own stencil into grid_f
  const bool same = cp.inv_dx_norm == mp.inv_dx;
  ReactOut o;
  o.active = 0;
  o.particle_norm = o.max_contrib = 0.0f;
  o.f[0] = o.f[1] = o.f[2] = 0.0f;
  o.g[0] = o.g[1] = o.g[2] = 0.0f;
  int kind = SCATTER_GLOBAL;
  bool checked = true;
  float4* mine = reinterpret_cast<float4*>(tile) + (threadIdx.x & (TILE_COPIES - 1)) * b.V;
  int round = -1, nrounds = 0;
  if (b.V > 0) {
    kind = SCATTER_TILE_RMW;
    round = scatter_rounds(sn, valid, b, tile, nrounds);
    tile_clear_n(tile, TILE_COPIES * 4 * b.V);
    const bool in_grid = b.lx >= 0 && b.ly >= 0 && b.lz >= 0 && b.lx + b.ex <= mp.gnx && b.ly + b.ey <= mp.gny &&
                         b.lz + b.ez <= mp.gnz;
    const bool in_lat =
        b.lx >= 0 && b.ly >= 0 && b.lz >= 0 && b.lx + b.ex <= lp.nx && b.ly + b.ey <= lp.ny && b.lz + b.ez <= lp.nz;
    checked = !(in_grid && (!(MODE && same) || in_lat));
  }
  // uniform over the wave: scalar branches
  kind = __builtin_amdgcn_readfirstlane(kind);
  checked = __builtin_amdgcn_readfirstlane((int)checked) != 0;
  if (valid) {
    o = p2g_react_particle<MODE>(lp, ls, mp, ms, cp, q, in, sn, same, checked, kind, round, nrounds, mine, b, gp);
    // a reaction on another spacing scatters on its own stencil into grid_f
    if (MODE && !same) o = reaction_particle(lp, ls, mp, ms, cp, MODE, q, gr);
  }
  if (b.V > 0) tile_flush_fused(reinterpret_cast<const float4*>(tile), b, gp, mp.gnx, mp.gny, mp.gnz);
  if (MODE) {
    const int act = block_sum_i(o.active, si);
    const float mpn = block_max_f(o.particle_norm, sf);
    const float mgc = block_max_f(o.max_contrib, sf);
    double net[3], netg[3];
    for (int d = 0; d < 3; ++d) net[d] = block_reduce((double)o.f[d], RED_SUM, sd);
    for (int d = 0; d < 3; ++d) netg[d] = block_reduce((double)o.g[d], RED_SUM, sd);
    if (threadIdx.x == 0) {
      rec[0] = (double)act;
      rec[1] = (double)mpn;
      rec[2] = (double)mgc;
      for (int d = 0; d < 3; ++d) {
        rec[3 + d] = net[d];
        rec[6 + d] = netg[d];
      }
    }
  }
}

template <int MODE>  // 0: no reaction, 1: penalty reaction, 2: moving-boundary reaction
__global__ void __launch_bounds__(PTPB) k_mpm_p2g_react(FsiLbmParams lp, FsiLbmState ls, FsiMpmParams mp,
                                                        FsiMpmState ms, FsiCouplingParams cp, double* rec) {
  __shared__ __align__(16) float tile[TILE_COPIES * 4 * FUSED_TILE_NODES];
  __shared__ double sd[PTPB / 64];
  __shared__ float sf[PTPB / 64];
  __shared__ int si[PTPB / 64];
  const long long q = (long long)blockIdx.x * PTPB + threadIdx.x;
  p2g_react_block<MODE>(lp, ls, mp, ms, cp, rec + (size_t)blockIdx.x * REACT_REC, q, q < mp.n_particles, tile,
                        sd, sf, si);
}

// Fold of the per-block reaction records into the counters, by one whole
// workgroup: lane l of its first wave sums records l, l + 64, ... and the
// workgroup reduction adds the lanes, so the order (and the result) is the
// same for every workgroup size and launch.
__device__ __forceinline__ void fold_react_records(const double* rec, int n_rec, FsiCouplingCounters* cc,
                                                   double* sd) {
  double a[REACT_REC];
#pragma unroll
  for (int e = 0; e < REACT_REC; ++e) a[e] = 0.0;
  for (int k = threadIdx.x; threadIdx.x < 64 && k < n_rec; k += 64) {
    // agent-scope loads: the records may come from other workgroups of the
    // same launch (the resident MPM subcycle)
    double r[REACT_REC];
#pragma unroll
    for (int e = 0; e < REACT_REC; ++e)
      r[e] = __hip_atomic_load(const_cast<double*>(rec) + (size_t)k * REACT_REC + e, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
    a[0] += r[0];
    a[1] = r[1] > a[1] ? r[1] : a[1];
    a[2] = r[2] > a[2] ? r[2] : a[2];
#pragma unroll
    for (int e = 3; e < REACT_REC; ++e) a[e] += r[e];
  }
  double t[REACT_REC];
#pragma unroll
  for (int e = 0; e < REACT_REC; ++e) t[e] = block_reduce(a[e], (e == 1 || e == 2) ? RED_MAX : RED_SUM, sd);
  if (threadIdx.x == 0) {
    cc->active_reaction_particle_count = (int32_t)t[0];
    cc->max_particle_reaction_bits = __float_as_uint((float)t[1]);
    cc->max_grid_reaction_bits = __float_as_uint((float)t[2]);
    for (int d = 0; d < 3; ++d) {
      cc->net_particle_reaction[d] = t[3 + d];
      cc->net_grid_reaction[d] = t[6 + d];
    }
  }
}

// grid_update, and (block 0) the fold of the reaction records into the counters.
__global__ void __launch_bounds__(TPB) k_mpm_grid_update_react(FsiMpmParams mp, FsiMpmState ms, const double* rec,
                                                               int n_rec, FsiCouplingCounters* cc) {
  __shared__ double sd[TPB / 64];
  const long long n = (long long)blockIdx.x * TPB + threadIdx.x;
  if (n < (long long)mp.gnx * mp.gny * mp.gnz) mpm_grid_node(mp, ms, n);
  if (rec && blockIdx.x == 0) fold_react_records(rec, n_rec, cc, sd);
}

// ---------------------------------------------------------------------------
// Resident MPM subcycle (fsi_mpm_subcycle_coupled): n substeps of
// fsi_mpm_substep_coupled in one launch.  One PTPB workgroup per particle
// block, every workgroup resident at once; each lane keeps its particle's x,
// v, C and F in registers for the whole subcycle.
//
// Substep s scatters into grid buffer buf(s) = s mod 3 of the caller's scratch
// and G2P reads it back, with the grid update applied per gathered node.
// With three buffers one grid barrier per substep, between the scatter and
// the gather, suffices: during its scatter of substep s a workgroup also
// clears the nodes it scattered to in substep s - 2 (the box of its stencils
// then), in buf(s + 1), which every gather of substep s - 2 finished reading
// before the barrier of substep s - 1 and which the scatter of substep s + 1
// writes only after the barrier of s.  The scratch is zero on entry and on
// exit: after a last barrier each workgroup clears its boxes of the last two
// substeps.  The state's grid is cleared at the start and receives, from the
// last gather, the velocity and mass of every node a workgroup's stencils
// cover (workgroups that share a node write the same values), so the state
// after the call is that of n calls of fsi_mpm_substep_coupled.
//
// Grid values cross workgroups (and XCDs) only through agent-scope atomics:
// the scatter's float adds, the clears as atomic stores, and the gather's
// atomic loads, which read the coherent level and not an XCD's L2.  Before a
// workgroup counts itself in at a barrier its wave waits for its own stores
// and adds to complete (s_waitcnt vmcnt(0)).  The barrier counts arrivals in
// eight words (workgroup b in word b mod 8); lanes 0..7 poll them, at most
// MPM_RES_MAX_POLLS times (then the timeout word is counted and the
// workgroup goes on: no launch can hang).  The last workgroup to leave resets
// the arrival words, so every launch starts from zero.
// ---------------------------------------------------------------------------
constexpr uint32_t MPM_RES_STRIDE = 64;  // words between counters (separate lines)
constexpr int MPM_RES_MAX_POLLS = 1 << 18;
constexpr uint32_t MPM_RES_EXIT = 8, MPM_RES_TIMEOUT = 9;
constexpr int MPM_RES_FILL = 4;  // tile nodes per lane whose loads the gather issues together

struct MpmResBuf {
  float* mom;  // momentum x, y, z (G apart)
  float* m;    // mass
};

__device__ __forceinline__ MpmResBuf mpm_res_buf(float* scratch, long long G, int k) {
  MpmResBuf b;
  b.mom = scratch + (size_t)k * 4 * G;
  b.m = b.mom + 3 * G;
  return b;
}

__device__ __forceinline__ void agent_store(float* a, float v) {
  __hip_atomic_store(a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float agent_load(const float* a) {
  return __hip_atomic_load(const_cast<float*>(a), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Barrier k (1, 2, ... within the launch) over nb workgroups.
__device__ __forceinline__ void mpm_res_barrier(uint32_t* sync, int nb, uint32_t k, int& bail) {
  __builtin_amdgcn_s_waitcnt(0);  // this wave's adds, stores and loads have completed
  __syncthreads();
  const uint32_t lane = threadIdx.x;
  if (lane == 0)
    __hip_atomic_fetch_add(sync + MPM_RES_STRIDE * (blockIdx.x & 7u), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (!bail) {
    // word w counts the workgroups b = w mod 8 of [0, nb)
    const uint32_t want = lane < 8 ? k * (((uint32_t)nb + 7u - lane) / 8u) : 0u;
    int polls = 0;
    while (true) {
      const uint32_t got =
          lane < 8 ? __hip_atomic_load(sync + MPM_RES_STRIDE * lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0u;
      if (__ballot(lane < 8 && (int32_t)(got - want) < 0) == 0) break;
      if (++polls > MPM_RES_MAX_POLLS) {
        if (lane == 0) atomicAdd(sync + MPM_RES_STRIDE * MPM_RES_TIMEOUT, 1u);
        bail = 1;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
  }
  __syncthreads();
}

// Agent-scope zero of nodes [lo, hi] (clipped to the grid) of buffer b.
__device__ __forceinline__ void mpm_res_clear_box(const MpmResBuf& b, const int* lo, const int* hi, int gnx, int gny,
                                                  int gnz, long long G) {
  int l[3], e[3];
  const int g[3] = {gnx, gny, gnz};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    l[d] = lo[d] < 0 ? 0 : lo[d];
    const int h = hi[d] > g[d] - 1 ? g[d] - 1 : hi[d];
    e[d] = h - l[d] + 1;
  }
  if (e[0] <= 0 || e[1] <= 0 || e[2] <= 0) return;
  const int V = e[0] * e[1] * e[2], plane = e[1] * e[2];
  for (int i = threadIdx.x; i < V; i += PTPB) {
    const int a = i / plane, r = i - a * plane, c = r / e[2], f = r - c * e[2];
    const long long n = ((long long)(l[0] + a) * gny + (l[1] + c)) * gnz + (l[2] + f);
    agent_store(b.mom + n, 0.0f);
    agent_store(b.mom + G + n, 0.0f);
    agent_store(b.mom + 2 * G + n, 0.0f);
    agent_store(b.m + n, 0.0f);
  }
}

template <int MODE>
__global__ void __launch_bounds__(PTPB) k_mpm_resident(FsiLbmParams lp, FsiLbmState ls, FsiMpmParams mp,
                                                       FsiMpmState ms, FsiCouplingParams cp, int n,
                                                       float* __restrict__ scratch, double* __restrict__ rec,
                                                       FsiCouplingCounters* cc, uint32_t* __restrict__ sync) {
  __shared__ __align__(16) float tile[TILE_COPIES * 4 * FUSED_TILE_NODES];
  __shared__ double sd[PTPB / 64];
  __shared__ float sf[PTPB / 64];
  __shared__ int si[PTPB / 64];
  const int nb = (int)gridDim.x;
  const long long P = mp.n_particles, G = (long long)mp.gnx * mp.gny * mp.gnz;
  const long long q = (long long)blockIdx.x * PTPB + threadIdx.x;
  const bool valid = q < P;
  const long long qs = valid ? q : 0;
  int bail = 0;
  // the start: the state's grid cleared, a slice per workgroup (the last
  // gather writes into it after n barriers)
  {
    const long long per = (G + nb - 1) / nb, n0 = per * blockIdx.x, n1 = n0 + per < G ? n0 + per : G;
    for (long long i = n0 + threadIdx.x; i < n1; i += PTPB) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        agent_store(ms.grid_v + d * G + i, 0.0f);
        agent_store(ms.grid_f + d * G + i, 0.0f);
      }
      agent_store(ms.grid_m + i, 0.0f);
    }
  }
  // the particle's state, held for the subcycle
  float x[3], v[3], C[9], F[9];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    x[d] = ms.x[d * P + qs];
    v[d] = ms.v[d * P + qs];
  }
  load_mat(ms.C, P, qs, C);
  load_mat(ms.F, P, qs, F);
  const float mass = ms.mass[qs], vol0 = ms.vol0[qs];
  // boxes of the last two substeps' stencils (uniform): [0] substep s - 1, [1] s - 2
  int blo[2][3], bhi[2][3];
#pragma unroll
  for (int k = 0; k < 2; ++k)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      blo[k][d] = 0;
      bhi[k][d] = -1;
    }
  uint32_t bar = 0;
  const bool same = cp.inv_dx_norm == mp.inv_dx;  // checked by the host
  for (int s = 0; s < n; ++s) {
    const MpmResBuf cur = mpm_res_buf(scratch, G, s % 3);
    const bool last = s == n - 1;
    const GlobalSink gp = {{cur.mom, cur.mom + G, cur.mom + 2 * G, cur.m, nullptr}};
    // ---- P2G (+ reaction) into buf(s), as p2g_react_block
    Stencil sn = stencil_at(x, mp.inv_dx);
    if (!valid) sn.base[0] = sn.base[1] = sn.base[2] = 0;
    int lo[3], hi[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      lo[d] = wave_min_i(valid ? sn.base[d] : 0x7FFFFFFF);
      hi[d] = wave_max_i(valid ? sn.base[d] + 2 : -0x7FFFFFFF);
    }
    const TileBox b = block_tile_box(sn.base, valid, FUSED_TILE_NODES);
    ReactOut o;
    o.active = 0;
    o.particle_norm = o.max_contrib = 0.0f;
    o.f[0] = o.f[1] = o.f[2] = 0.0f;
    o.g[0] = o.g[1] = o.g[2] = 0.0f;
    int kind = SCATTER_GLOBAL;
    bool checked = true;
    float4* mine = reinterpret_cast<float4*>(tile) + (threadIdx.x & (TILE_COPIES - 1)) * b.V;
    int round = -1, nrounds = 0;
    if (b.V > 0) {
      kind = SCATTER_TILE_RMW;
      round = scatter_rounds(sn, valid, b, tile, nrounds);
      tile_clear_n(tile, TILE_COPIES * 4 * b.V);
      const bool in_grid = b.lx >= 0 && b.ly >= 0 && b.lz >= 0 && b.lx + b.ex <= mp.gnx && b.ly + b.ey <= mp.gny &&
                           b.lz + b.ez <= mp.gnz;
      const bool in_lat = b.lx >= 0 && b.ly >= 0 && b.lz >= 0 && b.lx + b.ex <= lp.nx && b.ly + b.ey <= lp.ny &&
                          b.lz + b.ez <= lp.nz;
      checked = !(in_grid && (!(MODE && same) || in_lat));
    }
    kind = __builtin_amdgcn_readfirstlane(kind);
    checked = __builtin_amdgcn_readfirstlane((int)checked) != 0;
    if (valid) {
      P2gIn in;
#pragma unroll
      for (int e = 0; e < 9; ++e) {
        in.F[e] = F[e];
        in.C[e] = C[e];
      }
#pragma unroll
      for (int d = 0; d < 3; ++d) in.v[d] = v[d];
      in.mass = mass;
      in.vol0 = vol0;
      o = p2g_react_particle<MODE>(lp, ls, mp, ms, cp, q, in, sn, same, checked, kind, round, nrounds, mine, b, gp);
    }
    if (b.V > 0) tile_flush_fused(reinterpret_cast<const float4*>(tile), b, gp, mp.gnx, mp.gny, mp.gnz);
    // the nodes this workgroup scattered to two substeps ago, in buf(s + 1)
    if (s >= 2) mpm_res_clear_box(mpm_res_buf(scratch, G, (s + 1) % 3), blo[1], bhi[1], mp.gnx, mp.gny, mp.gnz, G);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      blo[1][d] = blo[0][d];
      bhi[1][d] = bhi[0][d];
      blo[0][d] = lo[d];
      bhi[0][d] = hi[d];
    }
    if (MODE && last) {  // the counters hold the last substep's reaction
      const int act = block_sum_i(o.active, si);
      const float mpn = block_max_f(o.particle_norm, sf);
      const float mgc = block_max_f(o.max_contrib, sf);
      double net[3], netg[3];
      for (int d = 0; d < 3; ++d) net[d] = block_reduce((double)o.f[d], RED_SUM, sd);
      for (int d = 0; d < 3; ++d) netg[d] = block_reduce((double)o.g[d], RED_SUM, sd);
      if (threadIdx.x == 0) {  // read by workgroup 0: agent-scope stores
        double* r = rec + (size_t)blockIdx.x * REACT_REC;
        const double vals[REACT_REC] = {(double)act, (double)mpn, (double)mgc, net[0], net[1], net[2],
                                        netg[0], netg[1], netg[2]};
        for (int e = 0; e < REACT_REC; ++e) __hip_atomic_store(r + e, vals[e], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    mpm_res_barrier(sync, nb, ++bar, bail);
    // ---- G2P from buf(s), the grid update applied per node.  The workgroup's
    // stencils are those of its scatter (the particles have not moved), so
    // with a tile the box's nodes are loaded once, coalesced, and their
    // velocities kept in LDS (the scatter's tile is free after its flush).
    const float zf[3] = {0.0f, 0.0f, 0.0f};  // grid_f stays zero: the reaction is in the momentum
    float4* vt = reinterpret_cast<float4*>(tile);
    if (b.V > 0) {
      // MPM_RES_FILL nodes per lane per pass, their loads all in flight
      // together (a node outside the grid, or past the box, reads node 0)
      const int plane = b.ey * b.ez;
      for (int i0 = 0; i0 < b.V; i0 += MPM_RES_FILL * PTPB) {
        float nm[MPM_RES_FILL], nv[MPM_RES_FILL][3];
        int ijk[MPM_RES_FILL][3];
        bool use[MPM_RES_FILL];
#pragma unroll
        for (int u = 0; u < MPM_RES_FILL; ++u) {
          const int i = i0 + u * PTPB + (int)threadIdx.x;
          const int a = i / plane, r = i - a * plane, e = r / b.ez, f = r - e * b.ez;
          ijk[u][0] = b.lx + a;
          ijk[u][1] = b.ly + e;
          ijk[u][2] = b.lz + f;
          use[u] = i < b.V && 0 <= ijk[u][0] && ijk[u][0] < mp.gnx && 0 <= ijk[u][1] && ijk[u][1] < mp.gny &&
                   0 <= ijk[u][2] && ijk[u][2] < mp.gnz;
          const long long nn = use[u] ? ((long long)ijk[u][0] * mp.gny + ijk[u][1]) * mp.gnz + ijk[u][2] : 0;
          nv[u][0] = agent_load(cur.mom + nn);
          nv[u][1] = agent_load(cur.mom + G + nn);
          nv[u][2] = agent_load(cur.mom + 2 * G + nn);
          nm[u] = agent_load(cur.m + nn);
        }
#pragma unroll
        for (int u = 0; u < MPM_RES_FILL; ++u) {
          const int i = i0 + u * PTPB + (int)threadIdx.x;
          float vel[3] = {nv[u][0], nv[u][1], nv[u][2]};
          if (use[u] && nm[u] > 0.0f) grid_node_velocity(mp, nm[u], nv[u], zf, ijk[u][0], ijk[u][1], ijk[u][2], vel);
          if (i < b.V) vt[i] = use[u] ? make_float4(vel[0], vel[1], vel[2], 0.0f) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
          if (last && use[u]) {  // the state's grid after the call: grid_update's output
            const long long nn = ((long long)ijk[u][0] * mp.gny + ijk[u][1]) * mp.gnz + ijk[u][2];
#pragma unroll
            for (int d = 0; d < 3; ++d) agent_store(ms.grid_v + d * G + nn, vel[d]);
            agent_store(ms.grid_m + nn, nm[u]);
          }
        }
      }
      __syncthreads();
    }
    if (valid) {
      int base[3];
      float fx[3], w[3][3];
#pragma unroll
      for (int d = 0; d < 3; ++d) bspline(x[d] * mp.inv_dx, base[d], fx[d], w[d]);
      float gv[27][3];
      uint32_t inm = 0u;
#pragma unroll
      for (int k = 0; k < 27; ++k) {
        const int I = base[0] + k / 9, J = base[1] + (k / 3) % 3, K = base[2] + k % 3;
        inm |= (0 <= I && I < mp.gnx && 0 <= J && J < mp.gny && 0 <= K && K < mp.gnz) ? (1u << k) : 0u;
      }
      if (b.V > 0) {
        // the nodes' velocities from the workgroup's tile, one LDS read per node
#pragma unroll
        for (int k = 0; k < 27; ++k) {
          const float4 t = vt[((base[0] + k / 9 - b.lx) * b.ey + (base[1] + (k / 3) % 3 - b.ly)) * b.ez +
                              (base[2] + k % 3 - b.lz)];
          gv[k][0] = t.x;
          gv[k][1] = t.y;
          gv[k][2] = t.z;
        }
      } else {
        // no tile: the particle's own nodes, all 108 loads in flight together
        float gm[27];
#pragma unroll
        for (int k = 0; k < 27; ++k) {
          const long long nn = ((inm >> k) & 1u) ? ((long long)(base[0] + k / 9) * mp.gny + (base[1] + (k / 3) % 3)) *
                                                          mp.gnz + (base[2] + k % 3)
                                                    : 0;
          gv[k][0] = agent_load(cur.mom + nn);
          gv[k][1] = agent_load(cur.mom + G + nn);
          gv[k][2] = agent_load(cur.mom + 2 * G + nn);
          gm[k] = agent_load(cur.m + nn);
        }
#pragma unroll
        for (int k = 0; k < 27; ++k) {
          if (gm[k] > 0.0f) {
            const float mom[3] = {gv[k][0], gv[k][1], gv[k][2]};
            grid_node_velocity(mp, gm[k], mom, zf, base[0] + k / 9, base[1] + (k / 3) % 3, base[2] + k % 3, gv[k]);
          }
        }
        if (last) {  // the state's grid after the call
#pragma unroll
          for (int k = 0; k < 27; ++k) {
            if (!((inm >> k) & 1u)) continue;
            const long long nn =
                ((long long)(base[0] + k / 9) * mp.gny + (base[1] + (k / 3) % 3)) * mp.gnz + (base[2] + k % 3);
#pragma unroll
            for (int d = 0; d < 3; ++d) agent_store(ms.grid_v + d * G + nn, gv[k][d]);
            agent_store(ms.grid_m + nn, gm[k]);
          }
        }
      }
      float xn[3], vn[3], Cn[9], Fn[9], Jn;
      g2p_update(mp, ms, q, fx, w, gv, inm, x, F, xn, vn, Cn, Fn, Jn);
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        x[d] = xn[d];
        v[d] = vn[d];
      }
#pragma unroll
      for (int e = 0; e < 9; ++e) {
        C[e] = Cn[e];
        F[e] = Fn[e];
      }
      if (last) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          ms.x[d * P + q] = x[d];
          ms.v[d * P + q] = v[d];
        }
        store_mat(ms.C, P, q, C);
        store_mat(ms.F, P, q, F);
        ms.Jp[q] = Jn;
      }
    }
  }
  // every gather of the last substep done: this workgroup's boxes of the last
  // two substeps cleared, so the scratch is zero again
  mpm_res_barrier(sync, nb, ++bar, bail);
  mpm_res_clear_box(mpm_res_buf(scratch, G, (n - 1) % 3), blo[0], bhi[0], mp.gnx, mp.gny, mp.gnz, G);
  if (n >= 2) mpm_res_clear_box(mpm_res_buf(scratch, G, (n - 2) % 3), blo[1], bhi[1], mp.gnx, mp.gny, mp.gnz, G);
  if (MODE && blockIdx.x == 0) fold_react_records(rec, nb, cc, sd);
  // the last workgroup out resets the arrival words
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t out =
        __hip_atomic_fetch_add(sync + MPM_RES_STRIDE * MPM_RES_EXIT, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (out == (uint32_t)nb - 1u) {
      for (uint32_t w = 0; w < 8; ++w) __hip_atomic_store(sync + MPM_RES_STRIDE * w, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(sync + MPM_RES_STRIDE * MPM_RES_EXIT, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// ---------------------------------------------------------------------------
// Particle binning (counting sort by stencil base) and the particle-ordered
// transfers.  Bin counts are integer atomics (order-free); each bin's list is
// then sorted by particle index, so start/order do not depend on the schedule.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(TPB) k_fill_i32(int32_t* a, long long n, int32_t v) {
  const long long i = (long long)blockIdx.x * TPB + threadIdx.x;
  if (i < n) a[i] = v;
}

__global__ void __launch_bounds__(TPB) k_bin_keys(const float* x, int32_t P, float inv, BinGrid bg, int32_t* count,
                                                  int32_t* key) {
  const long long p = (long long)blockIdx.x * TPB + threadIdx.x;
  if (p >= P) return;
  const int32_t b = particle_bin(bg, x, P, p, inv);
  key[p] = b;
  atomicAdd(count + b, 1);
}

constexpr int SCAN_TILE = 1024;  // elements per workgroup of the scan (4 per thread)

// Exclusive scan within 1024-element tiles of count[0, n) into start[], and
// each tile's total into part[tile].
__global__ void __launch_bounds__(TPB) k_scan_tiles(const int32_t* count, long long n, int32_t* start,
                                                    int32_t* part) {
  __shared__ int32_t sh[TPB];
  const long long base = (long long)blockIdx.x * SCAN_TILE + 4 * threadIdx.x;
  int32_t v[4];
  int32_t tsum = 0;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    v[e] = base + e < n ? count[base + e] : 0;
    tsum += v[e];
  }
  sh[threadIdx.x] = tsum;
  __syncthreads();
  for (int off = 1; off < TPB; off <<= 1) {  // inclusive Hillis-Steele over the thread sums
    const int32_t add = threadIdx.x >= (unsigned)off ? sh[threadIdx.x - off] : 0;
    __syncthreads();
    sh[threadIdx.x] += add;
    __syncthreads();
  }
  int32_t run = sh[threadIdx.x] - tsum;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    if (base + e < n) start[base + e] = run;
    run += v[e];
  }
  if (threadIdx.x == TPB - 1) part[blockIdx.x] = sh[TPB - 1];
}

// Exclusive scan of the tile totals in place (one workgroup); start[n] = total.
__global__ void __launch_bounds__(TPB) k_scan_parts(int32_t* part, int n_tiles, int32_t* start, long long n) {
  __shared__ int32_t sh[TPB];
  int32_t carry = 0;
  for (int c0 = 0; c0 < n_tiles; c0 += TPB) {
    const int i = c0 + (int)threadIdx.x;
    const int32_t v = i < n_tiles ? part[i] : 0;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < TPB; off <<= 1) {
      const int32_t add = threadIdx.x >= (unsigned)off ? sh[threadIdx.x - off] : 0;
      __syncthreads();
      sh[threadIdx.x] += add;
      __syncthreads();
    }
    if (i < n_tiles) part[i] = carry + sh[threadIdx.x] - v;
    const int32_t total = sh[TPB - 1];
    __syncthreads();
    carry += total;
  }
  if (threadIdx.x == 0) start[n] = carry;
}

__global__ void __launch_bounds__(TPB) k_scan_add(int32_t* start, long long n, const int32_t* part) {
  const long long i = (long long)blockIdx.x * TPB + threadIdx.x;
  if (i < n) start[i] += part[i / SCAN_TILE];
}

// Each particle takes a slot of its bin (count returns to zero).
__global__ void __launch_bounds__(TPB) k_bin_fill(const int32_t* key, int32_t P, const int32_t* start, int32_t* count,
                                                  int32_t* order) {
  const long long p = (long long)blockIdx.x * TPB + threadIdx.x;
  if (p >= P) return;
  const int32_t b = key[p];
  order[start[b] + atomicSub(count + b, 1) - 1] = (int32_t)p;
}

// Each bin's list in increasing particle index (insertion sort: a bin holds a
// few particles).
__global__ void __launch_bounds__(TPB) k_bin_sort(const int32_t* start, long long n_bins, int32_t* order) {
  const long long b = (long long)blockIdx.x * TPB + threadIdx.x;
  if (b >= n_bins) return;
  const int32_t s = start[b], e = start[b + 1];
  for (int32_t i = s + 1; i < e; ++i) {
    const int32_t v = order[i];
    int32_t j = i - 1;
    while (j >= s && order[j] > v) {
      order[j + 1] = order[j];
      --j;
    }
    order[j + 1] = v;
  }
}

// The particle pass of the ordered substep: terms, bin key and count, and the
// workgroup's reaction record (MODE 1 penalty, 2 moving boundary).
template <int MODE>
__global__ void __launch_bounds__(PTPB) k_ordered_terms(FsiLbmParams lp, FsiLbmState ls, FsiMpmParams mp,
                                                        FsiMpmState ms, FsiCouplingParams cp, float* terms, double* rec,
                                                        BinGrid bg, int32_t* count, int32_t* key) {
  __shared__ double sd[PTPB / 64];
  __shared__ float sf[PTPB / 64];
  __shared__ int si[PTPB / 64];
  const long long p = (long long)blockIdx.x * PTPB + threadIdx.x;
  ReactOut o;
  o.active = 0;
  o.particle_norm = o.max_contrib = 0.0f;
  o.f[0] = o.f[1] = o.f[2] = 0.0f;
  o.g[0] = o.g[1] = o.g[2] = 0.0f;
  if (p < mp.n_particles) {
    o = mpm_ordered_terms_particle(lp, ls, mp, ms, cp, MODE, p, terms);
    const int32_t b = particle_bin(bg, ms.x, mp.n_particles, p, mp.inv_dx);
    key[p] = b;
    atomicAdd(count + b, 1);
  }
  if (MODE) {
    const int act = block_sum_i(o.active, si);
    const float mpn = block_max_f(o.particle_norm, sf);
    const float mgc = block_max_f(o.max_contrib, sf);
    double net[3], netg[3];
    for (int d = 0; d < 3; ++d) net[d] = block_reduce((double)o.f[d], RED_SUM, sd);
    for (int d = 0; d < 3; ++d) netg[d] = block_reduce((double)o.g[d], RED_SUM, sd);
    if (threadIdx.x == 0) {
      double* r = rec + (size_t)blockIdx.x * REACT_REC;
      r[0] = (double)act;
      r[1] = (double)mpn;
      r[2] = (double)mgc;
      for (int d = 0; d < 3; ++d) {
        r[3 + d] = net[d];
        r[6 + d] = netg[d];
      }
    }
  }
}

// Node pass of the ordered substep; block 0 also folds the reaction records.
template <bool REACT>
__global__ void __launch_bounds__(TPB) k_ordered_nodes(FsiMpmParams mp, FsiMpmState ms, BinGrid bg,
                                                       const int32_t* start, const int32_t* order, const float* terms,
                                                       const double* rec, int n_rec, FsiCouplingCounters* cc) {
  __shared__ double sd[TPB / 64];
  const long long n = (long long)blockIdx.x * TPB + threadIdx.x;
  if (n < (long long)mp.gnx * mp.gny * mp.gnz) mpm_ordered_node(mp, ms, bg, start, order, terms, REACT, n);
  if (REACT && blockIdx.x == 0) fold_react_records(rec, n_rec, cc, sd);
}

__global__ void __launch_bounds__(TPB) k_project_ordered(FsiLbmParams lp, FsiLbmState ls, FsiMpmParams mp,
                                                         FsiMpmState ms, FsiCouplingParams cp, BinGrid bg,
                                                         const int32_t* start, const int32_t* order) {
  const long long c = (long long)blockIdx.x * TPB + threadIdx.x;
  if (c < (long long)lp.nx * lp.ny * lp.nz) project_ordered_cell(lp, ls, mp, ms, cp, bg, start, order, c);
}

__global__ void k_proj_counters_clear(FsiCouplingCounters* cc) {
  if (threadIdx.x == 0) {
    cc->projected_mass = 0.0;
    cc->projected_volume_raw = 0.0;
    cc->projected_volume_clamped = 0.0;
    cc->max_phi_raw_bits = 0;
    cc->active_cell_count = 0;
  }
}

__global__ void __launch_bounds__(TPB) k_mpm_grid_update(FsiMpmParams mp, FsiMpmState ms) {
  const long long n = (long long)blockIdx.x * TPB + threadIdx.x;
  if (n < (long long)mp.gnx * mp.gny * mp.gnz) mpm_grid_node(mp, ms, n);
}

__global__ void __launch_bounds__(PTPB) k_mpm_g2p(FsiMpmParams mp, FsiMpmState ms) {
  const long long q = (long long)blockIdx.x * PTPB + threadIdx.x;
  if (q < mp.n_particles) mpm_g2p_particle(mp, ms, q);
}

__global__ void __launch_bounds__(TPB) k_mpm_stats_partial(FsiMpmParams mp, FsiMpmState ms, FsiMpmStats* part) {
  __shared__ double sd[TPB / 64];
  double a[MPM_STATS_N];
#pragma unroll
  for (int f = 0; f < MPM_STATS_N; ++f) a[f] = red_identity(mpm_stats_op(f));
  for (long long q = (long long)blockIdx.x * TPB + threadIdx.x; q < mp.n_particles; q += (long long)gridDim.x * TPB)
    mpm_stats_particle(mp, ms, q, a);
  double* out = reinterpret_cast<double*>(part + blockIdx.x);
  for (int f = 0; f < MPM_STATS_N; ++f) {
    const double r = block_reduce(a[f], mpm_stats_op(f), sd);
    if (threadIdx.x == 0) out[f] = r;
  }
}

// ---------------------------------------------------------------------------
// Batched kernels.  The instance table and the block map are read-only for
// the launch (__restrict__): the descriptor fields a workgroup reads can stay
// in scalar registers across the stores it makes, instead of being loaded
// again after each one.  Dispatch slot blockIdx.x takes entry b of the block map
// (b = xcd_tile(...): each XCD gets one contiguous run of the map, i.e. of
// whole instances), which names the instance and the workgroup's index L in
// that instance's domain; its local work item is L * (block size) + thread.
// ---------------------------------------------------------------------------
#define FSI_BATCH_PROLOGUE(DOMAIN, BS)                                                     \
  const unsigned b_ = xcd_tile(blockIdx.x, gridDim.x);                                     \
  const int i_ = bmap[2 * b_], l_ = bmap[2 * b_ + 1];                                       \
  const FsiBatchInstance& I = inst[i_];                                                    \
  const long long e_ = (long long)l_ * (BS) + threadIdx.x;                                 \
  const bool in_ = e_ < batch_domain_size(I, DOMAIN);

// BOX-domain launches: each instance's workgroups (batch_domain_blocks) walk
// the cells or nodes of one of its boxes, clipped to the domain, k fastest,
// item b of the walk going to workgroup b / TPB mod that count.
#define FSI_BATCH_BOX_PROLOGUE                                               \
  const unsigned b_ = xcd_tile(blockIdx.x, gridDim.x);                       \
  const int i_ = bmap[2 * b_], l_ = bmap[2 * b_ + 1];                         \
  const FsiBatchInstance& I = inst[i_];                                      \
  const long long stride_ = batch_domain_blocks(I, FSI_BATCH_DOMAIN_BOX) * TPB;

struct BoxWalk {
  int lo[3], e[3];
  long long n;
};

__device__ __forceinline__ BoxWalk box_walk(const int32_t* r, int nx, int ny, int nz) {
  const int ext[3] = {nx, ny, nz};
  BoxWalk w;
  w.n = 1;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int lo = r[d] > 0 ? r[d] : 0, hi = r[3 + d] < ext[d] - 1 ? r[3 + d] : ext[d] - 1;
    w.lo[d] = lo;
    w.e[d] = hi >= lo ? hi - lo + 1 : 0;
    w.n *= w.e[d];
  }
  return w;
}

// Linear index (in the ny x nz domain) of item b of walk w.
__device__ __forceinline__ long long box_item(const BoxWalk& w, long long b, int ny, int nz) {
  const long long k = b % w.e[2], r = b / w.e[2], j = r % w.e[1], i = r / w.e[1];
  return ((w.lo[0] + i) * ny + (w.lo[1] + j)) * (long long)nz + (w.lo[2] + k);
}

// Grows box record b's `cur` over the stencils (base .. base + 2 per axis)
// of the wave's valid lanes: wave reductions, then one atomic per bound.
__device__ __forceinline__ void wave_grow_box(int32_t* b, const int* base, bool valid) {
  int lo[3], hi[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    lo[d] = wave_min_i(valid ? base[d] : 0x7FFFFFFF);
    hi[d] = wave_max_i(valid ? base[d] + 2 : -0x7FFFFFFF - 1);
  }
  if ((threadIdx.x & 63) == 0 && lo[0] <= hi[0]) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      atomicMin(b + BOX_CUR + d, lo[d]);
      atomicMax(b + BOX_CUR + 3 + d, hi[d]);
    }
  }
}

// Stencil base nodes of particle q (valid) at grid spacing 1 / inv.
__device__ __forceinline__ void stencil_base(const float* x, long long P, long long q, float inv, int* base) {
  float fx, w[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) bspline(x[d * P + q] * inv, base[d], fx, w);
}

// FIXED: the substep flags common to the launch's instances (STORED_MOMENTS,
// MOVING_WALL, CELL_FORCE), compiled in as in k_lbm_substep; ~0u = take them
// from each instance at run time (any mix of modes).  A penalty instance's
// substeps after the first take STORED_MOMENTS per cell (batch_moment_source).
template <uint32_t FIXED>
__global__ void __launch_bounds__(TPB, 5) kb_substep(const FsiBatchInstance* __restrict__ inst, const int32_t* __restrict__ bmap, int t) {
  const unsigned b_ = xcd_tile(blockIdx.x, gridDim.x);
  const int i_ = bmap[2 * b_], l_ = bmap[2 * b_ + 1];
  const FsiBatchInstance& I = inst[i_];
  // references into the instance table, not local copies: a copy indexed at
  // run time (F[cur], link_c[parity]) would live in scratch memory
  const FsiLbmParams& p = I.lp;
  const FsiLbmState& st = I.ls;
  // the single-instance launch geometry: l_ = i * tiles + tile of the (j,k) plane
  const unsigned plane = (unsigned)p.ny * (unsigned)p.nz, tiles = (plane + TPB - 1) / TPB;
  const unsigned i = (unsigned)l_ / tiles;
  const unsigned jk = ((unsigned)l_ - i * tiles) * TPB + threadIdx.x;
  if (jk >= plane) return;
  const unsigned j = jk / (unsigned)p.nz, k = jk - j * (unsigned)p.nz;
  constexpr uint32_t VAR = FSI_SUB_STORED_MOMENTS | FSI_SUB_MOVING_WALL | FSI_SUB_CELL_FORCE;
  uint32_t flags = FIXED == ~0u ? batch_substep_flags(I, t) : ((batch_substep_flags(I, t) & ~VAR) | FIXED);
  uint32_t oflags = flags;
  if (FIXED == ~0u) {
    batch_moment_source(I, t, flags, oflags, (int)i, (int)j, (int)k);
  } else if (FIXED == FSI_SUB_CELL_FORCE) {  // the penalty launch of substeps t >= 1
    oflags = flags | FSI_SUB_STORED_MOMENTS;
    if (batch_penalty_stored(I, (int)i, (int)j, (int)k)) flags |= FSI_SUB_STORED_MOMENTS;
  }
  const int cur = batch_cur(I, t), parity = t & 1;
  if (p.storage == FSI_STORE_F16_DEV)
    lbm_substep_cell<PopF16, true>(p, st, cur, parity, flags, (int)i, (int)j, (int)k, oflags);
  else if (p.storage == FSI_STORE_BF16_DEV)
    lbm_substep_cell<PopBF16, true>(p, st, cur, parity, flags, (int)i, (int)j, (int)k, oflags);
  else
    lbm_substep_cell<PopF32, true>(p, st, cur, parity, flags, (int)i, (int)j, (int)k, oflags);
}

__global__ void __launch_bounds__(TPB) kb_rebuild(const FsiBatchInstance* __restrict__ inst, const int32_t* __restrict__ bmap) {
  FSI_BATCH_PROLOGUE(FSI_BATCH_DOMAIN_CELLS, TPB)
  if (!in_) return;
  const int dyn = I.mode == FSI_BATCH_MODE_MOVING_BOUNDARY ? 1 : 0;
  FSI_BY_STORAGE(I.lp.storage, lbm_rebuild_cell<POP>(I.lp, I.ls, dyn, I.cur0, e_));
}

__global__ void __launch_bounds__(TPB) kb_moments(const FsiBatchInstance* __restrict__ inst, const int32_t* __restrict__ bmap, int t) {
  FSI_BATCH_PROLOGUE(FSI_BATCH_DOMAIN_CELLS, TPB)
  if (!in_) return;
  // moments after substep t: the populations are in F[cur0 ^ ((t + 1) & 1)]
  const uint32_t flags = I.cell_force ? FSI_SUB_CELL_FORCE : 0u;
  FSI_BY_STORAGE(I.lp.storage, lbm_moments_cell<POP>(I.lp, I.ls, batch_cur(I, t + 1), flags, e_));
}

__global__ void __launch_bounds__(TPB) kb_mb_begin(const FsiBatchInstance* __restrict__ inst, const int32_t* __restrict__ bmap) {
  FSI_BATCH_PROLOGUE(FSI_BATCH_DOMAIN_CELLS, TPB)
  if (e_ == 0) {
    FsiLbmCounters* k = I.ls.counters;
    k->bb_link_count = 0;
    k->bb_max_correction_bits = 0;
    k->mb_sample_count = 0;
    k->accum_norm_max_bits = 0;
    k->mean_norm_max_bits = 0;
    k->bb_net_solid_force[0] = k->bb_net_solid_force[1] = k->bb_net_solid_force[2] = 0.0;
  }
  if (in_) lbm_mb_begin_cell(I.lp, I.ls, e_);
}

__global__ void __launch_bounds__(TPB) kb_finalize(const FsiBatchInstance* __restrict__ inst, const int32_t* __restrict__ bmap) {
  __shared__ double sd[7 * TPB / 64];
  FSI_BATCH_PROLOGUE(FSI_BATCH_DOMAIN_CELLS, TPB)
  FinalizeOut o;
  o.links = 0;
  o.max_corr = o.mean_norm = o.runmax = 0.0f;
  o.hx = o.hy = o.hz = 0.0;
  const int last = I.n_substeps - 1;
  if (in_)
    FSI_BY_STORAGE(I.lp.storage, o = lbm_finalize_cell<POP>(I.lp, I.ls, batch_cur(I, I.n_substeps), last & 1,
                                                            I.n_substeps, 1, e_));
  finalize_block_record(o, I.ls.reduce_ws + (size_t)l_ * FIN_REC, sd);
}

// one workgroup per instance (INSTANCES domain): fold the finalize records
__global__ void __launch_bounds__(TPB) kb_finalize_fold(const FsiBatchInstance* __restrict__ inst, const int32_t* __restrict__ bmap) {
  __shared__ double sd[TPB / 64];
  const FsiBatchInstance& I = inst[bmap[2 * blockIdx.x]];
  const int n_rec = (int)batch_domain_blocks(I, FSI_BATCH_DOMAIN_CELLS);
  finalize_fold(I.ls.reduce_ws, n_rec, I.ls.counters, 1, I.n_substeps, sd);
}

// one workgroup per instance (INSTANCES domain); t = 0: projection box, 1: MPM box
__global__ void __launch_bounds__(64) kb_box_roll(const FsiBatchInstance* __restrict__ inst, const int32_t* __restrict__ bmap, int t) {
  const FsiBatchInstance& I = inst[bmap[2 * blockIdx.x]];
  if (threadIdx.x != 0) return;
  if (t == 0)
    box_roll(I.boxes + FSI_BATCH_BOX_PROJ, I.lp.nx, I.lp.ny, I.lp.nz);
  else
    box_roll(I.boxes + FSI_BATCH_BOX_MPM, I.mp.gnx, I.mp.gny, I.mp.gnz);
}

// LBM_RESIDENT: the moving-wall subcycle of the one instance the block map
// names, block b of it at map entry b, as k_lbm_resident.
__global__ void __launch_bounds__(RES_THREADS, 1) kb_lbm_resident(const FsiBatchInstance* __restrict__ inst,
                                                                  const int32_t* __restrict__ bmap, int) {
  const FsiBatchInstance& I = inst[bmap[2 * blockIdx.x]];
  res_subcycle_block<false>(I.lp, I.ls, I.cur0, I.n_substeps, BATCH_RESIDENT_FLAGS, I.res_halo, I.res_sync,
                            (uint32_t)bmap[2 * blockIdx.x + 1]);
}

// One workgroup per instance of the map (INSTANCES domain), its first lane.
__global__ void __launch_bounds__(64) kb_counters_clear(const FsiBatchInstance* inst, const int32_t* bmap, int which) {
  if (threadIdx.x == 0) batch_clear_counters(inst[bmap[2 * blockIdx.x]], which);
}

__global__ void __launch_bounds__(TPB) kb_proj_clear(const FsiBatchInstance* __restrict__ inst, const int32_t* __restrict__ bmap) {
  FSI_BATCH_BOX_PROLOGUE
  const BoxWalk w = box_walk(I.boxes + FSI_BATCH_BOX_PROJ + BOX_PREV, I.lp.nx, I.lp.ny, I.lp.nz);
  const long long N = batch_domain_size(I, FSI_BATCH_DOMAIN_CELLS);
  for (long long b = (long long)l_ * TPB + threadIdx.x; b < w.n; b += stride_) {
    const long long c = box_item(w, b, I.lp.ny, I.lp.nz);
    I.ls.solid_phi[c] = 0.0f;
    I.ls.solid_mass[c] = 0.0f;
    I.ls.solid_vel[c] = I.ls.solid_vel[N + c] = I.ls.solid_vel[2 * N + c] = 0.0f;
  }
}

__global__ void __launch_bounds__(PTPB) kb_proj_particles(const FsiBatchInstance* __restrict__ inst, const int32_t* __restrict__ bmap) {
  __shared__ __align__(16) float tile[TILE_COPIES * 5 * PROJ_TILE_NODES];
  FSI_BATCH_PROLOGUE(FSI_BATCH_DOMAIN_PARTICLES, PTPB)
  int base[3] = {0, 0, 0};
  if (in_) stencil_base(I.ms.x, I.mp.n_particles, e_, I.cp.inv_dx_norm, base);
  wave_grow_box(I.boxes + FSI_BATCH_BOX_PROJ, base, in_);
  proj_block(I.lp, I.ls, I.mp, I.ms, I.cp, e_, in_, tile);
}

__global__ void __launch_bounds__(TPB) kb_proj_normalize(const FsiBatchInstance* __restrict__ inst, const int32_t* __restrict__ bmap) {
  __shared__ double sd[5 * TPB / 64];
  FSI_BATCH_BOX_PROLOGUE
  const BoxWalk w = box_walk(I.boxes + FSI_BATCH_BOX_PROJ + BOX_CUR, I.lp.nx, I.lp.ny, I.lp.nz);
  float mp = 0.0f;
  int act = 0;
  double m = 0.0, vr = 0.0, vc = 0.0;
  for (long long b = (long long)l_ * TPB + threadIdx.x; b < w.n; b += stride_) {
    const ProjOut o = normalize_cell(I.lp, I.ls, I.cp, box_item(w, b, I.lp.ny, I.lp.nz));
    mp = o.phi_raw > mp ? o.phi_raw : mp;
    act += o.active;
    m += (double)o.mass;
    vr += (double)o.phi_raw * I.cp.cell_volume_norm;
    vc += (double)o.phi_clamped * I.cp.cell_volume_norm;
  }
  double v[5] = {(double)mp, (double)act, m, vr, vc};
  constexpr int op[5] = {RED_MAXF0, RED_SUMI, RED_SUM, RED_SUM, RED_SUM};
  block_reduce_n(v, op, sd);
  mp = (float)v[0];
  act = (int)v[1];
  m = v[2];
  vr = v[3];
  vc = v[4];
  if (threadIdx.x == 0) {
    FsiCouplingCounters* cc = I.cc;
    if (mp > 0.0f) atomicMax(&cc->max_phi_raw_bits, __float_as_uint(mp));
    if (act) atomicAdd(&cc->active_cell_count, act);
    if (m != 0.0) atomicAdd(&cc->projected_mass, m);
    if (vr != 0.0) atomicAdd(&cc->projected_volume_raw, vr);
    if (vc != 0.0) atomicAdd(&cc->projected_volume_clamped, vc);
  }
}

// Penalty pass of LBM step t.  From t = 1 on it first stores the moments of
// the cells it covers, from the populations after the last step and that
// step's cell force, as the moments pass after that step would have.
__global__ void __launch_bounds__(TPB) kb_penalty(const FsiBatchInstance* __restrict__ inst, const int32_t* __restrict__ bmap, int t) {
  __shared__ double sd[5 * TPB / 64];
  FSI_BATCH_BOX_PROLOGUE
  int32_t both[6];
  box_hull(I.boxes + FSI_BATCH_BOX_PROJ + BOX_CUR, I.boxes + FSI_BATCH_BOX_PROJ + BOX_PREV, both);
  const BoxWalk w = box_walk(both, I.lp.nx, I.lp.ny, I.lp.nz);
  float mfn = 0.0f;
  int act = 0;
  double net[3] = {0.0, 0.0, 0.0};
  if (t > 0 && I.lp.bc_x_right == FSI_BC_PRESSURE) {
    // the plane the pressure outlet reads, outside the hull: its moments too
    const long long plane = (long long)I.lp.ny * I.lp.nz, i0 = (long long)(I.lp.nx - 2) * plane;
    for (long long b = (long long)l_ * TPB + threadIdx.x; b < plane; b += stride_) {
      const int j = (int)(b / I.lp.nz), k = (int)(b % I.lp.nz);
      if (!batch_penalty_stored(I, I.lp.nx - 2, j, k))
        FSI_BY_STORAGE(I.lp.storage,
                       lbm_moments_cell<POP>(I.lp, I.ls, batch_cur(I, t), FSI_SUB_CELL_FORCE, i0 + b));
    }
  }
  for (long long b = (long long)l_ * TPB + threadIdx.x; b < w.n; b += stride_) {
    const long long c = box_item(w, b, I.lp.ny, I.lp.nz);
    if (t > 0) FSI_BY_STORAGE(I.lp.storage, lbm_moments_cell<POP>(I.lp, I.ls, batch_cur(I, t), FSI_SUB_CELL_FORCE, c));
    float f[3];
    int active;
    penalty_cell(I.lp, I.ls, I.cp, c, f, active);
    const float fn = norm3(f[0], f[1], f[2]);
    mfn = fn > mfn ? fn : mfn;
    act += active;
    for (int d = 0; d < 3; ++d) net[d] += (double)f[d];
  }
  double v[5] = {(double)mfn, (double)act, net[0], net[1], net[2]};
  constexpr int op[5] = {RED_MAXF0, RED_SUMI, RED_SUM, RED_SUM, RED_SUM};
  block_reduce_n(v, op, sd);
  mfn = (float)v[0];
  act = (int)v[1];
  for (int d = 0; d < 3; ++d) net[d] = v[2 + d];
  if (threadIdx.x == 0) {
    FsiCouplingCounters* cc = I.cc;
    if (act) atomicAdd(&cc->active_force_cell_count, act);
    if (mfn > 0.0f) {
      atomicMax(&cc->max_cell_force_bits, __float_as_uint(mfn));
      atomicMax(&cc->max_hydro_force_bits, __float_as_uint(mfn));
    }
    for (int d = 0; d < 3; ++d)
      if (net[d] != 0.0) {
        atomicAdd(&cc->net_cell_force[d], net[d]);
        atomicAdd(&cc->net_hydro_force[d], -net[d]);
      }
  }
}

__global__ void __launch_bounds__(TPB) kb_mpm_clear(const FsiBatchInstance* __restrict__ inst, const int32_t* __restrict__ bmap) {
  FSI_BATCH_BOX_PROLOGUE
  const BoxWalk w = box_walk(I.boxes + FSI_BATCH_BOX_MPM + BOX_PREV, I.mp.gnx, I.mp.gny, I.mp.gnz);
  const long long G = batch_domain_size(I, FSI_BATCH_DOMAIN_NODES);
  for (long long b = (long long)l_ * TPB + threadIdx.x; b < w.n; b += stride_) {
    const long long n = box_item(w, b, I.mp.gny, I.mp.gnz);
    I.ms.grid_v[n] = I.ms.grid_v[G + n] = I.ms.grid_v[2 * G + n] = 0.0f;
    I.ms.grid_m[n] = 0.0f;
    I.ms.grid_f[n] = I.ms.grid_f[G + n] = I.ms.grid_f[2 * G + n] = 0.0f;
  }
}

__global__ void __launch_bounds__(PTPB) kb_mpm_p2g(const FsiBatchInstance* __restrict__ inst, const int32_t* __restrict__ bmap) {
  __shared__ float tile[TILE_COPIES * 4 * TILE_NODES];
  FSI_BATCH_PROLOGUE(FSI_BATCH_DOMAIN_PARTICLES, PTPB)
  int base[3] = {0, 0, 0};
  if (in_) stencil_base(I.ms.x, I.mp.n_particles, e_, I.mp.inv_dx, base);
  wave_grow_box(I.boxes + FSI_BATCH_BOX_MPM, base, in_);
  p2g_block(I.mp, I.ms, e_, in_, tile);
}

__global__ void __launch_bounds__(PTPB) kb_mpm_reaction(const FsiBatchInstance* __restrict__ inst, const int32_t* __restrict__ bmap) {
  __shared__ double sd[PTPB / 64];
  __shared__ float sf[PTPB / 64];
  __shared__ int si[PTPB / 64];
  __shared__ float tile[TILE_COPIES * 3 * TILE_NODES];
  FSI_BATCH_PROLOGUE(FSI_BATCH_DOMAIN_PARTICLES, PTPB)
  const int mode = I.mode == FSI_BATCH_MODE_MOVING_BOUNDARY ? 2 : 1;
  int base[3] = {0, 0, 0};
  if (in_) stencil_base(I.ms.x, I.mp.n_particles, e_, I.cp.inv_dx_norm, base);
  wave_grow_box(I.boxes + FSI_BATCH_BOX_MPM, base, in_ && I.mode != FSI_BATCH_MODE_NONE);
  const ReactOut o = reaction_block(I.lp, I.ls, I.mp, I.ms, I.cp, mode, e_, in_ && I.mode != FSI_BATCH_MODE_NONE,
                                    tile);
  const int act = block_sum_i(o.active, si);
  const float mpn = block_max_f(o.particle_norm, sf);
  const float mgc = block_max_f(o.max_contrib, sf);
  double net[3];
  for (int d = 0; d < 3; ++d) net[d] = block_reduce((double)o.f[d], RED_SUM, sd);
  if (threadIdx.x == 0) {
    FsiCouplingCounters* cc = I.cc;
    if (act) atomicAdd(&cc->active_reaction_particle_count, act);
    if (mpn > 0.0f) atomicMax(&cc->max_particle_reaction_bits, __float_as_uint(mpn));
    if (mgc > 0.0f) atomicMax(&cc->max_grid_reaction_bits, __float_as_uint(mgc));
    for (int d = 0; d < 3; ++d)
      if (net[d] != 0.0) atomicAdd(&cc->net_particle_reaction[d], net[d]);
  }
}

__global__ void __launch_bounds__(PTPB) kb_mpm_p2g_react(const FsiBatchInstance* __restrict__ inst, const int32_t* __restrict__ bmap) {
  __shared__ __align__(16) float tile[TILE_COPIES * 4 * FUSED_TILE_NODES];
  __shared__ double sd[PTPB / 64];
  __shared__ float sf[PTPB / 64];
  __shared__ int si[PTPB / 64];
  FSI_BATCH_PROLOGUE(FSI_BATCH_DOMAIN_PARTICLES, PTPB)
  int base[3] = {0, 0, 0};
  if (in_) stencil_base(I.ms.x, I.mp.n_particles, e_, I.mp.inv_dx, base);
  wave_grow_box(I.boxes + FSI_BATCH_BOX_MPM, base, in_);
  if (I.mode != FSI_BATCH_MODE_NONE && I.cp.inv_dx_norm != I.mp.inv_dx) {
    // the reaction scatters on its own stencil into grid_f
    if (in_) stencil_base(I.ms.x, I.mp.n_particles, e_, I.cp.inv_dx_norm, base);
    wave_grow_box(I.boxes + FSI_BATCH_BOX_MPM, base, in_);
  }
  double* rec = I.react_ws + (size_t)l_ * REACT_REC;
  if (I.mode == FSI_BATCH_MODE_MOVING_BOUNDARY)
    p2g_react_block<2>(I.lp, I.ls, I.mp, I.ms, I.cp, rec, e_, in_, tile, sd, sf, si);
  else if (I.mode == FSI_BATCH_MODE_PENALTY)
    p2g_react_block<1>(I.lp, I.ls, I.mp, I.ms, I.cp, rec, e_, in_, tile, sd, sf, si);
  else
    p2g_react_block<0>(I.lp, I.ls, I.mp, I.ms, I.cp, rec, e_, in_, tile, sd, sf, si);
}

__global__ void __launch_bounds__(TPB) kb_mpm_grid_update(const FsiBatchInstance* __restrict__ inst, const int32_t* __restrict__ bmap) {
  FSI_BATCH_BOX_PROLOGUE
  const BoxWalk w = box_walk(I.boxes + FSI_BATCH_BOX_MPM + BOX_CUR, I.mp.gnx, I.mp.gny, I.mp.gnz);
  for (long long b = (long long)l_ * TPB + threadIdx.x; b < w.n; b += stride_)
    mpm_grid_node(I.mp, I.ms, box_item(w, b, I.mp.gny, I.mp.gnz));
}

__global__ void __launch_bounds__(TPB) kb_mpm_grid_update_react(const FsiBatchInstance* __restrict__ inst, const int32_t* __restrict__ bmap) {
  __shared__ double sd[TPB / 64];
  FSI_BATCH_BOX_PROLOGUE
  const BoxWalk w = box_walk(I.boxes + FSI_BATCH_BOX_MPM + BOX_CUR, I.mp.gnx, I.mp.gny, I.mp.gnz);
  for (long long b = (long long)l_ * TPB + threadIdx.x; b < w.n; b += stride_)
    mpm_grid_node(I.mp, I.ms, box_item(w, b, I.mp.gny, I.mp.gnz));
  if (l_ == 0 && I.mode != FSI_BATCH_MODE_NONE && I.mp.n_particles > 0)
    fold_react_records(I.react_ws, (int)((I.mp.n_particles + PTPB - 1) / PTPB), I.cc, sd);
}

__global__ void __launch_bounds__(PTPB) kb_mpm_g2p(const FsiBatchInstance* __restrict__ inst, const int32_t* __restrict__ bmap) {
  FSI_BATCH_PROLOGUE(FSI_BATCH_DOMAIN_PARTICLES, PTPB)
  if (in_) mpm_g2p_particle(I.mp, I.ms, e_);
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
namespace {
// Kernel of a batched operation: its entry and workgroup size.
struct BatchKernel {
  const void* fn;
  int block;
};

int batch_kernel(int op, int t, BatchKernel& k) {
  k.block = TPB;
  switch (op) {
    case FSI_BATCH_OP_COUNTERS_CLEAR: k.fn = (const void*)&kb_counters_clear; k.block = 64; break;
    case FSI_BATCH_OP_LBM_SUBSTEP: k.fn = (const void*)&kb_substep<~0u>; break;
    case FSI_BATCH_OP_LBM_SUBSTEP_MB:
      k.fn = t == 0 ? (const void*)&kb_substep<FSI_SUB_MOVING_WALL | FSI_SUB_STORED_MOMENTS>
                    : (const void*)&kb_substep<FSI_SUB_MOVING_WALL>;
      break;
    case FSI_BATCH_OP_LBM_SUBSTEP_PENALTY:
      k.fn = t == 0 ? (const void*)&kb_substep<FSI_SUB_CELL_FORCE | FSI_SUB_STORED_MOMENTS>
                    : (const void*)&kb_substep<FSI_SUB_CELL_FORCE>;
      break;
    case FSI_BATCH_OP_LBM_SUBSTEP_NONE:
      k.fn = t == 0 ? (const void*)&kb_substep<FSI_SUB_STORED_MOMENTS> : (const void*)&kb_substep<0u>;
      break;
    case FSI_BATCH_OP_LBM_REBUILD: k.fn = (const void*)&kb_rebuild; break;
    case FSI_BATCH_OP_LBM_MOMENTS: k.fn = (const void*)&kb_moments; break;
    case FSI_BATCH_OP_LBM_MB_BEGIN: k.fn = (const void*)&kb_mb_begin; break;
    case FSI_BATCH_OP_LBM_MB_FINALIZE: k.fn = (const void*)&kb_finalize; break;
    case FSI_BATCH_OP_LBM_MB_FINALIZE_FOLD: k.fn = (const void*)&kb_finalize_fold; break;
    case FSI_BATCH_OP_PROJ_CLEAR: k.fn = (const void*)&kb_proj_clear; break;
    case FSI_BATCH_OP_PROJ_PARTICLES: k.fn = (const void*)&kb_proj_particles; k.block = PTPB; break;
    case FSI_BATCH_OP_PROJ_NORMALIZE: k.fn = (const void*)&kb_proj_normalize; break;
    case FSI_BATCH_OP_PENALTY: k.fn = (const void*)&kb_penalty; break;
    case FSI_BATCH_OP_MPM_CLEAR: k.fn = (const void*)&kb_mpm_clear; break;
    case FSI_BATCH_OP_MPM_P2G: k.fn = (const void*)&kb_mpm_p2g; k.block = PTPB; break;
    case FSI_BATCH_OP_MPM_REACTION: k.fn = (const void*)&kb_mpm_reaction; k.block = PTPB; break;
    case FSI_BATCH_OP_MPM_GRID_UPDATE: k.fn = (const void*)&kb_mpm_grid_update; break;
    case FSI_BATCH_OP_MPM_G2P: k.fn = (const void*)&kb_mpm_g2p; k.block = PTPB; break;
    case FSI_BATCH_OP_MPM_P2G_REACT: k.fn = (const void*)&kb_mpm_p2g_react; k.block = PTPB; break;
    case FSI_BATCH_OP_MPM_GRID_UPDATE_REACT: k.fn = (const void*)&kb_mpm_grid_update_react; break;
    case FSI_BATCH_OP_BOX_ROLL: k.fn = (const void*)&kb_box_roll; k.block = 64; break;
    case FSI_BATCH_OP_LBM_RESIDENT: k.fn = (const void*)&kb_lbm_resident; k.block = RES_THREADS; break;
    default: return fail(FSI_E_INVALID, "unknown batch operation");
  }
  return FSI_OK;
}

// How many workgroups of the resident subcycle the current device holds at
// once (CUs x workgroups per CU), cached per device.
int res_capacity(bool force, long long& capacity) {
  static thread_local int dev_cached = -1;
  static thread_local long long cap[2] = {0, 0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fail(FSI_E_HIP, "hipGetDevice failed");
  if (dev != dev_cached) {
    int cus = 0, per_cu[3] = {0, 0, 0};
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu[0], k_lbm_resident<false>, RES_THREADS, 0) !=
            hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu[1], k_lbm_resident<true>, RES_THREADS, 0) !=
            hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu[2], kb_lbm_resident, RES_THREADS, 0) != hipSuccess)
      return fail(FSI_E_HIP, "occupancy query of the resident subcycle failed");
    // the batched kernel runs the unforced subcycle: both hold the smaller count
    cap[0] = (long long)cus * (per_cu[0] < per_cu[2] ? per_cu[0] : per_cu[2]);
    cap[1] = (long long)cus * per_cu[1];
    dev_cached = dev;
  }
  capacity = cap[force ? 1 : 0];
  return FSI_OK;
}

// An LBM_RESIDENT launch must fit the device at once (its block map holds one
// instance's blocks).
int check_resident_blocks(int op, int n_blocks) {
  if (op != FSI_BATCH_OP_LBM_RESIDENT) return FSI_OK;
  long long capacity = 0;
  if (const int rc = res_capacity(false, capacity)) return rc;
  if (n_blocks > capacity)
    return fail(FSI_E_UNSUPPORTED, "LBM_RESIDENT with " + std::to_string(n_blocks) +
                                       " workgroups, the device holds " + std::to_string(capacity));
  return FSI_OK;
}

// Argument values of a batched kernel and the pointer array a launch takes:
// (inst, map[, t]).
struct BatchArgs {
  const FsiBatchInstance* inst;
  const int32_t* map;
  int t;
  void* ptrs[3];
  BatchArgs(const FsiBatchInstance* i, const int32_t* m, int t_) : inst(i), map(m), t(t_) {
    ptrs[0] = &inst;
    ptrs[1] = &map;
    ptrs[2] = &t;
  }
};

struct BatchGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const FsiBatchInstance* inst = nullptr;
  int n_inst = 0;
  std::vector<hipGraphNode_t> nodes;
  std::vector<FsiBatchOp> ops;
};

void destroy_graph(BatchGraph* g) {
  if (!g) return;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  delete g;
}

// Dependencies of a graph node on `branch` (FsiBatchOp.branch), given the
// last node of the main chain (last[0]) and of each side chain so far: a main
// node follows the main chain, a side node its side chain (or, as its first
// node, the main chain), a join node all of them.
int batch_node_deps(int branch, const hipGraphNode_t* last, hipGraphNode_t* deps) {
  int n = 0;
  if (branch > FSI_BATCH_BRANCH_MAIN) {
    if (last[branch]) deps[n++] = last[branch];
    else if (last[0]) deps[n++] = last[0];
    return n;
  }
  if (last[0]) deps[n++] = last[0];
  if (branch == FSI_BATCH_BRANCH_JOIN)
    for (int c = 1; c <= FSI_BATCH_SIDE_CHAINS; ++c)
      if (last[c]) deps[n++] = last[c];
  return n;
}

// Kernel node parameters of op `o`; `a` holds the argument values (and must
// outlive the call that takes `p`).
void node_params(const FsiBatchOp& o, BatchArgs& a, hipKernelNodeParams& p) {
  BatchKernel k;
  batch_kernel(o.op, o.t, k);
  p.func = const_cast<void*>(k.fn);
  p.gridDim = dim3((unsigned)o.n_blocks);
  p.blockDim = dim3(k.block);
  p.sharedMemBytes = 0;
  p.kernelParams = a.ptrs;
  p.extra = nullptr;
}

// A launch list the graph can take: known operations with a block map and
// workgroups; for an update, the same operations and substep indices as the
// graph was built from.
int check_graph_ops(const BatchGraph* g, const FsiBatchOp* ops, int n_ops, bool update) {
  if (update && (size_t)n_ops != g->ops.size())
    return fail(FSI_E_INVALID, "batch graph update with another number of operations");
  for (int i = 0; i < n_ops; ++i) {
    BatchKernel k;
    if (batch_kernel(ops[i].op, ops[i].t, k) != FSI_OK) return FSI_E_INVALID;
    if (!ops[i].block_map || ops[i].n_blocks <= 0)
      return fail(FSI_E_INVALID, "batch graph operation without a block map or workgroups");
    if (const int rc = check_resident_blocks(ops[i].op, ops[i].n_blocks)) return rc;
    if (ops[i].branch < FSI_BATCH_BRANCH_JOIN || ops[i].branch > FSI_BATCH_SIDE_CHAINS)
      return fail(FSI_E_INVALID, "unknown batch graph branch");
    if (update && (ops[i].op != g->ops[(size_t)i].op || ops[i].t != g->ops[(size_t)i].t ||
                   ops[i].branch != g->ops[(size_t)i].branch))
      return fail(FSI_E_INVALID, "batch graph update with other operations");
  }
  return FSI_OK;
}
}  // namespace

namespace {
// Checks of a binning's buffers; the bin count must fit the int32 keys.
int check_bins(const FsiBins* b, long long nbins) {
  if (!b || !b->key || !b->start || !b->order || !b->work) return fail(FSI_E_INVALID, "null binning buffers");
  if (nbins + 2 > 0x7fffffffLL) return fail(FSI_E_INVALID, "grid too large for 32-bit bin keys");
  return FSI_OK;
}

// Clears the bin counts (before the keys are counted).
void bins_begin(long long nbins, const FsiBins& b, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_i32, dim3(blocks_for(nbins + 1)), dim3(TPB), 0, s, b.work, nbins + 1, 0);
}

// After the keys are counted (b.work[0 .. nbins]): scan into start, fill
// order, sort each bin by particle index.
void bins_finish(long long nbins, int32_t P, const FsiBins& b, hipStream_t s) {
  const long long n = nbins + 1;  // with the overflow bin
  const int tiles = (int)((n + SCAN_TILE - 1) / SCAN_TILE);
  int32_t* count = b.work;
  int32_t* part = b.work + n;
  hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(TPB), 0, s, count, n, b.start, part);
  hipLaunchKernelGGL(k_scan_parts, dim3(1), dim3(TPB), 0, s, part, tiles, b.start, n);
  hipLaunchKernelGGL(k_scan_add, dim3(blocks_for(n)), dim3(TPB), 0, s, b.start, n, (const int32_t*)part);
  if (P > 0) {
    hipLaunchKernelGGL(k_bin_fill, dim3(blocks_for(P)), dim3(TPB), 0, s, (const int32_t*)b.key, P,
                       (const int32_t*)b.start, count, b.order);
    hipLaunchKernelGGL(k_bin_sort, dim3(blocks_for(n)), dim3(TPB), 0, s, (const int32_t*)b.start, n, b.order);
  }
}
}  // namespace

// Scratch of fsi_mpm_substep_coupled_noalloc: room for the reaction records
// (FSI_COUPLED_WORKSPACE_DOUBLES) of FSI_NOALLOC_MAX_PARTICLES particles,
// 2.25 MiB, held by every GPU the library runs on until the process ends.  In
// host code `scratch` only names the symbol; its address on each device comes
// from hipGetSymbolAddress, kept per device below.
static __device__ __align__(256) double scratch[FSI_COUPLED_WORKSPACE_DOUBLES(FSI_NOALLOC_MAX_PARTICLES)];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[kMaxDevices] = {};

// The current device's address of `scratch`, looked up on the first call on
// that device (with it current) and kept.
int noalloc_scratch_ptr(double** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(FSI_E_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return fail(FSI_E_INVALID, "device index out of range for the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(FSI_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_dev[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_dev[dev];
  return FSI_OK;
}
}  // namespace

extern "C" {

const char* fsi_last_error(void) { return g_last_error.c_str(); }

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

int fsi_sync(void* stream) {
  hipError_t e = hipStreamSynchronize(as_stream(stream));
  if (e != hipSuccess) return fail(FSI_E_HIP, std::string("hipStreamSynchronize: ") + hipGetErrorString(e));
  return FSI_OK;
}

int fsi_lbm_init(const FsiLbmParams* p, const FsiLbmState* st, void* stream) {
  int rc = check_lbm(p, st);
  if (rc) return rc;
  const long long N = (long long)p->nx * p->ny * p->nz;
  FSI_BY_STORAGE(p->storage, hipLaunchKernelGGL(k_lbm_init<POP>, dim3(blocks_for(N)), dim3(TPB), 0, as_stream(stream), *p, *st));
  return check_launch("k_lbm_init");
}

int fsi_lbm_rebuild(const FsiLbmParams* p, const FsiLbmState* st, int dynamic, int cur, void* stream) {
  int rc = check_lbm(p, st);
  if (rc) return rc;
  if (cur != 0 && cur != 1) return fail(FSI_E_INVALID, "cur must be 0 or 1");
  if (dynamic && (!st->static_solid || !st->old_solid || !st->solid_phi || !st->solid_vel))
    return fail(FSI_E_INVALID, "dynamic rebuild needs static_solid, old_solid, solid_phi, solid_vel");
  if (!st->solid_vel) return fail(FSI_E_INVALID, "rebuild needs solid_vel");
  const long long N = (long long)p->nx * p->ny * p->nz;
  FSI_BY_STORAGE(p->storage, hipLaunchKernelGGL(k_lbm_rebuild<POP>, dim3(blocks_for(N)), dim3(TPB), 0, as_stream(stream), *p,
                                                *st, dynamic, cur));
  return check_launch("k_lbm_rebuild");
}

int fsi_lbm_substep(const FsiLbmParams* p, const FsiLbmState* st, int cur, int parity, uint32_t flags,
                    void* stream) {
  int rc = check_lbm(p, st);
  if (rc) return rc;
  if ((cur != 0 && cur != 1) || (parity != 0 && parity != 1))
    return fail(FSI_E_INVALID, "cur and parity must be 0 or 1");
  const bool mb = flags & FSI_SUB_MOVING_WALL;
  if (mb && (!st->link_c[0] || !st->link_c[1] || !st->solid_vel || !st->hydro_accum))
    return fail(FSI_E_INVALID, "moving-wall substep needs link_c, solid_vel, hydro_accum");
  if (mb && (flags & FSI_SUB_DIAG) && !st->link_corr)
    return fail(FSI_E_INVALID, "diagnostic substep needs link_corr");
  if ((flags & FSI_SUB_CELL_FORCE) && !st->cell_force) return fail(FSI_E_INVALID, "cell_force is null");
  const int bs = p->launch_block ? p->launch_block : TPB;
  if (bs != 128 && bs != 256 && bs != 512) return fail(FSI_E_INVALID, "launch_block must be 128, 256 or 512");
  if (bs != TPB && p->storage != FSI_STORE_F32)
    return fail(FSI_E_INVALID, "launch_block other than 256 is built for FP32 storage only");
  if (p->launch_minw != 0 && p->launch_minw != 1 && p->launch_minw != 5)
    return fail(FSI_E_INVALID, "launch_minw must be 0, 1 or 5");
  const unsigned tiles = (unsigned)(((long long)p->ny * p->nz + bs - 1) / bs);
  const dim3 grid(tiles * (unsigned)p->nx), block(bs);
  hipStream_t s = as_stream(stream);
  const uint32_t fixed = flags & (FSI_SUB_STORED_MOMENTS | FSI_SUB_MOVING_WALL | FSI_SUB_CELL_FORCE);
  const bool one_generation = (long long)grid.x * (bs / 64) <= 4LL * simds_on_device();
  const bool uncapped = p->launch_minw ? p->launch_minw == 1 : one_generation;
  switch (fixed) {
#define FSI_LAUNCH(V, W)                                                                                         \
  if (bs == 128)                                                                                                 \
    hipLaunchKernelGGL((k_lbm_substep<PopF32, (V), W, 128>), grid, block, 0, s, *p, *st, cur, parity, flags, tiles); \
  else if (bs == 512)                                                                                            \
    hipLaunchKernelGGL((k_lbm_substep<PopF32, (V), W, 512>), grid, block, 0, s, *p, *st, cur, parity, flags, tiles); \
  else                                                                                                           \
    FSI_BY_STORAGE(p->storage, hipLaunchKernelGGL((k_lbm_substep<POP, (V), W, 256>), grid, block, 0, s, *p, *st, cur, \
                                                  parity, flags, tiles));
#define FSI_CASE(V)        \
  case (V):                \
    if (uncapped) {        \
      FSI_LAUNCH(V, 1)     \
    } else {               \
      FSI_LAUNCH(V, 5)     \
    }                      \
    break;
    FSI_CASE(0u)
    FSI_CASE(1u)
    FSI_CASE(2u)
    FSI_CASE(3u)
    FSI_CASE(16u)
    FSI_CASE(17u)
    FSI_CASE(18u)
    FSI_CASE(19u)
#undef FSI_CASE
#undef FSI_LAUNCH
    default:
      return fail(FSI_E_INVALID, "unsupported substep flag combination");
  }
  return check_launch("k_lbm_substep");
}

int fsi_lbm_subcycle_resident(const FsiLbmParams* p, const FsiLbmState* st, int cur, int n, uint32_t flags,
                              float* halo, uint32_t* sync, void* stream) {
  int rc = check_lbm(p, st);
  if (rc) return rc;
  if (const char* why = res_unsupported(*p, *st, cur, n, flags, halo, sync)) return fail(FSI_E_UNSUPPORTED, why);
  if (!st->solid_vel) return fail(FSI_E_INVALID, "moving-wall subcycle needs solid_vel");
  const bool force = (flags & FSI_SUB_CELL_FORCE) != 0;
  if (force && !st->cell_force) return fail(FSI_E_INVALID, "cell_force is null");
  long long capacity = 0;
  if ((rc = res_capacity(force, capacity)) != FSI_OK) return rc;
  const int nblk = res_blocks(*p);
  if (nblk > capacity)
    return fail(FSI_E_UNSUPPORTED, "the lattice needs " + std::to_string(nblk) +
                                       " resident workgroups, the device holds " + std::to_string(capacity));
  hipStream_t s = as_stream(stream);
  if (force)
    hipLaunchKernelGGL(k_lbm_resident<true>, dim3(nblk), dim3(RES_THREADS), 0, s, *p, *st, cur, n, flags, halo, sync);
  else
    hipLaunchKernelGGL(k_lbm_resident<false>, dim3(nblk), dim3(RES_THREADS), 0, s, *p, *st, cur, n, flags, halo, sync);
  return check_launch("k_lbm_resident");
}

int fsi_lbm_open_bc(const FsiLbmParams* p, const FsiLbmState* st, int buf, uint32_t flags, void* stream) {
  int rc = check_lbm(p, st);
  if (rc) return rc;
  if (buf != 0 && buf != 1) return fail(FSI_E_INVALID, "buf must be 0 or 1");
  if (!st->counters) return fail(FSI_E_INVALID, "open_bc needs counters");
  if ((flags & FSI_SUB_CELL_FORCE) && !st->cell_force) return fail(FSI_E_INVALID, "cell_force is null");
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(k_open_bc_clear, dim3(1), dim3(64), 0, s, *st);
  const dim3 grid(blocks_for((long long)p->ny * p->nz), 2);
  FSI_BY_STORAGE(p->storage, hipLaunchKernelGGL(k_lbm_open_bc<POP>, grid, dim3(TPB), 0, s, *p, *st, buf, flags));
  return check_launch("k_lbm_open_bc");
}

int fsi_lbm_link_stats(const FsiLbmParams* p, const FsiLbmState* st, int cur, int parity, double* workspace_dev,
                       double* out_dev, void* stream) {
  int rc = check_lbm(p, st);
  if (rc) return rc;
  if ((cur != 0 && cur != 1) || (parity != 0 && parity != 1))
    return fail(FSI_E_INVALID, "cur and parity must be 0 or 1");
  if (!st->link_c[0] || !st->link_c[1] || !st->link_corr || !workspace_dev || !out_dev)
    return fail(FSI_E_INVALID, "link_stats needs link buffers, link_corr, workspace, out");
  hipStream_t s = as_stream(stream);
  FSI_BY_STORAGE(p->storage, hipLaunchKernelGGL(k_link_stats_partial<POP>, dim3(FSI_LINK_STATS_BLOCKS, 18), dim3(TPB), 0,
                                                s, *p, *st, cur, parity, workspace_dev));
  hipLaunchKernelGGL(k_link_stats_fold, dim3(1), dim3(64), 0, s, (const double*)workspace_dev,
                     (int)FSI_LINK_STATS_BLOCKS, out_dev);
  return check_launch("k_link_stats");
}

int fsi_lbm_moments(const FsiLbmParams* p, const FsiLbmState* st, int cur, uint32_t flags, void* stream) {
  int rc = check_lbm(p, st);
  if (rc) return rc;
  if (cur != 0 && cur != 1) return fail(FSI_E_INVALID, "cur must be 0 or 1");
  if ((flags & FSI_SUB_CELL_FORCE) && !st->cell_force) return fail(FSI_E_INVALID, "cell_force is null");
  const long long N = (long long)p->nx * p->ny * p->nz;
  FSI_BY_STORAGE(p->storage, hipLaunchKernelGGL(k_lbm_moments<POP>, dim3(blocks_for(N)), dim3(TPB), 0, as_stream(stream), *p,
                                                *st, cur, flags));
  return check_launch("k_lbm_moments");
}

int fsi_lbm_mb_begin(const FsiLbmParams* p, const FsiLbmState* st, void* stream) {
  int rc = check_lbm(p, st);
  if (rc) return rc;
  if (!st->hydro_accum || !st->counters)
    return fail(FSI_E_INVALID, "mb_begin needs hydro_accum, counters");
  const long long N = (long long)p->nx * p->ny * p->nz;
  hipLaunchKernelGGL(k_lbm_mb_begin, dim3(blocks_for(N)), dim3(TPB), 0, as_stream(stream), *p, *st);
  return check_launch("k_lbm_mb_begin");
}

int fsi_lbm_mb_finalize(const FsiLbmParams* p, const FsiLbmState* st, int cur, int parity, int n_samples,
                        int subcycled, void* stream) {
  int rc = check_lbm(p, st);
  if (rc) return rc;
  if ((cur != 0 && cur != 1) || (parity != 0 && parity != 1))
    return fail(FSI_E_INVALID, "cur and parity must be 0 or 1");
  if (!st->link_c[0] || !st->link_c[1] || !st->link_corr || !st->hydro_accum || !st->hydro_force || !st->counters ||
      !st->reduce_ws)
    return fail(FSI_E_INVALID, "mb_finalize needs link buffers, accumulator, hydro_force, counters, reduce_ws");
  if (n_samples < 1) return fail(FSI_E_INVALID, "n_samples must be positive");
  const long long N = (long long)p->nx * p->ny * p->nz;
  hipStream_t s = as_stream(stream);
  FSI_BY_STORAGE(p->storage, hipLaunchKernelGGL(k_lbm_finalize<POP>, dim3(blocks_for(N)), dim3(TPB), 0, s, *p, *st,
                                                cur, parity, n_samples, subcycled));
  hipLaunchKernelGGL(k_lbm_finalize_fold, dim3(1), dim3(TPB), 0, s, *st, (int)blocks_for(N), subcycled, n_samples);
  return check_launch("k_lbm_finalize");
}

int fsi_lbm_reduce_stats(const FsiLbmParams* p, const FsiLbmState* st, FsiLbmStats* workspace_dev,
                         FsiLbmStats* out_dev, void* stream) {
  int rc = check_lbm(p, st);
  if (rc) return rc;
  if (!workspace_dev || !out_dev || !st->solid_phi || !st->cell_force || !st->hydro_force)
    return fail(FSI_E_INVALID, "reduce_stats needs workspace, out, solid_phi, cell_force, hydro_force");
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(k_lbm_stats_partial, dim3(FSI_REDUCE_BLOCKS), dim3(TPB), 0, s, *p, *st, workspace_dev);
  hipLaunchKernelGGL(k_stats_final, dim3(LBM_STATS_N), dim3(64), 0, s, reinterpret_cast<const double*>(workspace_dev),
                     FSI_REDUCE_BLOCKS, LBM_STATS_N, 0, reinterpret_cast<double*>(out_dev));
  return check_launch("k_lbm_stats");
}

int fsi_project(const FsiLbmParams* lp, const FsiLbmState* ls, const FsiMpmParams* mp, const FsiMpmState* ms,
                const FsiCouplingParams* cp, FsiCouplingCounters* cc, void* stream) {
  int rc = check_lbm(lp, ls);
  if (rc) return rc;
  rc = check_mpm(mp, ms);
  if (rc) return rc;
  if (!cp || !cc || !ls->solid_phi || !ls->solid_mass || !ls->solid_vel)
    return fail(FSI_E_INVALID, "project needs coupling params/counters and projection fields");
  const long long N = (long long)lp->nx * lp->ny * lp->nz;
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(k_proj_clear, dim3(blocks_for(N)), dim3(TPB), 0, s, *lp, *ls, cc);
  if (mp->n_particles > 0)
    hipLaunchKernelGGL(k_proj_particles, dim3(pblocks_for(mp->n_particles)), dim3(PTPB), 0, s, *lp, *ls, *mp, *ms, *cp);
  hipLaunchKernelGGL(k_proj_normalize, dim3(blocks_for(N)), dim3(TPB), 0, s, *lp, *ls, *cp, cc);
  return check_launch("fsi_project");
}

int fsi_penalty_force(const FsiLbmParams* lp, const FsiLbmState* ls, const FsiCouplingParams* cp,
                      FsiCouplingCounters* cc, void* stream) {
  int rc = check_lbm(lp, ls);
  if (rc) return rc;
  if (!cp || !cc || !ls->cell_force || !ls->hydro_force || !ls->solid_phi || !ls->solid_vel)
    return fail(FSI_E_INVALID, "penalty needs coupling params/counters and force fields");
  const long long N = (long long)lp->nx * lp->ny * lp->nz;
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(k_penalty_clear, dim3(1), dim3(64), 0, s, cc);
  hipLaunchKernelGGL(k_penalty, dim3(blocks_for(N)), dim3(TPB), 0, s, *lp, *ls, *cp, cc);
  return check_launch("fsi_penalty_force");
}

int fsi_mpm_clear_grid(const FsiMpmParams* mp, const FsiMpmState* ms, void* stream) {
  int rc = check_mpm(mp, ms);
  if (rc) return rc;
  const long long G = (long long)mp->gnx * mp->gny * mp->gnz;
  hipLaunchKernelGGL(k_mpm_clear, dim3(blocks_for(G)), dim3(TPB), 0, as_stream(stream), *mp, *ms);
  return check_launch("fsi_mpm_clear_grid");
}

int fsi_mpm_p2g(const FsiMpmParams* mp, const FsiMpmState* ms, void* stream) {
  int rc = check_mpm(mp, ms);
  if (rc) return rc;
  hipStream_t s = as_stream(stream);
  if (mp->n_particles > 0)
    hipLaunchKernelGGL(k_mpm_p2g, dim3(pblocks_for(mp->n_particles)), dim3(PTPB), 0, s, *mp, *ms);
  return check_launch("fsi_mpm_p2g");
}

int fsi_mpm_reaction(const FsiLbmParams* lp, const FsiLbmState* ls, const FsiMpmParams* mp, const FsiMpmState* ms,
                     const FsiCouplingParams* cp, FsiCouplingCounters* cc, int mode, void* stream) {
  int rc = check_lbm(lp, ls);
  if (rc) return rc;
  rc = check_mpm(mp, ms);
  if (rc) return rc;
  if (mode != 1 && mode != 2) return fail(FSI_E_INVALID, "reaction mode must be 1 (penalty) or 2 (moving boundary)");
  if (!cp || !cc || !ls->hydro_force || !ls->solid_phi) return fail(FSI_E_INVALID, "reaction needs hydro_force, solid_phi");
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(k_reaction_clear, dim3(1), dim3(64), 0, s, cc);
  if (mp->n_particles > 0)
    hipLaunchKernelGGL(k_mpm_reaction, dim3(pblocks_for(mp->n_particles)), dim3(PTPB), 0, s, *lp, *ls, *mp, *ms, *cp, cc,
                       mode);
  return check_launch("fsi_mpm_reaction");
}

int fsi_mpm_substep_coupled(const FsiLbmParams* lp, const FsiLbmState* ls, const FsiMpmParams* mp,
                            const FsiMpmState* ms, const FsiCouplingParams* cp, FsiCouplingCounters* cc, int mode,
                            double* workspace, void* stream) {
  int rc = check_mpm(mp, ms);
  if (rc) return rc;
  if (mode < 0 || mode > 2) return fail(FSI_E_INVALID, "mode must be 0 (none), 1 (penalty) or 2 (moving boundary)");
  FsiLbmParams lpv;
  FsiLbmState lsv;
  FsiCouplingParams cpv;
  memset(&lpv, 0, sizeof(lpv));
  memset(&lsv, 0, sizeof(lsv));
  memset(&cpv, 0, sizeof(cpv));
  if (mode) {
    rc = check_lbm(lp, ls);
    if (rc) return rc;
    if (!cp || !cc || !workspace || !ls->hydro_force || !ls->solid_phi)
      return fail(FSI_E_INVALID, "the reaction needs coupling params, counters, workspace, hydro_force, solid_phi");
    lpv = *lp;
    lsv = *ls;
    cpv = *cp;
  }
  hipStream_t s = as_stream(stream);
  const long long G = (long long)mp->gnx * mp->gny * mp->gnz;
  const unsigned pb = pblocks_for(mp->n_particles);
  hipLaunchKernelGGL(k_mpm_clear, dim3(blocks_for(G)), dim3(TPB), 0, s, *mp, *ms);
  if (mp->n_particles > 0) {
    if (mode == 1)
      hipLaunchKernelGGL(k_mpm_p2g_react<1>, dim3(pb), dim3(PTPB), 0, s, lpv, lsv, *mp, *ms, cpv, workspace);
    else if (mode == 2)
      hipLaunchKernelGGL(k_mpm_p2g_react<2>, dim3(pb), dim3(PTPB), 0, s, lpv, lsv, *mp, *ms, cpv, workspace);
    else
      hipLaunchKernelGGL(k_mpm_p2g_react<0>, dim3(pb), dim3(PTPB), 0, s, lpv, lsv, *mp, *ms, cpv, workspace);
  }
  hipLaunchKernelGGL(k_mpm_grid_update_react, dim3(blocks_for(G)), dim3(TPB), 0, s, *mp, *ms,
                     (mode && mp->n_particles > 0) ? (const double*)workspace : (const double*)nullptr, (int)pb, cc);
  if (mode && mp->n_particles == 0) hipLaunchKernelGGL(k_reaction_clear, dim3(1), dim3(64), 0, s, cc);
  if (mp->n_particles > 0)
    hipLaunchKernelGGL(k_mpm_g2p, dim3(pb), dim3(PTPB), 0, s, *mp, *ms);
  return check_launch("fsi_mpm_substep_coupled");
}

int fsi_mpm_substep_coupled_noalloc(const FsiLbmParams* lp, const FsiLbmState* ls, const FsiMpmParams* mp,
                                    const FsiMpmState* ms, const FsiCouplingParams* cp, FsiCouplingCounters* cc,
                                    int mode, void* stream) {
  if (!mp) return fail(FSI_E_INVALID, "null MPM params");
  if (mp->n_particles > FSI_NOALLOC_MAX_PARTICLES)
    return fail(FSI_E_INVALID, "fsi_mpm_substep_coupled_noalloc: " + std::to_string(mp->n_particles) +
                                   " particles exceed its scratch limit of " +
                                   std::to_string(FSI_NOALLOC_MAX_PARTICLES) +
                                   " (use fsi_mpm_substep_coupled with a caller workspace)");
  double* ws = nullptr;
  if (mode) {
    const int rc = noalloc_scratch_ptr(&ws);
    if (rc) return rc;
  }
  return fsi_mpm_substep_coupled(lp, ls, mp, ms, cp, cc, mode, ws, stream);
}

int fsi_mpm_subcycle_coupled(const FsiLbmParams* lp, const FsiLbmState* ls, const FsiMpmParams* mp,
                             const FsiMpmState* ms, const FsiCouplingParams* cp, FsiCouplingCounters* cc, int mode,
                             int n, double* workspace, float* scratch, uint32_t* sync, void* stream) {
  static_assert(MPM_RES_STRIDE * MPM_RES_TIMEOUT == FSI_MPM_RES_TIMEOUT_WORD, "timeout word");
  static_assert(MPM_RES_STRIDE * (MPM_RES_TIMEOUT + 1) == FSI_MPM_RES_SYNC_WORDS, "sync words");
  int rc = check_mpm(mp, ms);
  if (rc) return rc;
  if (mode < 0 || mode > 2) return fail(FSI_E_INVALID, "mode must be 0 (none), 1 (penalty) or 2 (moving boundary)");
  if (n < 1) return fail(FSI_E_INVALID, "n must be at least 1");
  if (!scratch || !sync) return fail(FSI_E_INVALID, "the resident MPM subcycle needs scratch and sync");
  FsiLbmParams lpv;
  FsiLbmState lsv;
  FsiCouplingParams cpv;
  memset(&lpv, 0, sizeof(lpv));
  memset(&lsv, 0, sizeof(lsv));
  memset(&cpv, 0, sizeof(cpv));
  if (mode) {
    rc = check_lbm(lp, ls);
    if (rc) return rc;
    if (!cp || !cc || !workspace || !ls->hydro_force || !ls->solid_phi)
      return fail(FSI_E_INVALID, "the reaction needs coupling params, counters, workspace, hydro_force, solid_phi");
    if (cp->inv_dx_norm != mp->inv_dx)
      return fail(FSI_E_UNSUPPORTED, "the reaction samples the lattice at another spacing than the MPM grid");
    lpv = *lp;
    lsv = *ls;
    cpv = *cp;
  }
  if (mp->n_particles <= 0) return fail(FSI_E_UNSUPPORTED, "no particles");
  // every workgroup must be resident at once
  static thread_local int dev_cached = -1;
  static thread_local long long capacity[3] = {0, 0, 0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fail(FSI_E_HIP, "hipGetDevice failed");
  if (dev != dev_cached) {
    int cus = 0, per_cu[3] = {0, 0, 0};
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu[0], k_mpm_resident<0>, PTPB, 0) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu[1], k_mpm_resident<1>, PTPB, 0) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu[2], k_mpm_resident<2>, PTPB, 0) != hipSuccess)
      return fail(FSI_E_HIP, "occupancy query of the resident MPM subcycle failed");
    for (int m = 0; m < 3; ++m) capacity[m] = (long long)cus * per_cu[m];
    dev_cached = dev;
  }
  const unsigned pb = pblocks_for(mp->n_particles);
  if ((long long)pb > capacity[mode])
    return fail(FSI_E_UNSUPPORTED, std::to_string(pb) + " particle workgroups exceed the " +
                                       std::to_string(capacity[mode]) + " the device holds at once");
  hipStream_t s = as_stream(stream);
  if (mode == 1)
    hipLaunchKernelGGL(k_mpm_resident<1>, dim3(pb), dim3(PTPB), 0, s, lpv, lsv, *mp, *ms, cpv, n, scratch, workspace,
                       cc, sync);
  else if (mode == 2)
    hipLaunchKernelGGL(k_mpm_resident<2>, dim3(pb), dim3(PTPB), 0, s, lpv, lsv, *mp, *ms, cpv, n, scratch, workspace,
                       cc, sync);
  else
    hipLaunchKernelGGL(k_mpm_resident<0>, dim3(pb), dim3(PTPB), 0, s, lpv, lsv, *mp, *ms, cpv, n, scratch,
                       (double*)nullptr, (FsiCouplingCounters*)nullptr, sync);
  return check_launch("k_mpm_resident");
}

int fsi_bin_particles(const float* x, int32_t n_particles, float inv_dx, int32_t nx, int32_t ny, int32_t nz,
                      const FsiBins* bins, void* stream) {
  if (nx < 1 || ny < 1 || nz < 1 || n_particles < 0 || (n_particles > 0 && !x))
    return fail(FSI_E_INVALID, "fsi_bin_particles: bad grid or particles");
  const BinGrid bg = make_bins(nx, ny, nz);
  int rc = check_bins(bins, bg.nb);
  if (rc) return rc;
  hipStream_t s = as_stream(stream);
  bins_begin(bg.nb, *bins, s);
  if (n_particles > 0)
    hipLaunchKernelGGL(k_bin_keys, dim3(blocks_for(n_particles)), dim3(TPB), 0, s, x, n_particles, inv_dx, bg,
                       bins->work, bins->key);
  bins_finish(bg.nb, n_particles, *bins, s);
  return check_launch("fsi_bin_particles");
}

int fsi_mpm_substep_ordered(const FsiLbmParams* lp, const FsiLbmState* ls, const FsiMpmParams* mp,
                            const FsiMpmState* ms, const FsiCouplingParams* cp, FsiCouplingCounters* cc, int mode,
                            const FsiBins* bins, float* terms, double* workspace, void* stream) {
  int rc = check_mpm(mp, ms);
  if (rc) return rc;
  if (mode < 0 || mode > 2) return fail(FSI_E_INVALID, "mode must be 0 (none), 1 (penalty) or 2 (moving boundary)");
  const BinGrid bg = make_bins(mp->gnx, mp->gny, mp->gnz);
  rc = check_bins(bins, bg.nb);
  if (rc) return rc;
  if (mp->n_particles > 0 && !terms) return fail(FSI_E_INVALID, "fsi_mpm_substep_ordered: null terms buffer");
  FsiLbmParams lpv;
  FsiLbmState lsv;
  FsiCouplingParams cpv;
  memset(&lpv, 0, sizeof(lpv));
  memset(&lsv, 0, sizeof(lsv));
  memset(&cpv, 0, sizeof(cpv));
  if (mode) {
    rc = check_lbm(lp, ls);
    if (rc) return rc;
    if (!cp || !cc || !workspace || !ls->hydro_force || !ls->solid_phi)
      return fail(FSI_E_INVALID, "the reaction needs coupling params, counters, workspace, hydro_force, solid_phi");
    if (cp->inv_dx_norm != mp->inv_dx)
      return fail(FSI_E_INVALID, "fsi_mpm_substep_ordered: the lattice spacing must be the MPM grid's");
    lpv = *lp;
    lsv = *ls;
    cpv = *cp;
  }
  hipStream_t s = as_stream(stream);
  const long long G = (long long)mp->gnx * mp->gny * mp->gnz;
  const unsigned pb = pblocks_for(mp->n_particles);
  bins_begin(bg.nb, *bins, s);
  if (mp->n_particles > 0) {
    if (mode == 1)
      hipLaunchKernelGGL(k_ordered_terms<1>, dim3(pb), dim3(PTPB), 0, s, lpv, lsv, *mp, *ms, cpv, terms, workspace, bg,
                         bins->work, bins->key);
    else if (mode == 2)
      hipLaunchKernelGGL(k_ordered_terms<2>, dim3(pb), dim3(PTPB), 0, s, lpv, lsv, *mp, *ms, cpv, terms, workspace, bg,
                         bins->work, bins->key);
    else
      hipLaunchKernelGGL(k_ordered_terms<0>, dim3(pb), dim3(PTPB), 0, s, lpv, lsv, *mp, *ms, cpv, terms, workspace, bg,
                         bins->work, bins->key);
  }
  bins_finish(bg.nb, mp->n_particles, *bins, s);
  if (mode && mp->n_particles > 0)
    hipLaunchKernelGGL(k_ordered_nodes<true>, dim3(blocks_for(G)), dim3(TPB), 0, s, *mp, *ms, bg,
                       (const int32_t*)bins->start, (const int32_t*)bins->order, (const float*)terms,
                       (const double*)workspace, (int)pb, cc);
  else
    hipLaunchKernelGGL(k_ordered_nodes<false>, dim3(blocks_for(G)), dim3(TPB), 0, s, *mp, *ms, bg,
                       (const int32_t*)bins->start, (const int32_t*)bins->order, (const float*)terms,
                       (const double*)nullptr, 0, cc);
  if (mode && mp->n_particles == 0) hipLaunchKernelGGL(k_reaction_clear, dim3(1), dim3(64), 0, s, cc);
  if (mp->n_particles > 0) hipLaunchKernelGGL(k_mpm_g2p, dim3(pb), dim3(PTPB), 0, s, *mp, *ms);
  return check_launch("fsi_mpm_substep_ordered");
}

int fsi_project_ordered(const FsiLbmParams* lp, const FsiLbmState* ls, const FsiMpmParams* mp,
                        const FsiMpmState* ms, const FsiCouplingParams* cp, FsiCouplingCounters* cc,
                        const FsiBins* bins, void* stream) {
  int rc = check_lbm(lp, ls);
  if (rc) return rc;
  rc = check_mpm(mp, ms);
  if (rc) return rc;
  if (!cp || !cc || !ls->solid_phi || !ls->solid_mass || !ls->solid_vel)
    return fail(FSI_E_INVALID, "project needs coupling params/counters and projection fields");
  const BinGrid bg = make_bins(lp->nx, lp->ny, lp->nz);
  rc = check_bins(bins, bg.nb);
  if (rc) return rc;
  const long long N = (long long)lp->nx * lp->ny * lp->nz;
  hipStream_t s = as_stream(stream);
  bins_begin(bg.nb, *bins, s);
  if (mp->n_particles > 0)
    hipLaunchKernelGGL(k_bin_keys, dim3(blocks_for(mp->n_particles)), dim3(TPB), 0, s, (const float*)ms->x,
                       mp->n_particles, cp->inv_dx_norm, bg, bins->work, bins->key);
  bins_finish(bg.nb, mp->n_particles, *bins, s);
  hipLaunchKernelGGL(k_proj_counters_clear, dim3(1), dim3(64), 0, s, cc);
  hipLaunchKernelGGL(k_project_ordered, dim3(blocks_for(N)), dim3(TPB), 0, s, *lp, *ls, *mp, *ms, *cp, bg,
                     (const int32_t*)bins->start, (const int32_t*)bins->order);
  hipLaunchKernelGGL(k_proj_normalize, dim3(blocks_for(N)), dim3(TPB), 0, s, *lp, *ls, *cp, cc);
  return check_launch("fsi_project_ordered");
}

int fsi_mpm_grid_update(const FsiMpmParams* mp, const FsiMpmState* ms, void* stream) {
  int rc = check_mpm(mp, ms);
  if (rc) return rc;
  const long long G = (long long)mp->gnx * mp->gny * mp->gnz;
  hipLaunchKernelGGL(k_mpm_grid_update, dim3(blocks_for(G)), dim3(TPB), 0, as_stream(stream), *mp, *ms);
  return check_launch("fsi_mpm_grid_update");
}

int fsi_mpm_g2p(const FsiMpmParams* mp, const FsiMpmState* ms, void* stream) {
  int rc = check_mpm(mp, ms);
  if (rc) return rc;
  if (!ms->fixed_mask || !ms->fixed_x) return fail(FSI_E_INVALID, "g2p needs fixed_mask and fixed_x");
  if (mp->n_particles > 0)
    hipLaunchKernelGGL(k_mpm_g2p, dim3(pblocks_for(mp->n_particles)), dim3(PTPB), 0, as_stream(stream), *mp, *ms);
  return check_launch("fsi_mpm_g2p");
}

int fsi_mpm_reduce_stats(const FsiMpmParams* mp, const FsiMpmState* ms, FsiMpmStats* workspace_dev,
                         FsiMpmStats* out_dev, void* stream) {
  int rc = check_mpm(mp, ms);
  if (rc) return rc;
  if (!workspace_dev || !out_dev) return fail(FSI_E_INVALID, "reduce_stats needs workspace and out");
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(k_mpm_stats_partial, dim3(FSI_REDUCE_BLOCKS), dim3(TPB), 0, s, *mp, *ms, workspace_dev);
  hipLaunchKernelGGL(k_stats_final, dim3(MPM_STATS_N), dim3(64), 0, s, reinterpret_cast<const double*>(workspace_dev),
                     FSI_REDUCE_BLOCKS, MPM_STATS_N, 1, reinterpret_cast<double*>(out_dev));
  return check_launch("k_mpm_stats");
}


int fsi_batch_block_map(const FsiBatchInstance* inst_host, int n_inst, int domain, int mode_mask, int min_mpm,
                        int32_t* map_host) {
  if (!inst_host || n_inst < 0) return fail(FSI_E_INVALID, "null batch arguments");
  if (domain < FSI_BATCH_DOMAIN_CELLS || domain > FSI_BATCH_DOMAIN_BOX)
    return fail(FSI_E_INVALID, "unknown batch domain");
  long long total = 0;
  for (int i = 0; i < n_inst; ++i) {
    const FsiBatchInstance& I = inst_host[i];
    if (I.lp.open_bc != FSI_OPEN_BC_EQUILIBRIUM)
      return fail(FSI_E_INVALID, "the batched form runs the legacy x-face reset only (open_bc 0)");
    if (I.active)
      if (const char* why = batch_resident_invalid(I)) return fail(FSI_E_INVALID, why);
    if (!I.active || !(mode_mask & (1 << I.mode)) || I.mpm_substeps <= min_mpm) continue;
    const long long nb = batch_domain_blocks(I, domain);
    if (total + nb > 0x7fffffffLL) return fail(FSI_E_INVALID, "batch too large for one launch");
    if (map_host)
      for (long long b = 0; b < nb; ++b) {
        map_host[2 * (total + b)] = i;
        map_host[2 * (total + b) + 1] = (int32_t)b;
      }
    total += nb;
  }
  return (int)total;
}

int fsi_batch_launch(int op, const FsiBatchInstance* inst_dev, int n_inst, const int32_t* block_map_dev, int n_blocks,
                     int t, void* stream) {
  if (!inst_dev || n_inst <= 0) return fail(FSI_E_INVALID, "empty batch");
  BatchKernel k;
  if (batch_kernel(op, t, k) != FSI_OK) return FSI_E_INVALID;
  if (const int rc = check_resident_blocks(op, n_blocks)) return rc;
  if (!block_map_dev) return fail(FSI_E_INVALID, "null block map");
  if (n_blocks <= 0) return FSI_OK;
  BatchArgs a(inst_dev, block_map_dev, t);
  const hipError_t e = hipLaunchKernel(k.fn, dim3((unsigned)n_blocks), dim3(k.block), a.ptrs, 0, as_stream(stream));
  if (e != hipSuccess) return fail(FSI_E_HIP, hipGetErrorString(e));
  return check_launch("fsi_batch_launch");
}

int fsi_batch_graph_create(const FsiBatchOp* ops, int n_ops, const FsiBatchInstance* inst_dev, int n_inst,
                           void** graph_out) {
  if (!graph_out) return fail(FSI_E_INVALID, "null graph handle");
  *graph_out = nullptr;
  if (!ops || n_ops <= 0 || !inst_dev || n_inst <= 0) return fail(FSI_E_INVALID, "empty batch graph");
  BatchGraph* g = new BatchGraph();
  g->inst = inst_dev;
  g->n_inst = n_inst;
  g->ops.assign(ops, ops + n_ops);
  int rc = check_graph_ops(g, ops, n_ops, false);
  hipError_t e = hipSuccess;
  if (rc == FSI_OK) e = hipGraphCreate(&g->graph, 0);
  hipGraphNode_t last[FSI_BATCH_SIDE_CHAINS + 1] = {};
  for (int i = 0; rc == FSI_OK && e == hipSuccess && i < n_ops; ++i) {
    hipKernelNodeParams p;
    BatchArgs a(inst_dev, ops[i].block_map, ops[i].t);
    node_params(ops[i], a, p);
    hipGraphNode_t deps[FSI_BATCH_SIDE_CHAINS + 1];
    const int nd = batch_node_deps(ops[i].branch, last, deps);
    hipGraphNode_t node;
    e = hipGraphAddKernelNode(&node, g->graph, nd ? deps : nullptr, (size_t)nd, &p);
    if (e != hipSuccess) break;
    g->nodes.push_back(node);
    if (ops[i].branch > FSI_BATCH_BRANCH_MAIN) {
      last[ops[i].branch] = node;
    } else {
      last[0] = node;
      if (ops[i].branch == FSI_BATCH_BRANCH_JOIN)
        for (int c = 1; c <= FSI_BATCH_SIDE_CHAINS; ++c) last[c] = nullptr;
    }
  }
  if (rc == FSI_OK && e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  if (rc == FSI_OK && e != hipSuccess) rc = fail(FSI_E_HIP, hipGetErrorString(e));
  if (rc != FSI_OK) {
    destroy_graph(g);
    return rc;
  }
  *graph_out = g;
  return FSI_OK;
}

int fsi_batch_graph_update(void* graph, const FsiBatchOp* ops, int n_ops) {
  BatchGraph* g = static_cast<BatchGraph*>(graph);
  if (!g || !ops) return fail(FSI_E_INVALID, "null batch graph");
  const int rc = check_graph_ops(g, ops, n_ops, true);
  if (rc != FSI_OK) return rc;
  for (int i = 0; i < n_ops; ++i) {
    const FsiBatchOp& o = g->ops[(size_t)i];
    if (o.block_map == ops[i].block_map && o.n_blocks == ops[i].n_blocks) continue;
    hipKernelNodeParams p;
    BatchArgs a(g->inst, ops[i].block_map, ops[i].t);
    node_params(ops[i], a, p);
    const hipError_t e = hipGraphExecKernelNodeSetParams(g->exec, g->nodes[(size_t)i], &p);
    if (e != hipSuccess) return fail(FSI_E_HIP, hipGetErrorString(e));
    g->ops[(size_t)i] = ops[i];
  }
  return FSI_OK;
}

int fsi_batch_graph_launch(void* graph, void* stream) {
  BatchGraph* g = static_cast<BatchGraph*>(graph);
  if (!g) return fail(FSI_E_INVALID, "null batch graph");
  const hipError_t e = hipGraphLaunch(g->exec, as_stream(stream));
  if (e != hipSuccess) return fail(FSI_E_HIP, hipGetErrorString(e));
  return FSI_OK;
}

int fsi_batch_graph_destroy(void* graph) {
  destroy_graph(static_cast<BatchGraph*>(graph));
  return FSI_OK;
}

}  // extern "C"
