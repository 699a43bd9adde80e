// fsi_oracle.cpp - CPU restatement of the reference FSI substep.
//
// TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg load this library, as the reference answer or the
// timed CPU baseline; the product path never calls it.
//
// It restates lizhuoh9/MPM-LBM's Taichi kernels pass by pass, in the
// reference's own structure (separate colission / streaming / BC / streaming3
// passes over AoS (nx,ny,nz,19) fields, push streaming with the f -> F copy,
// dense 19x19 M and inv_M, sequential sums where the reference uses atomics),
// templated on the float type so the same code runs in FP32 (parity) and FP64
// (precision reference).  Each function cites the reference lines it follows;
// paths are relative to src/mpm_lbm/sim/ of the reference.
//
// Built with g++ -O2 (no -ffast-math); OpenMP only parallelises loops whose
// iterations are independent (no reductions), so results do not depend on
// the thread count.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <vector>

// ORACLE_MUTANT plants one known mistake for the tests that show the
// closed-form checks catch it (tests/test_oracle_regularized_bspline.py builds
// each variant into a temporary directory).  0, the default, is the oracle.
//   1  regularised collision: sign of the trace term of (e e - I/3) : Pi flipped
//   2  regularised collision: 9/2 Hermite factor written as 3
//   3  regularised collision: non-equilibrium relaxed by (1 - s_v/2)
//   4  regularised collision: Guo prefactor (1 - s_v/2) dropped
//   5  B-spline: middle weight centred on fx - 0.5
//   6  B-spline: outer weights swapped
#ifndef ORACLE_MUTANT
#define ORACLE_MUTANT 0
#endif

namespace {

constexpr int Q = 19;
// lbm/fluid.py:359-369
const int E[Q][3] = {{0, 0, 0},  {1, 0, 0},  {-1, 0, 0}, {0, 1, 0},   {0, -1, 0}, {0, 0, 1},  {0, 0, -1},
                     {1, 1, 0},  {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0},  {1, 0, 1},  {-1, 0, -1}, {1, 0, -1},
                     {-1, 0, 1}, {0, 1, 1},  {0, -1, -1}, {0, 1, -1}, {0, -1, 1}};
// lbm/fluid.py:252
const int LR[Q] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
// lbm/fluid.py:228-247
const int MI[Q][Q] = {
    {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1},
    {-1, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1},
    {1, -2, -2, -2, -2, -2, -2, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1},
    {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0},
    {0, -2, 2, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0},
    {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1},
    {0, 0, 0, -2, 2, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1},
    {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1},
    {0, 0, 0, 0, 0, -2, 2, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1},
    {0, 2, 2, -1, -1, -1, -1, 1, 1, 1, 1, 1, 1, 1, 1, -2, -2, -2, -2},
    {0, -2, -2, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, -2, -2, -2, -2},
    {0, 0, 0, 1, 1, -1, -1, 1, 1, 1, 1, -1, -1, -1, -1, 0, 0, 0, 0},
    {0, 0, 0, -1, -1, 1, 1, 1, 1, 1, 1, -1, -1, -1, -1, 0, 0, 0, 0},
    {0, 0, 0, 0, 0, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 0, 0, 0, 0},
    {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, -1, -1},
    {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0},
    {0, 0, 0, 0, 0, 0, 0, 1, -1, 1, -1, -1, 1, -1, 1, 0, 0, 0, 0},
    {0, 0, 0, 0, 0, 0, 0, -1, 1, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1},
    {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, -1, -1, 1, -1, 1, 1, -1}};

// inv_M = np.linalg.inv(M) in float64 (fluid.py:248), Gauss-Jordan with
// partial pivoting; the caller narrows it to the working type.
void invert_M(double out[Q][Q]) {
  double a[Q][2 * Q];
  for (int r = 0; r < Q; ++r)
    for (int c = 0; c < 2 * Q; ++c) a[r][c] = c < Q ? (double)MI[r][c] : (c - Q == r ? 1.0 : 0.0);
  for (int col = 0; col < Q; ++col) {
    int piv = col;
    for (int r = col + 1; r < Q; ++r)
      if (fabs(a[r][col]) > fabs(a[piv][col])) piv = r;
    if (piv != col)
      for (int c = 0; c < 2 * Q; ++c) std::swap(a[piv][c], a[col][c]);
    const double d = a[col][col];
    for (int c = 0; c < 2 * Q; ++c) a[col][c] /= d;
    for (int r = 0; r < Q; ++r) {
      if (r == col) continue;
      const double f = a[r][col];
      if (f == 0.0) continue;
      for (int c = 0; c < 2 * Q; ++c) a[r][c] -= f * a[col][c];
    }
  }
  for (int r = 0; r < Q; ++r)
    for (int c = 0; c < Q; ++c) out[r][c] = a[r][Q + c];
}

template <typename T>
struct Tables {
  T M[Q][Q], iM[Q][Q], w[Q], e[Q][3];
  Tables() {
    double inv[Q][Q];
    invert_M(inv);
    for (int r = 0; r < Q; ++r)
      for (int c = 0; c < Q; ++c) {
        M[r][c] = (T)MI[r][c];
        iM[r][c] = (T)inv[r][c];
      }
    // fluid.py:371-373
    for (int s = 0; s < Q; ++s) {
      w[s] = s == 0 ? (T)(1.0 / 3.0) : (s <= 6 ? (T)(1.0 / 18.0) : (T)(1.0 / 36.0));
      for (int d = 0; d < 3; ++d) e[s][d] = (T)E[s][d];
    }
  }
};

template <typename T>
const Tables<T>& tables() {
  static Tables<T> t;
  return t;
}

struct Dims {
  int nx, ny, nz;
  long long idx(int i, int j, int k) const { return ((long long)i * ny + j) * nz + k; }
  long long n() const { return (long long)nx * ny * nz; }
};

// fluid.py:321-326
template <typename T>
T feq(int k, T rho, const T* u) {
  const Tables<T>& tb = tables<T>();
  const T eu = tb.e[k][0] * u[0] + tb.e[k][1] * u[1] + tb.e[k][2] * u[2];
  const T uv = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  return tb.w[k] * rho * ((T)1.0 + (T)3.0 * eu + (T)4.5 * eu * eu - (T)1.5 * uv);
}

// fluid.py:420-429
inline void periodic(const Dims& d, int& i, int& j, int& k) {
  if (i < 0) i = d.nx - 1;
  if (i > d.nx - 1) i = 0;
  if (j < 0) j = d.ny - 1;
  if (j > d.ny - 1) j = 0;
  if (k < 0) k = d.nz - 1;
  if (k > d.nz - 1) k = 0;
}

template <typename T>
struct LbmArgs {
  int nx, ny, nz;
  T s_v, s_other, rho0;
  T fext[3];
  int bc_x_left, bc_x_right;
  T rho_bc_x_left, rho_bc_x_right;
  T vel_bc_x_left[3], vel_bc_x_right[3];
};

// init (fluid.py:329-344)
template <typename T>
void lbm_init(const LbmArgs<T>& a, T* F, T* f, T* rho, T* v) {
  const Dims d{a.nx, a.ny, a.nz};
  const T zero[3] = {0, 0, 0};
  for (long long c = 0; c < d.n(); ++c) {
    rho[c] = a.rho0;
    v[3 * c] = v[3 * c + 1] = v[3 * c + 2] = 0;
    for (int s = 0; s < Q; ++s) {
      f[c * Q + s] = feq<T>(s, a.rho0, zero);
      F[c * Q + s] = feq<T>(s, a.rho0, zero);
    }
  }
}

// colission (fluid.py:399-413) with meq_vec (:386-392), cal_local_force (:394-396)
template <typename T>
void colission(const LbmArgs<T>& a, const T* F, T* f, const T* rho, const T* v, const int8_t* solid,
               const T* cell_force) {
  const Tables<T>& tb = tables<T>();
  const Dims d{a.nx, a.ny, a.nz};
  const T S[Q] = {0, a.s_v, a.s_v, 0, a.s_other, 0, a.s_other, 0, a.s_other, a.s_v, a.s_v, a.s_v, a.s_v,
                  a.s_v, a.s_v, a.s_v, a.s_other, a.s_other, a.s_other};
#pragma omp parallel for schedule(static)
  for (long long c = 0; c < d.n(); ++c) {
    if (solid[c] != 0) continue;
    const T* Fc = F + c * Q;
    T m[Q];
    for (int r = 0; r < Q; ++r) {
      T acc = 0;
      for (int l = 0; l < Q; ++l) acc += tb.M[r][l] * Fc[l];
      m[r] = acc;
    }
    const T* u = v + 3 * c;
    T meq[Q] = {0};
    meq[0] = rho[c];
    meq[3] = u[0];
    meq[5] = u[1];
    meq[7] = u[2];
    meq[1] = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    meq[9] = (T)2 * u[0] * u[0] - u[1] * u[1] - u[2] * u[2];
    meq[11] = u[1] * u[1] - u[2] * u[2];
    meq[13] = u[0] * u[1];
    meq[14] = u[1] * u[2];
    meq[15] = u[0] * u[2];
    for (int r = 0; r < Q; ++r) m[r] -= S[r] * (m[r] - meq[r]);
    T frc[3];
    for (int k = 0; k < 3; ++k) frc[k] = a.fext[k] + (cell_force ? cell_force[3 * c + k] : (T)0);
    T g[Q];
    for (int l = 0; l < Q; ++l) {
      const T emu = (tb.e[l][0] - u[0]) * frc[0] + (tb.e[l][1] - u[1]) * frc[1] + (tb.e[l][2] - u[2]) * frc[2];
      const T eu = tb.e[l][0] * u[0] + tb.e[l][1] * u[1] + tb.e[l][2] * u[2];
      const T ef = tb.e[l][0] * frc[0] + tb.e[l][1] * frc[1] + tb.e[l][2] * frc[2];
      g[l] = tb.w[l] * (emu / (T)3.0 + (eu * ef) / (T)9.0);
    }
    for (int r = 0; r < Q; ++r) {
      T gm = 0;
      for (int l = 0; l < Q; ++l) gm += tb.M[r][l] * g[l];
      m[r] += ((T)1.0 - (T)0.5 * S[r]) * gm;
    }
    T* fc = f + c * Q;
    for (int r = 0; r < Q; ++r) {
      T acc = 0;
      for (int l = 0; l < Q; ++l) acc += tb.iM[r][l] * m[l];
      fc[r] = (T)0 + acc;
    }
  }
}

// Regularised BGK collision, the precision form's option (no reference
// counterpart), from J. Latt & B. Chopard, "Lattice Boltzmann method with
// regularized pre-collision distribution functions", Math. Comput. Simul. 72
// (2006) 165-168: f = feq + (1 - 1/tau) w (9/2) (e e - I/3) : Pi_neq, with
// Guo's BGK forcing term (1 - 1/(2 tau)) w [3 (e - u) + 9 (e.u) e] . G.
// tau = 1 / s_v; stored rho / v as in colission.
template <typename T>
void colission_regularized(const LbmArgs<T>& a, const T* F, T* f, const T* rho, const T* v, const int8_t* solid,
                           const T* cell_force) {
  const Tables<T>& tb = tables<T>();
  const Dims d{a.nx, a.ny, a.nz};
#pragma omp parallel for schedule(static)
  for (long long c = 0; c < d.n(); ++c) {
    if (solid[c] != 0) continue;
    const T* Fc = F + c * Q;
    const T* u = v + 3 * c;
    T fe[Q], Pi[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int s = 0; s < Q; ++s) {
      fe[s] = feq<T>(s, rho[c], u);
      const T dn = Fc[s] - fe[s];
      for (int p = 0; p < 3; ++p)
        for (int q = 0; q < 3; ++q) Pi[p][q] += tb.e[s][p] * tb.e[s][q] * dn;
    }
    const T tr = Pi[0][0] + Pi[1][1] + Pi[2][2];
    T G[3];
    for (int k = 0; k < 3; ++k) G[k] = a.fext[k] + (cell_force ? cell_force[3 * c + k] : (T)0);
    const T h2 = ORACLE_MUTANT == 2 ? (T)3 : (T)4.5;
    const T omega = ORACLE_MUTANT == 3 ? (T)1 - (T)0.5 * a.s_v : (T)1 - a.s_v;
    const T guo = ORACLE_MUTANT == 4 ? (T)1 : (T)1 - (T)0.5 * a.s_v;
    for (int s = 0; s < Q; ++s) {
      T qpi = ORACLE_MUTANT == 1 ? tr / (T)3 : -tr / (T)3;
      for (int p = 0; p < 3; ++p)
        for (int q = 0; q < 3; ++q) qpi += tb.e[s][p] * tb.e[s][q] * Pi[p][q];
      T out = fe[s] + omega * tb.w[s] * h2 * qpi;
      const T eu = tb.e[s][0] * u[0] + tb.e[s][1] * u[1] + tb.e[s][2] * u[2];
      T gterm = 0;
      for (int k = 0; k < 3; ++k) gterm += ((T)3 * (tb.e[s][k] - u[k]) + (T)9 * eu * tb.e[s][k]) * G[k];
      out += guo * tb.w[s] * gterm;
      f[c * Q + s] = out;
    }
  }
}

// The streaming target of fluid cell (i,j,k) in direction s: its periodic
// neighbour (fluid.py:420-429, :434-436, :493-495).  The link bounces back
// when that cell is solid (fluid.py:437-440, :496-519).
inline long long stream_target(const Dims& d, int i, int j, int k, int s) {
  int ii = i + E[s][0], jj = j + E[s][1], kk = k + E[s][2];
  periodic(d, ii, jj, kk);
  return d.idx(ii, jj, kk);
}

// streaming1: push with static half-way bounce-back (fluid.py:432-440)
template <typename T>
void streaming1(const LbmArgs<T>& a, const T* f, T* F, const int8_t* solid) {
  const Dims d{a.nx, a.ny, a.nz};
  for (int i = 0; i < d.nx; ++i)
    for (int j = 0; j < d.ny; ++j)
      for (int k = 0; k < d.nz; ++k) {
        const long long c = d.idx(i, j, k);
        if (solid[c] != 0) continue;
        for (int s = 0; s < Q; ++s) {
          const long long cp = stream_target(d, i, j, k, s);
          if (solid[cp] == 0)
            F[cp * Q + s] = f[c * Q + s];
          else
            F[c * Q + LR[s]] = f[c * Q + s];
        }
      }
}

template <typename T>
struct MbStats {
  long long link_count;
  T max_correction;
  T net_fluid_impulse[3];
  T net_solid_force[3];
  long long link_count_by_dir[Q];
  T fluid_impulse_by_dir[Q][3];
  T solid_force_by_dir[Q][3];
  T corr_abs_sum_by_dir[Q];
  T corr_abs_max_by_dir[Q];
};

// clear_moving_boundary_diagnostics + streaming_moving_bounceback +
// finalize_moving_boundary_diagnostics (fluid.py:444-458, :489-536).
// The reference's atomics become sequential sums in (i,j,k,s) order.
template <typename T>
void streaming_mb(const LbmArgs<T>& a, const T* f, T* F, const int8_t* solid, const T* rho, const T* solid_vel,
                  T* hydro_force, MbStats<T>* st) {
  const Tables<T>& tb = tables<T>();
  const Dims d{a.nx, a.ny, a.nz};
  memset(st, 0, sizeof(MbStats<T>));
  for (long long c = 0; c < d.n() * 3; ++c) hydro_force[c] = 0;
  for (int i = 0; i < d.nx; ++i)
    for (int j = 0; j < d.ny; ++j)
      for (int k = 0; k < d.nz; ++k) {
        const long long c = d.idx(i, j, k);
        if (solid[c] != 0) continue;
        for (int s = 0; s < Q; ++s) {
          const long long cp = stream_target(d, i, j, k, s);
          if (solid[cp] == 0) {
            F[cp * Q + s] = f[c * Q + s];
          } else {
            const T rl = rho[c];
            const T* uw = solid_vel + 3 * cp;
            const T corr = (T)-6.0 * tb.w[s] * rl * (tb.e[s][0] * uw[0] + tb.e[s][1] * uw[1] + tb.e[s][2] * uw[2]);
            const T bounced = f[c * Q + s] + corr;
            F[c * Q + LR[s]] = bounced;
            T fi[3];
            for (int q = 0; q < 3; ++q) {
              const T incoming = tb.e[s][q] * f[c * Q + s];
              const T outgoing = tb.e[LR[s]][q] * bounced;
              fi[q] = outgoing - incoming;
            }
            st->link_count_by_dir[s] += 1;
            const T ac = corr < 0 ? -corr : corr;
            st->corr_abs_sum_by_dir[s] += ac;
            st->corr_abs_max_by_dir[s] = std::max(st->corr_abs_max_by_dir[s], ac);
            for (int q = 0; q < 3; ++q) {
              hydro_force[3 * cp + q] += -fi[q];
              st->fluid_impulse_by_dir[s][q] += fi[q];
              st->solid_force_by_dir[s][q] += -fi[q];
            }
          }
        }
      }
  for (int s = 0; s < Q; ++s) {
    st->link_count += st->link_count_by_dir[s];
    st->max_correction = std::max(st->max_correction, st->corr_abs_max_by_dir[s]);
    for (int q = 0; q < 3; ++q) {
      st->net_fluid_impulse[q] += st->fluid_impulse_by_dir[s][q];
      st->net_solid_force[q] += st->solid_force_by_dir[s][q];
    }
  }
}

// Boundary_condition_legacy, x faces (fluid.py:927-958)
template <typename T>
void bc_legacy(const LbmArgs<T>& a, T* F, const T* v, const int8_t* solid) {
  const Dims d{a.nx, a.ny, a.nz};
  for (int j = 0; j < d.ny; ++j)
    for (int k = 0; k < d.nz; ++k) {
      const long long c0 = d.idx(0, j, k);
      if (a.bc_x_left == 1 && solid[c0] == 0) {
        const long long c1 = d.idx(1, j, k);
        const T* u = solid[c1] > 0 ? v + 3 * c1 : v + 3 * c0;
        for (int s = 0; s < Q; ++s) F[c0 * Q + s] = feq<T>(s, a.rho_bc_x_left, u);
      }
      if (a.bc_x_left == 2 && solid[c0] == 0)
        for (int s = 0; s < Q; ++s) F[c0 * Q + s] = feq<T>(s, (T)1.0, a.vel_bc_x_left);
    }
  for (int j = 0; j < d.ny; ++j)
    for (int k = 0; k < d.nz; ++k) {
      const long long cn = d.idx(d.nx - 1, j, k);
      if (a.bc_x_right == 1 && solid[cn] == 0) {
        const T* u = v + 3 * d.idx(d.nx - 2, j, k);
        for (int s = 0; s < Q; ++s) F[cn * Q + s] = feq<T>(s, a.rho_bc_x_right, u);
      }
      if (a.bc_x_right == 2 && solid[cn] == 0)
        for (int s = 0; s < Q; ++s) F[cn * Q + s] = feq<T>(s, (T)1.0, a.vel_bc_x_right);
    }
}

// Open boundaries of Boundary_condition (fluid.py:538-570) other than the
// legacy reset, on the x faces: the regularized reconstruction of the unknown
// populations (the ones streamed in across the face, fluid.py:26-27) from the
// inner neighbour's non-equilibrium part (fluid.py:1027-1030, :1349-1395), its
// limited form with clip counters (:1060-1104, :1397-1442), and the convective
// pressure outlet (:1106-1111 with :1051-1058, :1445-1493).  The neighbour's
// rho, v are those of the previous streaming3; its populations are the
// post-stream ones this step.
enum { OB_LEGACY = 0, OB_REGULARIZED = 1, OB_REGULARIZED_LIMITED = 2, OB_CONVECTIVE = 3 };

template <typename T>
struct ObArgs {
  int mode, limiter, floor_enabled;
  T rho_min, rho_max, u_max, noneq_cap, floor;
};

// rho clips, velocity clips, non-equilibrium clips, population floors,
// reconstructed populations (the reference's ob_*_count_step)
struct ObCounts {
  long long c[5];
};

static const int UNKNOWN_X_MIN[5] = {1, 7, 9, 11, 13};
static const int UNKNOWN_X_MAX[5] = {2, 8, 10, 12, 14};

// _regularized_population / _limited_regularized_population
template <typename T>
T ob_regularized(const ObArgs<T>& ob, bool limited, int s, T target_rho, const T* target_u, const T* Fn, T rho_n,
                 const T* v_n, ObCounts* k) {
  T rl = target_rho;
  T ul[3] = {target_u[0], target_u[1], target_u[2]};
  const bool lim = limited && ob.limiter;
  if (lim) {
    k->c[4] += 1;
    if (rl < ob.rho_min) {
      rl = ob.rho_min;
      k->c[0] += 1;
    }
    if (rl > ob.rho_max) {
      rl = ob.rho_max;
      k->c[0] += 1;
    }
    const T norm = sqrt(ul[0] * ul[0] + ul[1] * ul[1] + ul[2] * ul[2]);
    if (norm > ob.u_max) {
      const T sc = ob.u_max / norm;
      for (int q = 0; q < 3; ++q) ul[q] = ul[q] * sc;
      k->c[1] += 1;
    }
  }
  T noneq = Fn[s] - feq<T>(s, rho_n, v_n);
  if (lim) {
    if (noneq > ob.noneq_cap) {
      noneq = ob.noneq_cap;
      k->c[2] += 1;
    }
    if (noneq < -ob.noneq_cap) {
      noneq = -ob.noneq_cap;
      k->c[2] += 1;
    }
  }
  T out = feq<T>(s, rl, ul) + noneq;
  if (lim && ob.floor_enabled && out < ob.floor) {
    out = ob.floor;
    k->c[3] += 1;
  }
  return out;
}

template <typename T>
void bc_open(const LbmArgs<T>& a, const ObArgs<T>& ob, T* F, const T* rho, const T* v, const int8_t* solid,
             ObCounts* k) {
  const Dims d{a.nx, a.ny, a.nz};
  const bool limited = ob.mode == OB_REGULARIZED_LIMITED;
  for (int q = 0; q < 5; ++q) k->c[q] = 0;
  for (int j = 0; j < d.ny; ++j)
    for (int kk = 0; kk < d.nz; ++kk) {
      const long long c0 = d.idx(0, j, kk);
      if ((a.bc_x_left != 1 && a.bc_x_left != 2) || solid[c0] != 0) continue;
      const long long cn = solid[d.idx(1, j, kk)] > 0 ? c0 : d.idx(1, j, kk);
      T target_rho;
      const T* target_u;
      if (a.bc_x_left == 1) {
        target_rho = a.rho_bc_x_left;
        target_u = v + 3 * cn;
      } else {
        target_rho = rho[cn] <= (T)1.0e-6 ? a.rho0 : rho[cn];
        target_u = a.vel_bc_x_left;
      }
      for (int u = 0; u < 5; ++u) {
        const int s = UNKNOWN_X_MIN[u];
        F[c0 * Q + s] = ob_regularized<T>(ob, limited, s, target_rho, target_u, F + cn * Q, rho[cn], v + 3 * cn, k);
      }
    }
  for (int j = 0; j < d.ny; ++j)
    for (int kk = 0; kk < d.nz; ++kk) {
      const long long c0 = d.idx(d.nx - 1, j, kk);
      if ((a.bc_x_right != 1 && a.bc_x_right != 2) || solid[c0] != 0) continue;
      const bool inner_solid = solid[d.idx(d.nx - 2, j, kk)] > 0;
      const long long cn = inner_solid ? c0 : d.idx(d.nx - 2, j, kk);
      if (a.bc_x_right == 1 && ob.mode == OB_CONVECTIVE) {
        // _convective_outlet_population: linear extrapolation, then the floor
        const long long c2 = inner_solid ? c0 : d.idx(d.nx - 3, j, kk);
        for (int u = 0; u < 5; ++u) {
          const int s = UNKNOWN_X_MAX[u];
          T x = F[cn * Q + s];
          if (solid[c2] == 0) x = (T)2.0 * F[cn * Q + s] - F[c2 * Q + s];
          if (ob.limiter && ob.floor_enabled && x < ob.floor) x = ob.floor;
          F[c0 * Q + s] = x;
        }
        continue;
      }
      T target_rho;
      const T* target_u;
      if (a.bc_x_right == 1) {
        target_rho = a.rho_bc_x_right;
        target_u = v + 3 * cn;
      } else {
        target_rho = rho[cn] <= (T)1.0e-6 ? a.rho0 : rho[cn];
        target_u = a.vel_bc_x_right;
      }
      for (int u = 0; u < 5; ++u) {
        const int s = UNKNOWN_X_MAX[u];
        F[c0 * Q + s] = ob_regularized<T>(ob, limited, s, target_rho, target_u, F + cn * Q, rho[cn], v + 3 * cn, k);
      }
    }
}

// streaming3 (fluid.py:1795-1814)
template <typename T>
void streaming3(const LbmArgs<T>& a, const T* F, T* f, T* rho, T* v, const int8_t* solid, const T* cell_force) {
  const Tables<T>& tb = tables<T>();
  const Dims d{a.nx, a.ny, a.nz};
#pragma omp parallel for schedule(static)
  for (long long c = 0; c < d.n(); ++c) {
    if (solid[c] == 0) {
      T r = 0;
      T u[3] = {0, 0, 0};
      for (int s = 0; s < Q; ++s) f[c * Q + s] = F[c * Q + s];
      for (int s = 0; s < Q; ++s) r += f[c * Q + s];
      for (int s = 0; s < Q; ++s)
        for (int q = 0; q < 3; ++q) u[q] += tb.e[s][q] * f[c * Q + s];
      for (int q = 0; q < 3; ++q) {
        const T fq = a.fext[q] + (cell_force ? cell_force[3 * c + q] : (T)0);
        u[q] /= r;
        u[q] += (fq / (T)2) / r;
      }
      rho[c] = r;
      for (int q = 0; q < 3; ++q) v[3 * c + q] = u[q];
    } else {
      rho[c] = a.rho0;
      v[3 * c] = v[3 * c + 1] = v[3 * c + 2] = 0;
    }
  }
}

// update_dynamic_solid + reinitialize_new_fluid_cells (fluid.py:1848-1877)
template <typename T>
void dynamic_solid(const LbmArgs<T>& a, T threshold, const T* phi, const int8_t* static_solid, int8_t* solid,
                   int8_t* old_solid, int8_t* reinit, const T* solid_vel, T* F, T* f, T* rho, T* v) {
  const Dims d{a.nx, a.ny, a.nz};
  for (long long c = 0; c < d.n(); ++c) {
    const int8_t prev = solid[c];
    old_solid[c] = prev;
    const bool dyn = phi[c] >= threshold;
    const int8_t now = (static_solid[c] != 0 || dyn) ? 1 : 0;
    solid[c] = now;
    reinit[c] = (prev != 0 && now == 0) ? 1 : 0;
  }
  for (long long c = 0; c < d.n(); ++c) {
    if (reinit[c] != 1) continue;
    T rv[3] = {0, 0, 0};
    const T* sv = solid_vel + 3 * c;
    if (sqrt(sv[0] * sv[0] + sv[1] * sv[1] + sv[2] * sv[2]) > 0) {
      rv[0] = sv[0];
      rv[1] = sv[1];
      rv[2] = sv[2];
    }
    rho[c] = a.rho0;
    for (int q = 0; q < 3; ++q) v[3 * c + q] = rv[q];
    for (int s = 0; s < Q; ++s) {
      const T fe = feq<T>(s, a.rho0, rv);
      f[c * Q + s] = fe;
      F[c * Q + s] = fe;
    }
    reinit[c] = 0;
  }
}

// --------------------------------------------------------------------------
// MPM (mpm/solid.py)
// --------------------------------------------------------------------------
template <typename T>
struct MpmArgs {
  int gnx, gny, gnz;
  int n_particles;
  int bound;
  int damping_enabled, lock_z_enabled;
  T dx, inv_dx, dt, mu, la;
  T gravity[3];
  T velocity_damping;
  T domain_max[3];
};

template <typename T>
void weights(T X, int& base, T& fx, T w[3]) {
  base = (int)(X - (T)0.5);  // ti.cast(f32 -> i32) truncates
  fx = X - (T)base;
  w[0] = (T)0.5 * ((T)1.5 - fx) * ((T)1.5 - fx);
  const T c1 = ORACLE_MUTANT == 5 ? fx - (T)0.5 : fx - (T)1.0;
  w[1] = (T)0.75 - c1 * c1;
  w[2] = (T)0.5 * (fx - (T)0.5) * (fx - (T)0.5);
  if (ORACLE_MUTANT == 6) std::swap(w[0], w[2]);
}

template <typename T>
T det3(const T* A) {
  return A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) + A[2] * (A[3] * A[7] - A[4] * A[6]);
}

// R = U V^T of ti.svd(F) (solid.py:232-233).  Restated from the polar
// decomposition: for det F > 0, R is the orthogonal polar factor of F,
// computed here by Higham's scaled Newton iteration
//   X <- (g X + X^{-T} / g) / 2,  g = (|det X|)^(-1/3)
// (N. J. Higham, "Computing the polar decomposition - with applications",
// SIAM J. Sci. Stat. Comput. 7 (1986)), iterated in double to convergence.
// For det F <= 0, where ti.svd returns U, V with det = +1 and a negative
// smallest singular value, R = Q V diag(1,1,-1) V^T with Q the polar factor
// of F and V's last column the right singular vector of the smallest
// singular value; V comes from a Jacobi eigen-decomposition of F^T F.
inline void jacobi_eig(const double A0[9], double lam[3], double V[9]) {
  double A[9];
  memcpy(A, A0, sizeof(A));
  for (int e = 0; e < 9; ++e) V[e] = (e % 4 == 0) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 50; ++sweep) {
    const double off = fabs(A[1]) + fabs(A[2]) + fabs(A[5]);
    if (off < 1e-300) break;
    for (int pq = 0; pq < 3; ++pq) {
      const int p = pq == 2 ? 1 : 0, q = pq == 0 ? 1 : 2;
      const double apq = A[p * 3 + q];
      if (fabs(apq) < 1e-300) continue;
      const double th = (A[q * 3 + q] - A[p * 3 + p]) / (2 * apq);
      const double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1));
      const double cs = 1 / sqrt(t * t + 1), sn = t * cs;
      for (int r = 0; r < 3; ++r) {
        const double x = A[r * 3 + p], y = A[r * 3 + q];
        A[r * 3 + p] = cs * x - sn * y;
        A[r * 3 + q] = sn * x + cs * y;
      }
      for (int c = 0; c < 3; ++c) {
        const double x = A[p * 3 + c], y = A[q * 3 + c];
        A[p * 3 + c] = cs * x - sn * y;
        A[q * 3 + c] = sn * x + cs * y;
      }
      for (int r = 0; r < 3; ++r) {
        const double x = V[r * 3 + p], y = V[r * 3 + q];
        V[r * 3 + p] = cs * x - sn * y;
        V[r * 3 + q] = sn * x + cs * y;
      }
    }
  }
  lam[0] = A[0];
  lam[1] = A[4];
  lam[2] = A[8];
}

inline void inv_transpose(const double* X, double* Y) {
  const double d = det3<double>(X);
  Y[0] = (X[4] * X[8] - X[5] * X[7]) / d;
  Y[1] = (X[5] * X[6] - X[3] * X[8]) / d;
  Y[2] = (X[3] * X[7] - X[4] * X[6]) / d;
  Y[3] = (X[2] * X[7] - X[1] * X[8]) / d;
  Y[4] = (X[0] * X[8] - X[2] * X[6]) / d;
  Y[5] = (X[1] * X[6] - X[0] * X[7]) / d;
  Y[6] = (X[1] * X[5] - X[2] * X[4]) / d;
  Y[7] = (X[2] * X[3] - X[0] * X[5]) / d;
  Y[8] = (X[0] * X[4] - X[1] * X[3]) / d;
}

inline void polar_rotation(const double* F, double* R) {
  const double dF = det3<double>(F);
  double X[9];
  memcpy(X, F, sizeof(X));
  for (int it = 0; it < 100; ++it) {
    double Y[9];
    inv_transpose(X, Y);
    const double g = pow(fabs(det3<double>(X)), -1.0 / 3.0);
    double dmax = 0;
    for (int e = 0; e < 9; ++e) {
      const double nx = 0.5 * (g * X[e] + Y[e] / g);
      dmax = std::max(dmax, fabs(nx - X[e]));
      X[e] = nx;
    }
    if (dmax < 1e-15) break;
  }
  if (dF > 0) {
    memcpy(R, X, sizeof(X));
    return;
  }
  // det F <= 0: X is the polar factor with det -1; rotate back by the
  // reflection across the smallest right singular vector.
  double A[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) A[r * 3 + c] = F[r] * F[c] + F[3 + r] * F[3 + c] + F[6 + r] * F[6 + c];
  double lam[3], V[9];
  jacobi_eig(A, lam, V);
  int m = 0;
  for (int q = 1; q < 3; ++q)
    if (lam[q] < lam[m]) m = q;
  const double vm[3] = {V[0 * 3 + m], V[1 * 3 + m], V[2 * 3 + m]};
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      double acc = 0;
      for (int q = 0; q < 3; ++q) acc += X[r * 3 + q] * ((q == c ? 1.0 : 0.0) - 2.0 * vm[q] * vm[c]);
      R[r * 3 + c] = acc;
    }
}

// clear_grid + p2g (solid.py:213-249); particles in index order, sequential sums.
template <typename T>
void p2g(const MpmArgs<T>& a, const T* x, const T* v, const T* C, const T* F, T* Jp, const T* mass, const T* vol0,
         T* grid_v, T* grid_m, T* grid_f) {
  const long long G = (long long)a.gnx * a.gny * a.gnz;
  for (long long n = 0; n < G; ++n) {
    grid_v[3 * n] = grid_v[3 * n + 1] = grid_v[3 * n + 2] = 0;
    grid_m[n] = 0;
    grid_f[3 * n] = grid_f[3 * n + 1] = grid_f[3 * n + 2] = 0;
  }
  for (int p = 0; p < a.n_particles; ++p) {
    int base[3];
    T fx[3], w[3][3];
    for (int d = 0; d < 3; ++d) weights<T>(x[3 * p + d] * a.inv_dx, base[d], fx[d], w[d]);
    const T* Fp = F + 9 * p;
    double Fd[9], Rd[9];
    for (int e = 0; e < 9; ++e) Fd[e] = (double)Fp[e];
    polar_rotation(Fd, Rd);
    // F - R from the exact polar factor (evaluated in double, then narrowed):
    // the reference's ti.svd f32 iteration is not available, and at small
    // strain an f32-rounded R would leave F - R with only a few valid digits.
    T FmR[9];
    for (int e = 0; e < 9; ++e) FmR[e] = (T)(Fd[e] - Rd[e]);
    const T J = det3<T>(Fp);
    Jp[p] = J;
    T FiT[9];
    FiT[0] = (Fp[4] * Fp[8] - Fp[5] * Fp[7]) / J;
    FiT[1] = (Fp[5] * Fp[6] - Fp[3] * Fp[8]) / J;
    FiT[2] = (Fp[3] * Fp[7] - Fp[4] * Fp[6]) / J;
    FiT[3] = (Fp[2] * Fp[7] - Fp[1] * Fp[8]) / J;
    FiT[4] = (Fp[0] * Fp[8] - Fp[2] * Fp[6]) / J;
    FiT[5] = (Fp[1] * Fp[6] - Fp[0] * Fp[7]) / J;
    FiT[6] = (Fp[1] * Fp[5] - Fp[2] * Fp[4]) / J;
    FiT[7] = (Fp[2] * Fp[3] - Fp[0] * Fp[5]) / J;
    FiT[8] = (Fp[0] * Fp[4] - Fp[1] * Fp[3]) / J;
    T P[9];
    for (int e = 0; e < 9; ++e) P[e] = (T)2.0 * a.mu * FmR[e] + a.la * (J - (T)1.0) * J * FiT[e];
    T aff[9];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        T pf = 0;
        for (int q = 0; q < 3; ++q) pf += P[r * 3 + q] * Fp[c * 3 + q];
        aff[r * 3 + c] = -a.dt * vol0[p] * pf * a.inv_dx * a.inv_dx + mass[p] * C[9 * p + r * 3 + c];
      }
    for (int oa = 0; oa < 3; ++oa)
      for (int ob = 0; ob < 3; ++ob)
        for (int oc = 0; oc < 3; ++oc) {
          const int I = base[0] + oa, J2 = base[1] + ob, K = base[2] + oc;
          if (!(0 <= I && I < a.gnx && 0 <= J2 && J2 < a.gny && 0 <= K && K < a.gnz)) continue;
          const T dpos[3] = {((T)oa - fx[0]) * a.dx, ((T)ob - fx[1]) * a.dx, ((T)oc - fx[2]) * a.dx};
          const T wt = w[0][oa] * w[1][ob] * w[2][oc];
          const long long n = ((long long)I * a.gny + J2) * a.gnz + K;
          for (int r = 0; r < 3; ++r) {
            const T mom = mass[p] * v[3 * p + r] + (aff[r * 3] * dpos[0] + aff[r * 3 + 1] * dpos[1] + aff[r * 3 + 2] * dpos[2]);
            grid_v[3 * n + r] += wt * mom;
          }
          grid_m[n] += wt * mass[p];
        }
  }
}

// grid_update + apply_boundary_at_node (solid.py:252-259, :97-111)
template <typename T>
void grid_update(const MpmArgs<T>& a, T* grid_v, const T* grid_m, const T* grid_f) {
  for (int I = 0; I < a.gnx; ++I)
    for (int J = 0; J < a.gny; ++J)
      for (int K = 0; K < a.gnz; ++K) {
        const long long n = ((long long)I * a.gny + J) * a.gnz + K;
        if (!(grid_m[n] > 0)) continue;
        const int idx[3] = {I, J, K};
        const int ext[3] = {a.gnx, a.gny, a.gnz};
        for (int d = 0; d < 3; ++d) {
          T vd = grid_v[3 * n + d] / grid_m[n];
          vd += a.dt * a.gravity[d];
          vd += a.dt * grid_f[3 * n + d] / grid_m[n];
          grid_v[3 * n + d] = vd;
        }
        for (int d = 0; d < 3; ++d) {
          if (idx[d] < a.bound && grid_v[3 * n + d] < 0) grid_v[3 * n + d] = 0;
          if (idx[d] > ext[d] - a.bound && grid_v[3 * n + d] > 0) grid_v[3 * n + d] = 0;
        }
      }
}

// g2p + clamp_particle_to_domain + apply_step112_dynamics_controls + pin (solid.py:262-315, :114-124)
template <typename T>
void g2p(const MpmArgs<T>& a, T* x, T* v, T* C, T* F, T* Jp, const int8_t* fixed_mask, const T* fixed_x,
         const T* grid_v) {
  for (int p = 0; p < a.n_particles; ++p) {
    int base[3];
    T fx[3], w[3][3];
    for (int d = 0; d < 3; ++d) weights<T>(x[3 * p + d] * a.inv_dx, base[d], fx[d], w[d]);
    T nv[3] = {0, 0, 0}, nC[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int oa = 0; oa < 3; ++oa)
      for (int ob = 0; ob < 3; ++ob)
        for (int oc = 0; oc < 3; ++oc) {
          const int I = base[0] + oa, J2 = base[1] + ob, K = base[2] + oc;
          if (!(0 <= I && I < a.gnx && 0 <= J2 && J2 < a.gny && 0 <= K && K < a.gnz)) continue;
          const T dpos[3] = {((T)oa - fx[0]) * a.dx, ((T)ob - fx[1]) * a.dx, ((T)oc - fx[2]) * a.dx};
          const T wt = w[0][oa] * w[1][ob] * w[2][oc];
          const long long n = ((long long)I * a.gny + J2) * a.gnz + K;
          for (int r = 0; r < 3; ++r) {
            nv[r] += wt * grid_v[3 * n + r];
            for (int c = 0; c < 3; ++c) nC[r * 3 + c] += (T)4.0 * wt * grid_v[3 * n + r] * dpos[c] * a.inv_dx * a.inv_dx;
          }
        }
    for (int d = 0; d < 3; ++d) {
      v[3 * p + d] = nv[d];
      x[3 * p + d] += a.dt * nv[d];
    }
    const T eps = a.dx;
    for (int d = 0; d < 3; ++d) {
      if (x[3 * p + d] < eps) {
        x[3 * p + d] = eps;
        if (v[3 * p + d] < 0) v[3 * p + d] = 0;
      }
      if (x[3 * p + d] > a.domain_max[d] - eps) {
        x[3 * p + d] = a.domain_max[d] - eps;
        if (v[3 * p + d] > 0) v[3 * p + d] = 0;
      }
    }
    for (int e = 0; e < 9; ++e) C[9 * p + e] = nC[e];
    T Fn[9];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        T acc = 0;
        for (int q = 0; q < 3; ++q) acc += ((r == q ? (T)1 : (T)0) + a.dt * nC[r * 3 + q]) * F[9 * p + q * 3 + c];
        Fn[r * 3 + c] = acc;
      }
    for (int e = 0; e < 9; ++e) F[9 * p + e] = Fn[e];
    Jp[p] = det3<T>(F + 9 * p);
    if (a.damping_enabled) {
      T df = (T)1.0 - a.velocity_damping * a.dt;
      df = df > 0 ? df : 0;
      for (int d = 0; d < 3; ++d) v[3 * p + d] *= df;
    }
    if (a.lock_z_enabled) {
      x[3 * p + 2] = fixed_x[3 * p + 2];
      v[3 * p + 2] = 0;
      for (int d = 0; d < 3; ++d) {
        C[9 * p + 2 * 3 + d] = 0;
        C[9 * p + d * 3 + 2] = 0;
        F[9 * p + 2 * 3 + d] = 0;
        F[9 * p + d * 3 + 2] = 0;
      }
      F[9 * p + 8] = 1;
      Jp[p] = det3<T>(F + 9 * p);
    }
    if (fixed_mask[p] != 0) {
      for (int d = 0; d < 3; ++d) {
        x[3 * p + d] = fixed_x[3 * p + d];
        v[3 * p + d] = 0;
      }
      for (int e = 0; e < 9; ++e) {
        C[9 * p + e] = 0;
        F[9 * p + e] = (e % 4 == 0) ? (T)1 : (T)0;
      }
      Jp[p] = 1;
    }
  }
}

// --------------------------------------------------------------------------
// Coupling (coupling/projection.py, penalty.py, moving_boundary.py)
// --------------------------------------------------------------------------
template <typename T>
struct CplArgs {
  T inv_dx_norm, cell_volume_norm, vel_scale, eps_mass;
  T beta_lbm, penalty_phi_min, force_cap_lbm;
  T force_density_scale, reaction_scale;
  T mb_phi_min, mb_force_cap_norm;
  T area_scale;  // link-area transfer's global factor (1 for the engineering transfer)
};

template <typename T>
struct ProjStats {
  T projected_mass, projected_volume_raw, projected_volume_clamped, max_phi_raw;
  long long active_cell_count;
};

// clear_projection + project_particles + normalize_projection (projection.py:20-97)
template <typename T>
void project(const LbmArgs<T>& la, const MpmArgs<T>& ma, const CplArgs<T>& ca, const T* x, const T* v,
             const T* Jp, const T* mass, const T* vol0, T* phi, T* smass, T* svel, ProjStats<T>* st) {
  const Dims d{la.nx, la.ny, la.nz};
  for (long long c = 0; c < d.n(); ++c) {
    phi[c] = 0;
    smass[c] = 0;
    svel[3 * c] = svel[3 * c + 1] = svel[3 * c + 2] = 0;
  }
  for (int p = 0; p < ma.n_particles; ++p) {
    int base[3];
    T fx[3], w[3][3];
    for (int q = 0; q < 3; ++q) weights<T>(x[3 * p + q] * ca.inv_dx_norm, base[q], fx[q], w[q]);
    const T J = std::max(Jp[p], (T)0);
    const T vol = vol0[p] * J;
    T vl[3];
    for (int q = 0; q < 3; ++q) vl[q] = v[3 * p + q] * ca.vel_scale;
    for (int oa = 0; oa < 3; ++oa)
      for (int ob = 0; ob < 3; ++ob)
        for (int oc = 0; oc < 3; ++oc) {
          const int I = base[0] + oa, J2 = base[1] + ob, K = base[2] + oc;
          if (!(0 <= I && I < d.nx && 0 <= J2 && J2 < d.ny && 0 <= K && K < d.nz)) continue;
          const T wt = w[0][oa] * w[1][ob] * w[2][oc];
          const T mc = wt * mass[p];
          const T vf = wt * vol / ca.cell_volume_norm;
          const long long c = d.idx(I, J2, K);
          phi[c] += vf;
          smass[c] += mc;
          for (int q = 0; q < 3; ++q) svel[3 * c + q] += mc * vl[q];
        }
  }
  memset(st, 0, sizeof(*st));
  for (long long c = 0; c < d.n(); ++c) {
    const T pr = phi[c];
    st->max_phi_raw = std::max(st->max_phi_raw, pr);
    if (smass[c] > ca.eps_mass)
      for (int q = 0; q < 3; ++q) svel[3 * c + q] = svel[3 * c + q] / smass[c];
    else
      for (int q = 0; q < 3; ++q) svel[3 * c + q] = 0;
    const T pc = std::min((T)1, std::max((T)0, pr));
    phi[c] = pc;
    if (pc > (T)1e-8) st->active_cell_count += 1;
    st->projected_mass += smass[c];
    st->projected_volume_raw += pr * ca.cell_volume_norm;
    st->projected_volume_clamped += pc * ca.cell_volume_norm;
  }
}

// clear_force_fields + build_penalty_force (penalty.py:42-88)
template <typename T>
void penalty_force(const LbmArgs<T>& la, const CplArgs<T>& ca, const int8_t* solid, const T* phi, const T* rho,
                   const T* v, const T* svel, T* cell_force, T* hydro_force, T* out_stats /*active,maxc,maxh,net6*/) {
  const Dims d{la.nx, la.ny, la.nz};
  for (int q = 0; q < 9; ++q) out_stats[q] = 0;
  for (long long c = 0; c < d.n(); ++c) {
    T f[3] = {0, 0, 0};
    if (solid[c] == 0 && phi[c] > ca.penalty_phi_min) {
      for (int q = 0; q < 3; ++q) f[q] = ca.beta_lbm * phi[c] * rho[c] * (svel[3 * c + q] - v[3 * c + q]);
      const T n = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
      if (n > ca.force_cap_lbm)
        for (int q = 0; q < 3; ++q) f[q] = f[q] * (ca.force_cap_lbm / n);
      out_stats[0] += 1;
    }
    for (int q = 0; q < 3; ++q) {
      cell_force[3 * c + q] = f[q];
      hydro_force[3 * c + q] = -f[q];
      out_stats[3 + q] += f[q];
      out_stats[6 + q] += -f[q];
    }
    const T n = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    out_stats[1] = std::max(out_stats[1], n);
    out_stats[2] = std::max(out_stats[2], n);
  }
}

// add_lbm_reaction_to_mpm_grid (penalty.py:90-129), mode 1, and
// add_moving_boundary_reaction_to_mpm_grid (moving_boundary.py:58-111), mode 2.
template <typename T>
void reaction(const LbmArgs<T>& la, const MpmArgs<T>& ma, const CplArgs<T>& ca, int mode, const T* x, const T* Jp,
              const T* vol0, const T* hydro, const T* phi, T* grid_f, T* out_stats /*active,maxp,maxg,netp3,netg3*/) {
  const Dims d{la.nx, la.ny, la.nz};
  for (int q = 0; q < 9; ++q) out_stats[q] = 0;
  for (int p = 0; p < ma.n_particles; ++p) {
    int base[3];
    T fx[3], w[3][3];
    for (int q = 0; q < 3; ++q) weights<T>(x[3 * p + q] * ca.inv_dx_norm, base[q], fx[q], w[q]);
    T h[3] = {0, 0, 0}, sphi = 0;
    for (int oa = 0; oa < 3; ++oa)
      for (int ob = 0; ob < 3; ++ob)
        for (int oc = 0; oc < 3; ++oc) {
          const int I = base[0] + oa, J2 = base[1] + ob, K = base[2] + oc;
          if (!(0 <= I && I < d.nx && 0 <= J2 && J2 < d.ny && 0 <= K && K < d.nz)) continue;
          const T wt = w[0][oa] * w[1][ob] * w[2][oc];
          const long long c = d.idx(I, J2, K);
          for (int q = 0; q < 3; ++q) h[q] += wt * hydro[3 * c + q];
          sphi += wt * phi[c];
        }
    if (mode == 2 && !(sphi > ca.mb_phi_min)) continue;
    const T J = std::max(Jp[p], (T)0);
    const T vol = vol0[p] * J;
    T pf[3];
    for (int q = 0; q < 3; ++q) pf[q] = h[q] * ca.force_density_scale * vol * ca.reaction_scale;
    if (mode == 2) {
      // add_link_area_reaction_to_mpm_grid (link_area.py:150-157) multiplies
      // the engineering force by its area scale last
      for (int q = 0; q < 3; ++q) pf[q] = pf[q] * ca.area_scale;
      T n = sqrt(pf[0] * pf[0] + pf[1] * pf[1] + pf[2] * pf[2]);
      if (n > ca.mb_force_cap_norm) {
        for (int q = 0; q < 3; ++q) pf[q] = pf[q] * (ca.mb_force_cap_norm / n);
        n = ca.mb_force_cap_norm;
      }
      if (n > 0) out_stats[0] += 1;
      out_stats[1] = std::max(out_stats[1], n);
      for (int q = 0; q < 3; ++q) out_stats[3 + q] += pf[q];
    }
    for (int oa = 0; oa < 3; ++oa)
      for (int ob = 0; ob < 3; ++ob)
        for (int oc = 0; oc < 3; ++oc) {
          const int I = base[0] + oa, J2 = base[1] + ob, K = base[2] + oc;
          if (!(0 <= I && I < ma.gnx && 0 <= J2 && J2 < ma.gny && 0 <= K && K < ma.gnz)) continue;
          const T wt = w[0][oa] * w[1][ob] * w[2][oc];
          const long long n = ((long long)I * ma.gny + J2) * ma.gnz + K;
          T cn = 0;
          for (int q = 0; q < 3; ++q) {
            const T cq = wt * pf[q];
            grid_f[3 * n + q] += cq;
            cn += cq * cq;
            out_stats[6 + q] += cq;
          }
          out_stats[2] = std::max(out_stats[2], (T)sqrt(cn));
        }
  }
}

}  // namespace

// --------------------------------------------------------------------------
// C ABI (ctypes): one entry per reference kernel, suffix _f32 / _f64.
// --------------------------------------------------------------------------
#define ORC_EXPORT extern "C" __attribute__((visibility("default")))

#define ORC_DEFINE(T, SFX)                                                                                          \
  ORC_EXPORT void orc_lbm_init_##SFX(const LbmArgs<T>* a, T* F, T* f, T* rho, T* v) { lbm_init<T>(*a, F, f, rho, v); } \
  ORC_EXPORT void orc_colission_##SFX(const LbmArgs<T>* a, const T* F, T* f, const T* rho, const T* v,             \
                                      const int8_t* solid, const T* cell_force) {                                   \
    colission<T>(*a, F, f, rho, v, solid, cell_force);                                                            \
  }                                                                                                                 \
  ORC_EXPORT void orc_colission_regularized_##SFX(const LbmArgs<T>* a, const T* F, T* f, const T* rho,       \
                                                  const T* v, const int8_t* solid, const T* cell_force) {         \
    colission_regularized<T>(*a, F, f, rho, v, solid, cell_force);                                                \
  }                                                                                                                 \
  ORC_EXPORT void orc_streaming1_##SFX(const LbmArgs<T>* a, const T* f, T* F, const int8_t* solid) {              \
    streaming1<T>(*a, f, F, solid);                                                                               \
  }                                                                                                                 \
  ORC_EXPORT void orc_streaming_mb_##SFX(const LbmArgs<T>* a, const T* f, T* F, const int8_t* solid, const T* rho, \
                                         const T* svel, T* hydro, MbStats<T>* st) {                                 \
    streaming_mb<T>(*a, f, F, solid, rho, svel, hydro, st);                                                       \
  }                                                                                                                 \
  ORC_EXPORT void orc_bc_legacy_##SFX(const LbmArgs<T>* a, T* F, const T* v, const int8_t* solid) {               \
    bc_legacy<T>(*a, F, v, solid);                                                                                \
  }                                                                                                                 \
  ORC_EXPORT void orc_bc_open_##SFX(const LbmArgs<T>* a, const ObArgs<T>* ob, T* F, const T* rho, const T* v,     \
                                    const int8_t* solid, ObCounts* k) {                                             \
    bc_open<T>(*a, *ob, F, rho, v, solid, k);                                                                      \
  }                                                                                                                 \
  ORC_EXPORT void orc_streaming3_##SFX(const LbmArgs<T>* a, const T* F, T* f, T* rho, T* v, const int8_t* solid,   \
                                       const T* cell_force) {                                                       \
    streaming3<T>(*a, F, f, rho, v, solid, cell_force);                                                           \
  }                                                                                                                 \
  ORC_EXPORT void orc_dynamic_solid_##SFX(const LbmArgs<T>* a, T thr, const T* phi, const int8_t* static_solid,    \
                                          int8_t* solid, int8_t* old_solid, int8_t* reinit, const T* svel, T* F,    \
                                          T* f, T* rho, T* v) {                                                     \
    dynamic_solid<T>(*a, thr, phi, static_solid, solid, old_solid, reinit, svel, F, f, rho, v);                  \
  }                                                                                                                 \
  ORC_EXPORT void orc_p2g_##SFX(const MpmArgs<T>* a, const T* x, const T* v, const T* C, const T* F, T* Jp,        \
                                const T* mass, const T* vol0, T* gv, T* gm, T* gf) {                                \
    p2g<T>(*a, x, v, C, F, Jp, mass, vol0, gv, gm, gf);                                                           \
  }                                                                                                                 \
  ORC_EXPORT void orc_grid_update_##SFX(const MpmArgs<T>* a, T* gv, const T* gm, const T* gf) {                   \
    grid_update<T>(*a, gv, gm, gf);                                                                               \
  }                                                                                                                 \
  ORC_EXPORT void orc_g2p_##SFX(const MpmArgs<T>* a, T* x, T* v, T* C, T* F, T* Jp, const int8_t* fm,             \
                                const T* fxp, const T* gv) {                                                        \
    g2p<T>(*a, x, v, C, F, Jp, fm, fxp, gv);                                                                      \
  }                                                                                                                 \
  ORC_EXPORT void orc_project_##SFX(const LbmArgs<T>* la, const MpmArgs<T>* ma, const CplArgs<T>* ca, const T* x,  \
                                    const T* v, const T* Jp, const T* mass, const T* vol0, T* phi, T* smass,        \
                                    T* svel, ProjStats<T>* st) {                                                    \
    project<T>(*la, *ma, *ca, x, v, Jp, mass, vol0, phi, smass, svel, st);                                         \
  }                                                                                                                 \
  ORC_EXPORT void orc_penalty_force_##SFX(const LbmArgs<T>* la, const CplArgs<T>* ca, const int8_t* solid,         \
                                          const T* phi, const T* rho, const T* v, const T* svel, T* cf, T* hf,      \
                                          T* out) {                                                                 \
    penalty_force<T>(*la, *ca, solid, phi, rho, v, svel, cf, hf, out);                                            \
  }                                                                                                                 \
  ORC_EXPORT void orc_reaction_##SFX(const LbmArgs<T>* la, const MpmArgs<T>* ma, const CplArgs<T>* ca, int mode,  \
                                     const T* x, const T* Jp, const T* vol0, const T* hydro, const T* phi, T* gf,   \
                                     T* out) {                                                                      \
    reaction<T>(*la, *ma, *ca, mode, x, Jp, vol0, hydro, phi, gf, out);                                           \
  }                                                                                                                 \
  ORC_EXPORT void orc_polar_##SFX(const T* F, T* R) {                                                              \
    double Fd[9], Rd[9];                                                                                          \
    for (int e = 0; e < 9; ++e) Fd[e] = (double)F[e];                                                             \
    polar_rotation(Fd, Rd);                                                                                       \
    for (int e = 0; e < 9; ++e) R[e] = (T)Rd[e];                                                                  \
  }                                                                                                                 \
  ORC_EXPORT void orc_bspline_##SFX(int n, const T* X, int* base, T* fx, T* w) {                                   \
    for (int i = 0; i < n; ++i) weights<T>(X[i], base[i], fx[i], w + 3 * i);                                        \
  }                                                                                                                 \
  ORC_EXPORT void orc_inv_M_##SFX(T* out) {                                                                        \
    const Tables<T>& tb = tables<T>();                                                                            \
    for (int r = 0; r < Q; ++r)                                                                                   \
      for (int c = 0; c < Q; ++c) out[r * Q + c] = tb.iM[r][c];                                                   \
  }

ORC_DEFINE(float, f32)
ORC_DEFINE(double, f64)

ORC_EXPORT int orc_mbstats_size_f32() { return (int)sizeof(MbStats<float>); }
ORC_EXPORT int orc_mbstats_size_f64() { return (int)sizeof(MbStats<double>); }
ORC_EXPORT int orc_mutant() { return ORACLE_MUTANT; }

// The bounce-back links of every fluid cell, as the streaming passes classify
// them: bit s (1..18) of mask[c] is set when the streaming target of fluid
// cell c in direction s is solid; solid cells get 0.
ORC_EXPORT void orc_link_masks(int nx, int ny, int nz, const int8_t* solid, uint32_t* mask) {
  const Dims d{nx, ny, nz};
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) {
        const long long c = d.idx(i, j, k);
        uint32_t m = 0;
        if (solid[c] == 0)
          for (int s = 1; s < Q; ++s)
            if (solid[stream_target(d, i, j, k, s)] != 0) m |= 1u << s;
        mask[c] = m;
      }
}
