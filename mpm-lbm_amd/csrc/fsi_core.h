// fsi_core.h - per-cell / per-particle bodies of the FSI substep kernels.
//
// Every function here is FSI_HD: the HIP build (fsi_kernels.hip) calls them
// from __global__ kernels on gfx950, and the CPU emulation build
// (fsi_emu.cpp, test-only) calls the very same bodies from serial loops so
// the kernel arithmetic and indexing can be checked on a machine without a
// GPU.  Atomic updates go through fsi_atomic_* which map to hardware atomics
// on the GPU and plain read-modify-write in the serial emulation.
//
// Reference semantics followed (lizhuoh9/MPM-LBM, src/mpm_lbm/sim/...):
//   D3Q19 ordering, weights, LR table ........ lbm/fluid.py:252, :359-373
//   feq ....................................... lbm/fluid.py:321-326
//   MRT moments M, meq, Guo forcing ........... lbm/fluid.py:228-277, :386-413
//   streaming + half-way / moving-wall BB ..... lbm/fluid.py:432-440, :489-519
//   legacy x-face boundary conditions ......... lbm/fluid.py:927-958
//   streaming3 moments ........................ lbm/fluid.py:1795-1814
//   dynamic solid / reinit .................... lbm/fluid.py:1848-1877
//   MPM p2g / grid_update / g2p ............... mpm/solid.py:213-315
//   projection / penalty / MB reaction ........ coupling/{projection,penalty,moving_boundary}.py
#pragma once

#include <stdint.h>
#include <math.h>
#include <string.h>
#include "../../include/fsi_hip.h"

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#define FSI_HD __host__ __device__ __forceinline__
#define FSI_HDM __host__ __device__ __forceinline__
// Evaluate each float operation rounded on its own, in the written order, as
// the oracle's plain C++ does: the library is built with
// -ffp-contract=fast-honor-pragmas, so these turn a*b+c -> fma fusion off
// (inside a function body, or from BEGIN to END at file scope) and back on.
// The build without fma fusion (FSI_EXACT_BUILD, -ffp-contract=off) keeps it
// off after END as well, so no kernel of that build fuses anything.
#define FSI_NO_CONTRACT _Pragma("clang fp contract(off)")
#define FSI_NO_CONTRACT_BEGIN _Pragma("clang fp contract(off)")
#ifdef FSI_EXACT_BUILD
#define FSI_NO_CONTRACT_END _Pragma("clang fp contract(off)")
#else
#define FSI_NO_CONTRACT_END _Pragma("clang fp contract(fast)")
#endif
#else
#define FSI_HD static inline
#define FSI_HDM inline
#define FSI_NO_CONTRACT
#define FSI_NO_CONTRACT_BEGIN
#define FSI_NO_CONTRACT_END
#endif

namespace fsi {

constexpr int Q = 19;
constexpr int NLINK = 18;  // link directions 1..18 (direction 0 never leaves the cell)

// cellinfo bits
constexpr uint32_t CI_SOLID = 1u << 31;   // cell is solid
constexpr uint32_t CI_MOVING = 1u << 30;  // fluid cell: has a link to a solid cell of nonzero velocity;
                                          // solid cell: its own solid velocity is nonzero
constexpr uint32_t CI_MASK = 0x7FFFEu;    // bits 1..18: link (fluid) / incoming (solid) directions

// D3Q19 velocity set in the reference's order (fluid.py:359-369).
constexpr int EX[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
constexpr int EY[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
constexpr int EZ[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
constexpr int LR[Q] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};

FSI_HD float wgt(int s) {
  return s == 0 ? (1.0f / 3.0f) : (s <= 6 ? (1.0f / 18.0f) : (1.0f / 36.0f));
}

// feq(k, rho, u) = w_k rho (1 + 3 e.u + 4.5 (e.u)^2 - 1.5 u.u)   (fluid.py:321-326)
FSI_HD float feq(int s, float rho, float ux, float uy, float uz) {
  const float eu = (float)EX[s] * ux + (float)EY[s] * uy + (float)EZ[s] * uz;
  const float uv = ux * ux + uy * uy + uz * uz;
  return wgt(s) * rho * (1.0f + 3.0f * eu + 4.5f * eu * eu - 1.5f * uv);
}

// ---------------------------------------------------------------------------
// Population storage.  FP32 stores F_s; the 16-bit forms store the deviation
// d_s = F_s - w_s from the rest-state weights, rounded to nearest even, and
// every load / collision runs in FP32 (F_s = w_s + d_s).
// ---------------------------------------------------------------------------
FSI_HD uint32_t f32_bits(float v) {
  uint32_t u;
  memcpy(&u, &v, 4);
  return u;
}
FSI_HD float bits_f32(uint32_t u) {
  float v;
  memcpy(&v, &u, 4);
  return v;
}
FSI_HD uint16_t f32_to_bf16(float v) {
  const uint32_t u = f32_bits(v);
  if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu)) return (uint16_t)((u >> 16) | 0x40u);  // quiet NaN
  return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
FSI_HD float bf16_to_f32(uint16_t h) { return bits_f32((uint32_t)h << 16); }

#ifdef __HIPCC__
FSI_HD uint16_t f32_to_f16(float v) { return __half_as_ushort(__float2half_rn(v)); }
FSI_HD float f16_to_f32(uint16_t h) { return __half2float(__ushort_as_half(h)); }
#else
// IEEE binary16 with round-to-nearest-even (what v_cvt_f16_f32 does on gfx950).
static inline uint16_t f32_to_f16(float v) {
  const uint32_t u = f32_bits(v);
  const uint32_t sign = (u >> 16) & 0x8000u;
  const uint32_t ax = u & 0x7FFFFFFFu;
  if (ax >= 0x7F800000u) return (uint16_t)(sign | 0x7C00u | (ax > 0x7F800000u ? 0x200u : 0u));
  if (ax >= 0x477FF000u) return (uint16_t)(sign | 0x7C00u);  // rounds to infinity
  if (ax < 0x38800000u) {                                   // subnormal or zero in binary16
    if (ax < 0x33000000u) return (uint16_t)sign;            // below half the smallest subnormal
    const uint32_t mant = (ax & 0x007FFFFFu) | 0x00800000u;
    const int shift = 126 - (int)(ax >> 23);                // 14..24
    const uint32_t q = mant >> shift, rem = mant & ((1u << shift) - 1u), half = 1u << (shift - 1);
    const uint32_t r = q + ((rem > half || (rem == half && (q & 1u))) ? 1u : 0u);
    return (uint16_t)(sign | r);
  }
  const uint32_t e = (ax >> 23) - 112u;
  const uint32_t m = ax & 0x007FFFFFu;
  uint32_t h = (e << 10) | (m >> 13);
  const uint32_t rem = m & 0x1FFFu;
  if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) ++h;
  return (uint16_t)(sign | h);
}
static inline float f16_to_f32(uint16_t h) {
  const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
  const uint32_t e = (h >> 10) & 0x1Fu, m = h & 0x3FFu;
  if (e == 0) {
    const float v = (float)m * 5.9604644775390625e-8f;  // m * 2^-24
    return sign ? -v : v;
  }
  if (e == 31) return bits_f32(sign | 0x7F800000u | (m << 13));
  return bits_f32(sign | ((e + 112u) << 23) | (m << 13));
}
#endif

// A value that every lane of the wave holds (the lattice size), shown to be
// uniform.  A uniform value that reaches a use through divergent branches can
// end up in vector registers, and the plane bases derived from it then take
// 64-bit per-lane address arithmetic at every load and store.
FSI_HD uint32_t wave_uniform(uint32_t v) {
#if defined(__HIP_DEVICE_COMPILE__)
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
#else
  return v;
#endif
}

// An instruction-scheduling fence (device): the code before it is scheduled
// before the code after it, which bounds how many loads are in flight (and
// registers live) at once.
FSI_HD void sched_fence() {
#if defined(__HIP_DEVICE_COMPILE__)
  __builtin_amdgcn_sched_barrier(0);
#endif
}

// Element c of a population plane through a 32-bit byte offset.  A lattice
// holds at most 2^30 cells (checked by the C ABI), so c * sizeof(T) < 2^32:
// the plane base (uniform: buffer + s * N) stays in scalar registers and each
// lane adds one zero-extended 32-bit offset, which gfx950 global loads and
// stores take directly (saddr + voffset form).
//
// Every plane lives in device global memory.  On the GPU the access is made
// through a global-address-space pointer, so a plane base the kernel read
// from memory (the batched kernels' instance table) still gets global
// loads and stores instead of flat ones.
#if defined(__HIP_DEVICE_COMPILE__)
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T& at32(T* plane, uint32_t c) {
  typedef __attribute__((address_space(1))) T GT;
  typedef __attribute__((address_space(1))) const char GC;
  return *(GT*)((GC*)plane + (size_t)(uint32_t)(c * (uint32_t)sizeof(T)));
}
#else
template <typename T>
inline T& at32(T* plane, uint32_t c) {
  return *(T*)((const char*)plane + (size_t)(uint32_t)(c * (uint32_t)sizeof(T)));
}
#endif

struct PopF32 {
  static constexpr bool kF32 = true;
  float* p;
  FSI_HDM float get(long long N, int s, long long c) const { return p[s * N + c]; }
  FSI_HDM void put(long long N, int s, long long c, float v) const { p[s * N + c] = v; }
  FSI_HDM float ld(uint32_t N, int s, uint32_t c) const { return at32(p + (size_t)s * N, c); }
  FSI_HDM void st(uint32_t N, int s, uint32_t c, float v) const { at32(p + (size_t)s * N, c) = v; }
};
struct PopF16 {
  static constexpr bool kF32 = false;
  uint16_t* p;
  FSI_HDM float get(long long N, int s, long long c) const { return wgt(s) + f16_to_f32(p[s * N + c]); }
  FSI_HDM void put(long long N, int s, long long c, float v) const { p[s * N + c] = f32_to_f16(v - wgt(s)); }
  FSI_HDM float ld(uint32_t N, int s, uint32_t c) const { return wgt(s) + f16_to_f32(at32(p + (size_t)s * N, c)); }
  FSI_HDM void st(uint32_t N, int s, uint32_t c, float v) const { at32(p + (size_t)s * N, c) = f32_to_f16(v - wgt(s)); }
};
struct PopBF16 {
  static constexpr bool kF32 = false;
  uint16_t* p;
  FSI_HDM float get(long long N, int s, long long c) const { return wgt(s) + bf16_to_f32(p[s * N + c]); }
  FSI_HDM void put(long long N, int s, long long c, float v) const { p[s * N + c] = f32_to_bf16(v - wgt(s)); }
  FSI_HDM float ld(uint32_t N, int s, uint32_t c) const { return wgt(s) + bf16_to_f32(at32(p + (size_t)s * N, c)); }
  FSI_HDM void st(uint32_t N, int s, uint32_t c, float v) const { at32(p + (size_t)s * N, c) = f32_to_bf16(v - wgt(s)); }
};

// Plane s of fp32 populations (other storages have no float planes).
template <typename Pop>
FSI_HD const float* pop_plane_f32(const Pop& P, uint32_t N, int s) {
  return reinterpret_cast<const float*>(P.p) + (size_t)s * N;
}

template <typename Pop>
FSI_HD Pop pop_of(void* base) {
  Pop q;
  q.p = reinterpret_cast<decltype(q.p)>(base);
  return q;
}

// ---------------------------------------------------------------------------
// Moment transforms.  M (fluid.py:228-247) has entries in {0,+-1,+-2}; its
// inverse has the same sparsity as M^T with entries k/72 (for k in Z), so both
// products are written out over the 9 opposite-direction pairs
// (a, LR[a]) as sums P_a = F_a + F_LR(a) and differences Q_a = F_a - F_LR(a).
// Row by row these equal the dense products of the reference exactly in
// rational arithmetic; in fp32 they differ only by rounding.
// ---------------------------------------------------------------------------
FSI_HD void moments_M(const float* F, float* m) {
  const float P1 = F[1] + F[2], Q1 = F[1] - F[2];
  const float P3 = F[3] + F[4], Q3 = F[3] - F[4];
  const float P5 = F[5] + F[6], Q5 = F[5] - F[6];
  const float P7 = F[7] + F[8], Q7 = F[7] - F[8];
  const float P9 = F[9] + F[10], Q9 = F[9] - F[10];
  const float P11 = F[11] + F[12], Q11 = F[11] - F[12];
  const float P13 = F[13] + F[14], Q13 = F[13] - F[14];
  const float P15 = F[15] + F[16], Q15 = F[15] - F[16];
  const float P17 = F[17] + F[18], Q17 = F[17] - F[18];
  const float axis = P1 + P3 + P5;
  const float diag = P7 + P9 + P11 + P13 + P15 + P17;
  const float dxy = P7 + P9, dxz = P11 + P13, dyz = P15 + P17;
  m[0] = F[0] + axis + diag;
  m[1] = -F[0] + diag;
  m[2] = F[0] - 2.0f * axis + diag;
  const float jx = Q7 + Q9 + Q11 + Q13;
  const float jy = Q7 - Q9 + Q15 + Q17;
  const float jz = Q11 - Q13 + Q15 - Q17;
  m[3] = Q1 + jx;
  m[4] = -2.0f * Q1 + jx;
  m[5] = Q3 + jy;
  m[6] = -2.0f * Q3 + jy;
  m[7] = Q5 + jz;
  m[8] = -2.0f * Q5 + jz;
  m[9] = 2.0f * P1 - P3 - P5 + dxy + dxz - 2.0f * dyz;
  m[10] = -2.0f * P1 + P3 + P5 + dxy + dxz - 2.0f * dyz;
  m[11] = P3 - P5 + dxy - dxz;
  m[12] = -P3 + P5 + dxy - dxz;
  m[13] = P7 - P9;
  m[14] = P15 - P17;
  m[15] = P11 - P13;
  m[16] = Q7 + Q9 - Q11 - Q13;
  m[17] = -Q7 + Q9 + Q15 + Q17;
  m[18] = Q11 - Q13 - Q15 + Q17;
}

// f += M^{-1} r   (M^{-1} written out exactly; see moments_M)
FSI_HD void add_invM(const float* r, float* f) {
  const float a18 = (r[0] - r[2]) * (1.0f / 18.0f);
  const float c0 = r[0] * (1.0f / 36.0f) + r[1] * (1.0f / 24.0f) + r[2] * (1.0f / 72.0f);
  const float t910a = (r[9] - r[10]) * (1.0f / 12.0f);
  const float t910b = (r[9] - r[10]) * (1.0f / 24.0f);
  const float s910 = (r[9] + r[10]) * (1.0f / 48.0f);
  const float s910b = (r[9] + r[10]) * (1.0f / 24.0f);
  const float d1112 = (r[11] - r[12]) * (1.0f / 8.0f);
  const float s1112 = (r[11] + r[12]) * (1.0f / 16.0f);
  const float jx = r[3] * (1.0f / 12.0f) + r[4] * (1.0f / 24.0f);
  const float jy = r[5] * (1.0f / 12.0f) + r[6] * (1.0f / 24.0f);
  const float jz = r[7] * (1.0f / 12.0f) + r[8] * (1.0f / 24.0f);
  const float r16 = r[16] * 0.125f, r17 = r[17] * 0.125f, r18 = r[18] * 0.125f;

  f[0] += r[0] * (1.0f / 3.0f) - r[1] * 0.5f + r[2] * (1.0f / 6.0f);
  {  // (1,2)
    const float S = a18 + t910a, D = (r[3] - r[4]) * (1.0f / 6.0f);
    f[1] += S + D;
    f[2] += S - D;
  }
  {  // (3,4)
    const float S = a18 - t910b + d1112, D = (r[5] - r[6]) * (1.0f / 6.0f);
    f[3] += S + D;
    f[4] += S - D;
  }
  {  // (5,6)
    const float S = a18 - t910b - d1112, D = (r[7] - r[8]) * (1.0f / 6.0f);
    f[5] += S + D;
    f[6] += S - D;
  }
  const float cxy = c0 + s910 + s1112;
  const float cxz = c0 + s910 - s1112;
  const float cyz = c0 - s910b;
  {  // (7,8)
    const float S = cxy + r[13] * 0.25f, D = jx + jy + r16 - r17;
    f[7] += S + D;
    f[8] += S - D;
  }
  {  // (9,10)
    const float S = cxy - r[13] * 0.25f, D = jx - jy + r16 + r17;
    f[9] += S + D;
    f[10] += S - D;
  }
  {  // (11,12)
    const float S = cxz + r[15] * 0.25f, D = jx + jz - r16 + r18;
    f[11] += S + D;
    f[12] += S - D;
  }
  {  // (13,14)
    const float S = cxz - r[15] * 0.25f, D = jx - jz - r16 - r18;
    f[13] += S + D;
    f[14] += S - D;
  }
  {  // (15,16)
    const float S = cyz + r[14] * 0.25f, D = jy + jz + r17 - r18;
    f[15] += S + D;
    f[16] += S - D;
  }
  {  // (17,18)
    const float S = cyz - r[14] * 0.25f, D = jy - jz + r17 + r18;
    f[17] += S + D;
    f[18] += S - D;
  }
}

// MRT collision with Guo forcing (fluid.py:399-413):
//   m = M F;  m -= S (m - meq(rho,u));  m += (1 - S/2) M g;  f = M^{-1} m
// evaluated as f = F + M^{-1} [ -S (M F - meq) + (1 - S/2) M g ], which is the
// same map because M^{-1} M = I.  Relaxation vector (fluid.py:304):
//   [0, sv, sv, 0, so, 0, so, 0, so, sv x7, so x3].
// meq (fluid.py:386-392): m0 = rho, m1 = u.u, m3/5/7 = u, m9 = 2ux^2-uy^2-uz^2,
//   m11 = uy^2-uz^2, m13 = ux uy, m14 = uy uz, m15 = ux uz, others 0.
// Guo term as written in the reference (fluid.py:408):
//   g_l = w_l ((e_l - u).F / 3 + (e_l.u)(e_l.F) / 9).
FSI_HD void mrt_collide_m(const float* F, const float* m, float ux, float uy, float uz, bool has_force,
                          float gx, float gy, float gz, float sv, float so, float* f) {
  float r[Q];
  const float uu = ux * ux + uy * uy + uz * uz;
  r[0] = 0.0f;
  r[1] = -sv * (m[1] - uu);
  r[2] = -sv * m[2];
  r[3] = 0.0f;
  r[4] = -so * m[4];
  r[5] = 0.0f;
  r[6] = -so * m[6];
  r[7] = 0.0f;
  r[8] = -so * m[8];
  r[9] = -sv * (m[9] - (2.0f * ux * ux - uy * uy - uz * uz));
  r[10] = -sv * m[10];
  r[11] = -sv * (m[11] - (uy * uy - uz * uz));
  r[12] = -sv * m[12];
  r[13] = -sv * (m[13] - ux * uy);
  r[14] = -sv * (m[14] - uy * uz);
  r[15] = -sv * (m[15] - ux * uz);
  r[16] = -so * m[16];
  r[17] = -so * m[17];
  r[18] = -so * m[18];
  if (has_force) {
    float g[Q];
    const float uF = ux * gx + uy * gy + uz * gz;
#pragma unroll
    for (int l = 0; l < Q; ++l) {
      const float eF = (float)EX[l] * gx + (float)EY[l] * gy + (float)EZ[l] * gz;
      const float eu = (float)EX[l] * ux + (float)EY[l] * uy + (float)EZ[l] * uz;
      g[l] = wgt(l) * ((eF - uF) * (1.0f / 3.0f) + (eu * eF) * (1.0f / 9.0f));
    }
    float mg[Q];
    moments_M(g, mg);
    const float hv = 1.0f - 0.5f * sv, ho = 1.0f - 0.5f * so;
    r[0] += mg[0];
    r[1] += hv * mg[1];
    r[2] += hv * mg[2];
    r[3] += mg[3];
    r[4] += ho * mg[4];
    r[5] += mg[5];
    r[6] += ho * mg[6];
    r[7] += mg[7];
    r[8] += ho * mg[8];
#pragma unroll
    for (int k = 9; k <= 15; ++k) r[k] += hv * mg[k];
    r[16] += ho * mg[16];
    r[17] += ho * mg[17];
    r[18] += ho * mg[18];
  }
#pragma unroll
  for (int s = 0; s < Q; ++s) f[s] = F[s];
  add_invM(r, f);
}

FSI_HD void mrt_collide(const float* F, float rho, float ux, float uy, float uz, bool has_force,
                        float gx, float gy, float gz, float sv, float so, float* f) {
  (void)rho;
  float m[Q];
  moments_M(F, m);
  mrt_collide_m(F, m, ux, uy, uz, has_force, gx, gy, gz, sv, so, f);
}

// Regularised BGK (Latt & Chopard, Math. Comput. Simul. 72 (2006) 165):
// f = feq + (1 - 1/tau) f1, with f1_s = w_s 9/2 (e_s e_s - I/3) : Pi_neq and
// Pi_neq = sum_s (F_s - feq_s) e_s e_s, plus Guo's BGK forcing
// (1 - 1/(2 tau)) w_s [3 (e_s - u) + 9 (e_s . u) e_s] . G.  No reference
// counterpart: an option of the precision form, tau = 1/s_v.
FSI_HD void regularized_collide(const float* F, float rho, float ux, float uy, float uz, bool has_force,
                                float gx, float gy, float gz, float sv, float* f) {
  float fe[Q];
  float pxx = 0.f, pyy = 0.f, pzz = 0.f, pxy = 0.f, pxz = 0.f, pyz = 0.f;
#pragma unroll
  for (int s = 0; s < Q; ++s) {
    fe[s] = feq(s, rho, ux, uy, uz);
    const float d = F[s] - fe[s];
    const float ex = (float)EX[s], ey = (float)EY[s], ez = (float)EZ[s];
    pxx += ex * ex * d;
    pyy += ey * ey * d;
    pzz += ez * ez * d;
    pxy += ex * ey * d;
    pxz += ex * ez * d;
    pyz += ey * ez * d;
  }
  const float tr3 = (pxx + pyy + pzz) * (1.0f / 3.0f);
  const float omega = 1.0f - sv;
  const float fo = 1.0f - 0.5f * sv;
#pragma unroll
  for (int s = 0; s < Q; ++s) {
    const float ex = (float)EX[s], ey = (float)EY[s], ez = (float)EZ[s];
    // (e e - I/3) : Pi_neq
    const float qpi = ex * ex * pxx + ey * ey * pyy + ez * ez * pzz +
                      2.0f * (ex * ey * pxy + ex * ez * pxz + ey * ez * pyz) - tr3;
    float v = fe[s] + omega * wgt(s) * 4.5f * qpi;
    if (has_force) {
      const float eu = ex * ux + ey * uy + ez * uz;
      const float gxs = 3.0f * (ex - ux) + 9.0f * eu * ex;
      const float gys = 3.0f * (ey - uy) + 9.0f * eu * ey;
      const float gzs = 3.0f * (ez - uz) + 9.0f * eu * ez;
      v += fo * wgt(s) * (gxs * gx + gys * gy + gzs * gz);
    }
    f[s] = v;
  }
}

// streaming3 moments (fluid.py:1795-1814): rho = sum F, u = (sum e F)/rho + (force/2)/rho,
// evaluated as u = (sum e F + force/2) * (1/rho) with the sums taken from the
// MRT moment vector (m0 = rho, m3/m5/m7 = momentum): the fused substep uses
// the same expressions, so stored and recomputed moments agree bit for bit.
FSI_HD void moments_of(const float* F, float gx, float gy, float gz, float& rho, float& ux,
                       float& uy, float& uz) {
  float m[Q];
  moments_M(F, m);
  rho = m[0];
  const float inv = 1.0f / rho;
  ux = (m[3] + 0.5f * gx) * inv;
  uy = (m[5] + 0.5f * gy) * inv;
  uz = (m[7] + 0.5f * gz) * inv;
}

// ---------------------------------------------------------------------------
// Lattice indexing
// ---------------------------------------------------------------------------
struct Grid {
  int nx, ny, nz;
  long long N;
};

FSI_HD Grid make_grid(const FsiLbmParams& p) {
  Grid g;
  g.nx = p.nx;
  g.ny = p.ny;
  g.nz = p.nz;
  g.N = (long long)p.nx * p.ny * p.nz;
  return g;
}

FSI_HD int wrap(int i, int n) { return i < 0 ? n - 1 : (i > n - 1 ? 0 : i); }

FSI_HD long long cell_of(const Grid& g, int i, int j, int k) {
  return ((long long)i * g.ny + j) * g.nz + k;
}

// Cell c of the lattice (0 <= c < N <= 2^30, checked by the C ABI) in 32-bit
// division: a 64-bit division is a long software sequence on the GPU.
FSI_HD void coords_of(const Grid& g, long long c, int& i, int& j, int& k) {
  const uint32_t cc = (uint32_t)c, nz = (uint32_t)g.nz, ny = (uint32_t)g.ny;
  const uint32_t t = cc / nz;
  k = (int)(cc - t * nz);
  const uint32_t u = t / ny;
  j = (int)(t - u * ny);
  i = (int)u;
}

// periodic_index(i + e_s) (fluid.py:420-429)
FSI_HD long long neighbor(const Grid& g, int i, int j, int k, int s) {
  return cell_of(g, wrap(i + EX[s], g.nx), wrap(j + EY[s], g.ny), wrap(k + EZ[s], g.nz));
}

// a fluid cell on an x face whose populations the legacy BC overwrites
FSI_HD bool bc_face(const FsiLbmParams& p, int i) {
  return (i == 0 && p.bc_x_left != FSI_BC_PERIODIC) ||
         (i == p.nx - 1 && p.bc_x_right != FSI_BC_PERIODIC);
}

// ---------------------------------------------------------------------------
// Atomics / order-independent reductions
// ---------------------------------------------------------------------------
#ifdef __HIPCC__
// a lands in device global memory (grid and lattice arrays): addressed as
// such, the add is a global atomic even when the pointer was read from memory
__device__ __forceinline__ void atomic_add_f32(float* a, float v) {
  typedef __attribute__((address_space(1))) float GF;
  __hip_atomic_fetch_add((GF*)a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void atomic_add_f64(double* a, double v) { atomicAdd(a, v); }
__device__ __forceinline__ void atomic_max_nonneg(uint32_t* a, float v) {
  atomicMax(a, __float_as_uint(v));
}
__device__ __forceinline__ void atomic_add_i32(int32_t* a, int32_t v) { atomicAdd(a, v); }
#define FSI_ATOMIC_HD __device__ __forceinline__
#else
static inline uint32_t f2u(float v) {
  uint32_t u;
  memcpy(&u, &v, 4);
  return u;
}
static inline void atomic_add_f32(float* a, float v) { *a += v; }
static inline void atomic_add_f64(double* a, double v) { *a += v; }
static inline void atomic_max_nonneg(uint32_t* a, float v) {
  const uint32_t u = f2u(v);
  if (u > *a) *a = u;
}
static inline void atomic_add_i32(int32_t* a, int32_t v) { *a += v; }
#define FSI_ATOMIC_HD static inline
#endif

FSI_HD float norm3(float x, float y, float z) { return sqrtf(x * x + y * y + z * z); }

// ---------------------------------------------------------------------------
// LBM per-cell bodies
// ---------------------------------------------------------------------------

// init (fluid.py:329-344): F = feq(rho0, 0) in both buffers, rho = rho0, u = 0.
template <typename Pop>
FSI_HD void lbm_init_cell(const FsiLbmParams& p, const FsiLbmState& st, long long c) {
  const Grid g = make_grid(p);
  const Pop F0 = pop_of<Pop>(st.F[0]), F1 = pop_of<Pop>(st.F[1]);
#pragma unroll
  for (int s = 0; s < Q; ++s) {
    const float v = feq(s, p.rho0, 0.0f, 0.0f, 0.0f);
    F0.put(g.N, s, c, v);
    F1.put(g.N, s, c, v);
  }
  st.rho[c] = p.rho0;
  st.vel[c] = 0.0f;
  st.vel[g.N + c] = 0.0f;
  st.vel[2 * g.N + c] = 0.0f;
}

FSI_HD bool new_solid_flag(const FsiLbmParams& p, const FsiLbmState& st, long long c, int dynamic) {
  if (!dynamic) return st.solid[c] != 0;
  return st.static_solid[c] != 0 || st.solid_phi[c] >= p.dynamic_solid_threshold;
}

// update_dynamic_solid + reinitialize_new_fluid_cells (fluid.py:1848-1877) and
// the link masks of the streaming passes.  Neighbour flags are evaluated from
// static_solid / solid_phi directly, so every cell can be processed in one
// pass without waiting for its neighbours' flag writes.
template <typename Pop>
FSI_HD void lbm_rebuild_cell(const FsiLbmParams& p, const FsiLbmState& st, int dynamic, int cur,
                             long long c) {
  const Grid g = make_grid(p);
  int i, j, k;
  coords_of(g, c, i, j, k);
  const bool prev = st.solid[c] != 0;
  const bool now = new_solid_flag(p, st, c, dynamic);
  const Pop Fc = pop_of<Pop>(st.F[cur]);
  const Pop Fo = pop_of<Pop>(st.F[1 - cur]);
  if (dynamic) {
    st.old_solid[c] = prev ? 1 : 0;
    st.solid[c] = now ? 1 : 0;
    if (prev && !now) {
      // solid -> fluid: reset to feq(rho0, solid_vel or 0)
      float vx = st.solid_vel[c], vy = st.solid_vel[g.N + c], vz = st.solid_vel[2 * g.N + c];
      if (!(norm3(vx, vy, vz) > 0.0f)) vx = vy = vz = 0.0f;
#pragma unroll
      for (int s = 0; s < Q; ++s) {
        const float v = feq(s, p.rho0, vx, vy, vz);
        Fc.put(g.N, s, c, v);
        Fo.put(g.N, s, c, v);
      }
      st.rho[c] = p.rho0;
      st.vel[c] = vx;
      st.vel[g.N + c] = vy;
      st.vel[2 * g.N + c] = vz;
    } else if (!prev && now) {
      // fluid -> solid: the populations freeze at their current value
#pragma unroll
      for (int s = 0; s < Q; ++s) Fo.put(g.N, s, c, Fc.get(g.N, s, c));
    }
  }
  uint32_t info = 0;
  if (now) {
    info = CI_SOLID;
    if (st.solid_vel && (st.solid_vel[c] != 0.0f || st.solid_vel[g.N + c] != 0.0f || st.solid_vel[2 * g.N + c] != 0.0f))
      info |= CI_MOVING;
#pragma unroll
    for (int s = 1; s < Q; ++s) {
      // incoming link: the fluid cell c - e_s streams direction s into c
      const long long x = neighbor(g, i, j, k, LR[s]);
      if (!new_solid_flag(p, st, x, dynamic)) info |= 1u << s;
    }
  } else {
#pragma unroll
    for (int s = 1; s < Q; ++s) {
      const long long y = neighbor(g, i, j, k, s);
      if (new_solid_flag(p, st, y, dynamic)) {
        info |= 1u << s;
        if (st.solid_vel[y] != 0.0f || st.solid_vel[g.N + y] != 0.0f ||
            st.solid_vel[2 * g.N + y] != 0.0f)
          info |= CI_MOVING;
      }
    }
  }
  st.cellinfo[c] = info;
}

// Link scalars of the previous substep, seen from wall cell y.  The fluid cell
// x = y - e_s of link (x, s) bounced f_s into its own slot F_next[x][LR[s]],
// which no other cell writes; in the following substep that buffer is the
// read-only source.  The link's force on the wall is e_s c_s with
// c_s = f_s + bounced = 2 bounced - corr (fluid.py:510-519), so a wall at rest
// (corr = 0) reads c_s = 2 F_src[x][LR[s]] directly.  Fluid cells that bounce
// off a moving wall, and fluid cells on a BC face (whose populations the BC
// overwrites), store c_s in link_c instead; y takes it from there when it is
// itself moving (CI_MOVING on a solid cell) or x lies on a BC face.
// Per-axis BC-face flags of the planes i-1, i, i+1 select the case per link.
struct FaceFlags {
  bool m, o, p;
  FSI_HDM bool of_source(int s) const {  // is x = y - e_s on a BC face?
    return EX[s] > 0 ? m : (EX[s] < 0 ? p : o);
  }
};

FSI_HD FaceFlags face_flags(const FsiLbmParams& p, int i) {
  FaceFlags f;
  f.m = bc_face(p, i == 0 ? p.nx - 1 : i - 1);
  f.o = bc_face(p, i);
  f.p = bc_face(p, i == p.nx - 1 ? 0 : i + 1);
  return f;
}

// fp64 subcycle accumulator of cell y: 4 doubles (x, y, z, running max of |accum|).
FSI_HD void wall_accumulate(double* A, uint32_t y, float fx, float fy, float fz) {
  double* a = A + 4 * (size_t)y;
  const double ax = a[0] + (double)fx, ay = a[1] + (double)fy, az = a[2] + (double)fz;
  a[0] = ax;
  a[1] = ay;
  a[2] = az;
  const double n = (double)norm3((float)ax, (float)ay, (float)az);
  if (n > a[3]) a[3] = n;
}

// The same update from an accumulator record `a` already loaded, written to
// record o.
FSI_HD void wall_accumulate_into(double* o, const double* a, float fx, float fy, float fz) {
  const double ax = a[0] + (double)fx, ay = a[1] + (double)fy, az = a[2] + (double)fz;
  const double n = (double)norm3((float)ax, (float)ay, (float)az);
  o[0] = ax;
  o[1] = ay;
  o[2] = az;
  o[3] = n > a[3] ? n : a[3];
}

FSI_HD void wall_accumulate_loaded(double* A, uint32_t y, const double* a, float fx, float fy, float fz) {
  wall_accumulate_into(A + 4 * (size_t)y, a, fx, fy, fz);
}

// Previous-step velocity of cell c as the reference's `v` field holds it at
// Boundary_condition time: the stored moments, or the moments of F_src; 0 on
// a solid cell.  Every load is issued at once and none waits on another (the
// solid flag decides only at the end); 32-bit addressing.
template <typename Pop>
FSI_HD void prev_velocity_flat(const FsiLbmParams& p, const FsiLbmState& st, const Pop& Fs, uint32_t flags,
                               uint32_t c, float& ux, float& uy, float& uz) {
  const uint32_t N = (uint32_t)p.nx * (uint32_t)p.ny * (uint32_t)p.nz;
  const bool solid = (at32(st.cellinfo, c) & CI_SOLID) != 0;
  if (flags & FSI_SUB_STORED_MOMENTS) {
    ux = at32(st.vel, c);
    uy = at32(st.vel + N, c);
    uz = at32(st.vel + 2 * (size_t)N, c);
  } else {
    float Fn[Q];
#pragma unroll
    for (int s = 0; s < Q; ++s) Fn[s] = Fs.ld(N, s, c);
    float gx = p.force[0], gy = p.force[1], gz = p.force[2];
    if (flags & FSI_SUB_CELL_FORCE) {
      gx += at32(st.cell_force, c);
      gy += at32(st.cell_force + N, c);
      gz += at32(st.cell_force + 2 * (size_t)N, c);
    }
    float r;
    moments_of(Fn, gx, gy, gz, r, ux, uy, uz);
  }
  if (solid) ux = uy = uz = 0.0f;
}

// Periodic neighbour offsets of cell (i,j,k) (fluid.py:420-429) as signed
// 32-bit element deltas per axis; d(s) = index(i+ex, j+ey, k+ez) - index(i,j,k).
struct Nbr {
  int32_t xp, xm, yp, ym, zp, zm;
  FSI_HDM int32_t d(int s) const {
    return (EX[s] > 0 ? xp : (EX[s] < 0 ? xm : 0)) + (EY[s] > 0 ? yp : (EY[s] < 0 ? ym : 0)) +
           (EZ[s] > 0 ? zp : (EZ[s] < 0 ? zm : 0));
  }
};

FSI_HD Nbr nbr_of(const FsiLbmParams& p, int i, int j, int k) {
  const int32_t sx = p.ny * p.nz, sy = p.nz;
  Nbr n;
  n.xp = i == p.nx - 1 ? -(p.nx - 1) * sx : sx;
  n.xm = i == 0 ? (p.nx - 1) * sx : -sx;
  n.yp = j == p.ny - 1 ? -(p.ny - 1) * sy : sy;
  n.ym = j == 0 ? (p.ny - 1) * sy : -sy;
  n.zp = k == p.nz - 1 ? -(p.nz - 1) : 1;
  n.zm = k == 0 ? (p.nz - 1) : -1;
  return n;
}

// colission (fluid.py:399-413) of fluid cell c from its post-stream
// populations F: the force is LBMConfig.force plus cell_force under
// FSI_SUB_CELL_FORCE; the cell collides with the stored rho, vel under
// FSI_SUB_STORED_MOMENTS, else with the streaming3 moments (fluid.py:1795-1814)
// taken from the moment vector as in moments_of.  Returns the post-collision
// populations f and the density and velocity used.
FSI_HD void collide_cell(const FsiLbmParams& p, const FsiLbmState& st, uint32_t flags, uint32_t N, uint32_t c,
                         const float* F, float* f, float& rho, float& ux, float& uy, float& uz) {
  float gx = p.force[0], gy = p.force[1], gz = p.force[2];
  if (flags & FSI_SUB_CELL_FORCE) {
    gx += at32(st.cell_force, c);
    gy += at32(st.cell_force + N, c);
    gz += at32(st.cell_force + 2 * (size_t)N, c);
  }
  const bool has_force = (gx != 0.0f) || (gy != 0.0f) || (gz != 0.0f);
  float m[Q];
  moments_M(F, m);
  if (flags & FSI_SUB_STORED_MOMENTS) {
    rho = at32(st.rho, c);
    ux = at32(st.vel, c);
    uy = at32(st.vel + N, c);
    uz = at32(st.vel + 2 * (size_t)N, c);
  } else {
    rho = m[0];
    const float inv = 1.0f / rho;
    ux = (m[3] + 0.5f * gx) * inv;
    uy = (m[5] + 0.5f * gy) * inv;
    uz = (m[7] + 0.5f * gz) * inv;
  }
  if (p.collision == FSI_COLLIDE_REGULARIZED)
    regularized_collide(F, rho, ux, uy, uz, has_force, gx, gy, gz, p.s_v, f);
  else
    mrt_collide_m(F, m, ux, uy, uz, has_force, gx, gy, gz, p.s_v, p.s_other, f);
}

// Ladd moving-wall corrections of fluid cell c's links (fluid.py:489-519):
// cr[s] = -6 w_s rho (e_s . u_wall(c + e_s)) on each link s of a cell next to
// a moving wall, 0 otherwise (wall_corr of wall_dot, in that order).
FSI_HD float wall_dot(int s, float wx, float wy, float wz) {
  return (float)EX[s] * wx + (float)EY[s] * wy + (float)EZ[s] * wz;
}
FSI_HD float wall_corr(int s, float rho, float dot) { return -6.0f * wgt(s) * rho * dot; }

// e_s . u_wall(c + e_s) of each link s of fluid cell c, into dot[1..18] (0 on
// the other directions).  The wall velocities are loaded 6 links at a time,
// each load branch-free (a direction without a link reads the cell itself),
// so a cell next to the moving solid waits for memory 3 times rather than
// once per link.
FSI_HD void moving_wall_dots(const FsiLbmState& st, uint32_t N, uint32_t c, uint32_t info, const Nbr& nb,
                             float* dot) {
#pragma unroll
  for (int g0 = 1; g0 < Q; g0 += 6) {
    float w3[6][3];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      const uint32_t y = (info & (1u << (g0 + q))) ? c + (uint32_t)nb.d(g0 + q) : c;
      w3[q][0] = at32(st.solid_vel, y);
      w3[q][1] = at32(st.solid_vel + N, y);
      w3[q][2] = at32(st.solid_vel + 2 * (size_t)N, y);
    }
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      const int s = g0 + q;
      dot[s] = (info & (1u << s)) ? wall_dot(s, w3[q][0], w3[q][1], w3[q][2]) : 0.0f;
    }
  }
}

FSI_HD void moving_wall_corrections(const FsiLbmState& st, uint32_t N, uint32_t c, uint32_t info, const Nbr& nb,
                                    bool moving, float rho, float* cr) {
#pragma unroll
  for (int s = 0; s < Q; ++s) cr[s] = 0.0f;
  if (!moving) return;
  float dot[Q];
  moving_wall_dots(st, N, c, info, nb, dot);
#pragma unroll
  for (int s = 1; s < Q; ++s) cr[s] = wall_corr(s, rho, dot[s]);
}

// Boundary_condition_legacy (fluid.py:927-958) of BC-face cell (i,j,k): the
// density and velocity its 19 populations are reset to.  (ux,uy,uz): the
// cell's own velocity (a left pressure face keeps it); (obx,oby,obz): that of
// cell (nx-2, j, k) behind a right pressure face.  The right face is applied
// last.
FSI_HD void face_reset_state(const FsiLbmParams& p, const FsiLbmState& st, int i, int j, int k, float ux, float uy,
                             float uz, float obx, float oby, float obz, float& brho, float& bx, float& by,
                             float& bz) {
  const Grid g = make_grid(p);
  brho = 1.0f;
  bx = by = bz = 0.0f;
  if (i == 0 && p.bc_x_left != FSI_BC_PERIODIC) {
    if (p.bc_x_left == FSI_BC_VELOCITY) {
      brho = 1.0f;
      bx = p.vel_bc_x_left[0];
      by = p.vel_bc_x_left[1];
      bz = p.vel_bc_x_left[2];
    } else {
      brho = p.rho_bc_x_left;
      const long long c1 = cell_of(g, (g.nx > 1 ? 1 : 0), j, k);
      if (st.cellinfo[c1] & CI_SOLID) {
        bx = by = bz = 0.0f;
      } else {
        bx = ux;
        by = uy;
        bz = uz;
      }
    }
  }
  if (i == g.nx - 1 && p.bc_x_right != FSI_BC_PERIODIC) {
    if (p.bc_x_right == FSI_BC_VELOCITY) {
      brho = 1.0f;
      bx = p.vel_bc_x_right[0];
      by = p.vel_bc_x_right[1];
      bz = p.vel_bc_x_right[2];
    } else {
      brho = p.rho_bc_x_right;
      bx = obx;
      by = oby;
      bz = obz;
    }
  }
}

template <typename Pop>
FSI_HD void lbm_fluid_cell(const FsiLbmParams& p, const FsiLbmState& st, const Pop& Fs, int cur, int parity,
                           uint32_t flags, int i, int j, int k, uint32_t N, uint32_t c, uint32_t info, const Nbr& nb,
                           const FaceFlags& ff, const float* F, uint32_t outlet_flags);

FSI_HD uint32_t cell_info(const FsiLbmState& st, uint32_t c) { return at32(st.cellinfo, c); }

// Wall cell c (solid, with links) of a moving-wall substep: load the link
// scalars c_s of the previous substep into Fw[1..18] and the cell's
// accumulator record into acc.  A direction without a link reads element
// idle of its plane instead (masked at the sum): the cell itself when the
// wave's other lanes load its row anyway, else element 0, one line that all
// such lanes share.
template <typename Pop>
FSI_HD void wall_links_load(const FsiLbmState& st, const Pop& Fs, int parity, uint32_t N, uint32_t c,
                            uint32_t info, const Nbr& nb, const FaceFlags& ff, float* Fw, double* acc,
                            uint32_t idle) {
  const float* link = st.link_c[1 - parity];
  const bool wall_moving = (info & CI_MOVING) != 0;
  Fw[0] = 0.0f;
  if constexpr (Pop::kF32) {
    // fp32 populations: each link's source is a float plane either way, so
    // one branch-free load per direction (the factor 2 of a population read
    // back is applied at the sum) and no load waits for another
#pragma unroll
    for (int s = 1; s < Q; ++s) {
      const bool has = (info & (1u << s)) != 0;
      const uint32_t x = has ? c + (uint32_t)nb.d(LR[s]) : idle;
      const float* src = (wall_moving || ff.of_source(s)) ? link + (size_t)(s - 1) * N
                                                          : pop_plane_f32(Fs, N, LR[s]);
      Fw[s] = at32(src, x);
    }
  } else {
#pragma unroll
    for (int s = 1; s < Q; ++s) {
      Fw[s] = 0.0f;
      if (info & (1u << s)) {
        const uint32_t x = c + (uint32_t)nb.d(LR[s]);
        if (wall_moving || ff.of_source(s))
          Fw[s] = at32(link + (size_t)(s - 1) * N, x);
        else
          Fw[s] = 2.0f * Fs.ld(N, LR[s], x);
      }
    }
  }
  const double* a = st.hydro_accum + 4 * (size_t)c;
  acc[0] = a[0];
  acc[1] = a[1];
  acc[2] = a[2];
  acc[3] = a[3];
}

// The force of wall cell c's links, sum_s e_s c_s, added to its accumulator.
template <typename Pop>
FSI_HD void wall_links_accumulate(const FsiLbmState& st, uint32_t c, uint32_t info, const FaceFlags& ff,
                                  const float* Fw, const double* acc) {
  const bool wall_moving = (info & CI_MOVING) != 0;
  float fx = 0.0f, fy = 0.0f, fz = 0.0f;
#pragma unroll
  for (int s = 1; s < Q; ++s) {
    float cs = Fw[s];
    if constexpr (Pop::kF32) cs = !(info & (1u << s)) ? 0.0f : ((wall_moving || ff.of_source(s)) ? cs : 2.0f * cs);
    fx += (float)EX[s] * cs;
    fy += (float)EY[s] * cs;
    fz += (float)EZ[s] * cs;
  }
  wall_accumulate_loaded(st.hydro_accum, c, acc, fx, fy, fz);
}

// One fused LBM substep for cell (i,j,k) (push form): collide F[cur] at the
// cell, push the post-collision populations to the neighbours in F[1-cur],
// bounce them back at solid neighbours (half-way, with the Ladd moving-wall
// correction when FSI_SUB_MOVING_WALL), and apply the legacy x-face BCs.
// Every population of every fluid cell of F[1-cur] is written exactly once:
// by the upstream fluid neighbour, by the cell's own bounce-back, or by the BC.
// Wall (solid) cells gather the previous substep's link forces instead.
// All lattice addressing is 32-bit (at most 2^30 cells, checked by the ABI).
// outlet_flags: the flags the pressure outlet reads the velocity behind it
// with (~0u: `flags`); the batched penalty substep takes the stored moments
// of some cells and the moments of F_src of the others.
template <typename Pop, bool SPEC = false>
FSI_HD void lbm_substep_cell(const FsiLbmParams& p, const FsiLbmState& st, int cur, int parity,
                             uint32_t flags, int i, int j, int k, uint32_t outlet_flags = ~0u) {
  const uint32_t N = (uint32_t)p.nx * (uint32_t)p.ny * (uint32_t)p.nz;
  const uint32_t c = ((uint32_t)i * (uint32_t)p.ny + (uint32_t)j) * (uint32_t)p.nz + (uint32_t)k;
  const uint32_t info = at32(st.cellinfo, c);
  const Pop Fs = pop_of<Pop>(st.F[cur]);
  // SPEC: the populations are loaded with the flags, before the flags say
  // whether they are needed (a solid cell drops them): one memory round trip
  // for a fluid cell instead of two, for some extra traffic.  It pays where
  // several wave generations share the chip (a large lattice, the batched
  // form); a lattice that fits in one generation is faster without it.
  float F[Q];
  if (SPEC) {
#pragma unroll
    for (int s = 0; s < Q; ++s) F[s] = Fs.ld(N, s, c);
  }
  const Nbr nb = nbr_of(p, i, j, k);
  const bool mb = (flags & FSI_SUB_MOVING_WALL) != 0;
  const bool solid = (info & CI_SOLID) != 0;
  const bool gather = solid && mb && (flags & FSI_SUB_GATHER_PREV) && (info & CI_MASK);
  if (solid && !gather) return;
  const FaceFlags ff = face_flags(p, i);
  // One load phase for both kinds of cell, so that a wave holding fluid and
  // wall cells waits for memory once: fluid cells load their populations,
  // wall cells the link scalars c_s of their links (and their accumulator
  // record), and the fluid work is issued before the wall work.  With SPEC
  // the wall cells' loads go to registers of their own, Fw, so they do not
  // wait for the populations' loads into F to land.
  float Fw_[SPEC ? Q : 1];
  float* Fw = SPEC ? Fw_ : F;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  if (!solid) {
    if (!SPEC) {
#pragma unroll
      for (int s = 0; s < Q; ++s) F[s] = Fs.ld(N, s, c);
    }
  } else {
    wall_links_load<Pop>(st, Fs, parity, N, c, info, nb, ff, Fw, acc, c);
  }
  if (!solid) {
    lbm_fluid_cell<Pop>(p, st, Fs, cur, parity, flags, i, j, k, N, c, info, nb, ff, F,
                        outlet_flags == ~0u ? flags : outlet_flags);
    return;
  }
  wall_links_accumulate<Pop>(st, c, info, ff, Fw, acc);
}

// Collide, push, bounce back and apply the x-face BCs for fluid cell c whose
// post-stream populations F are already loaded (the rest of lbm_substep_cell).
template <typename Pop>
FSI_HD void lbm_fluid_cell(const FsiLbmParams& p, const FsiLbmState& st, const Pop& Fs, int cur, int parity,
                           uint32_t flags, int i, int j, int k, uint32_t N, uint32_t c, uint32_t info, const Nbr& nb,
                           const FaceFlags& ff, const float* F, uint32_t outlet_flags) {
  const bool mb = (flags & FSI_SUB_MOVING_WALL) != 0;
  const Pop Fd = pop_of<Pop>(st.F[1 - cur]);
  // The pressure outlet's reset needs the velocity of the cell behind it
  // (fluid.py:944-951): its loads go out now, with nothing waiting on them
  // until the reset, instead of as a second memory round trip at the end.
  const bool legacy_reset = p.open_bc == FSI_OPEN_BC_EQUILIBRIUM;
  const bool outlet_p = legacy_reset && i == p.nx - 1 && p.bc_x_right == FSI_BC_PRESSURE;
  float obx = 0.0f, oby = 0.0f, obz = 0.0f;
  if (outlet_p) prev_velocity_flat(p, st, Fs, outlet_flags, c - (uint32_t)(p.ny * p.nz), obx, oby, obz);
  float f[Q], rho, ux, uy, uz;
  collide_cell(p, st, flags, N, c, F, f, rho, ux, uy, uz);

  // The legacy reset rewrites every population of a face cell here; the open
  // modes leave the face cells streaming like inner cells (their unknown
  // populations are rewritten by the open-boundary pass after the substep).
  const bool legacy_bc = p.open_bc == FSI_OPEN_BC_EQUILIBRIUM;
  const bool self_face = ff.o && legacy_bc;
  const bool ok_xp = !(ff.p && legacy_bc);
  const bool ok_xm = !(ff.m && legacy_bc);
  float* link = mb ? st.link_c[parity] : nullptr;
  const bool diag = mb && (flags & FSI_SUB_DIAG);
  const bool moving = mb && (info & CI_MOVING);
  // walls cannot read these links back from F (a boundary rewrites them)
  const bool keep_links = mb && (moving || ff.o);
  // moving-wall corrections of the links, before the pushes
  float cr[Q];
  moving_wall_corrections(st, N, c, info, nb, moving, rho, cr);
  N = wave_uniform(N);
#pragma unroll
  for (int s = 0; s < Q; ++s) {
    if (s != 0 && (info & (1u << s))) {
      const float corr = cr[s];
      const float bounced = f[s] + corr;
      if (!self_face) Fd.st(N, LR[s], c, bounced);
      if (keep_links) at32(link + (size_t)(s - 1) * N, c) = f[s] + bounced;
      if (diag) at32(st.link_corr + (size_t)(s - 1) * N, c) = corr;
    } else {
      const bool ok = EX[s] > 0 ? ok_xp : (EX[s] < 0 ? ok_xm : !self_face);
      if (ok) Fd.st(N, s, c + (uint32_t)nb.d(s), f[s]);
    }
  }

  if (self_face) {
    float brho, bx, by, bz;
    face_reset_state(p, st, i, j, k, ux, uy, uz, obx, oby, obz, brho, bx, by, bz);
#pragma unroll
    for (int s = 0; s < Q; ++s) Fd.st(N, s, c, feq(s, brho, bx, by, bz));
  }
}

// ---------------------------------------------------------------------------
// Resident moving-wall subcycle (pull form).
//
// A whole subcycle of moving-wall substeps runs with the lattice on chip.
// Each substep has two phases with an exchange between them.  Slot (s, x)
// holds the value cell x + e_s takes as its population s:
//   A. every fluid cell c collides (as lbm_fluid_cell) and fills slot (s, c)
//      for each direction s = 1..18 with its post-collision f_s, or, on a link
//      (the neighbour y = c + e_s is solid), with the link scalar
//      c_s = f_s + (f_s + corr_s) that the wall y sums.  On a link it also
//      fills slot (LR[s], y), which only c reads (c = y + e_LR[s]) and which
//      the solid y leaves free, with its bounce-back f_s + corr_s;
//   B. every fluid cell takes each population q = 1..18 from slot (q, c - e_q)
//      and keeps its own f_0; a cell on a BC face resets all 19 to the
//      boundary equilibrium; a wall cell y sums e_s c_s, c_s from slot
//      (s, y - e_s), over its incoming links into its accumulator.
// Values, bounce-backs, face resets and wall sums are the expressions of the
// push form (lbm_substep_cell), so the two forms give the same bits.  A link
// into a wall at rest has corr_s = +-0, so f_s + (f_s + corr_s) equals
// 2 (f_s + corr_s), the value the push form's wall reads back from the
// populations.  The wall sum of substep t is added during substep t instead
// of t + 1, which keeps the order of the additions; the last substep's links
// are left to the finalize, as after the push form.
//
// Block decomposition: a workgroup owns RES_BI x RES_BJ x RES_NZ cells
// (i-planes x j-rows x a whole k-row of RES_NZ = nz cells), one cell per
// work-item; wave w holds row (li, lj) = (w / RES_BJ, w % RES_BJ) and lane k
// cell k of it.  Slots of cells in the block are in LDS, one plane per
// direction; a slot (s, x) of a cell x in another block is element x of halo
// plane s in device memory.
// ---------------------------------------------------------------------------
constexpr int RES_BI = 4, RES_BJ = 4, RES_NZ = 64;
constexpr int RES_THREADS = RES_BI * RES_BJ * RES_NZ;

FSI_HD bool res_shape_ok(int nx, int ny, int nz) {
  return nz == RES_NZ && nx >= RES_BI && ny >= RES_BJ && nx % RES_BI == 0 && ny % RES_BJ == 0;
}

FSI_HD int res_blocks(const FsiLbmParams& p) { return (p.nx / RES_BI) * (p.ny / RES_BJ); }

struct ResLane {
  int li, lj, k;  // position in the block
  int i, j;       // lattice position (k is shared)
  uint32_t c;     // cell
};

FSI_HD ResLane res_lane(const FsiLbmParams& p, uint32_t block, uint32_t t) {
  const uint32_t nbj = (uint32_t)p.ny / RES_BJ;
  const uint32_t bi = block / nbj, bj = block - bi * nbj;
  const uint32_t w = wave_uniform(t / RES_NZ);  // one row per wave: li, lj, i, j are wave-uniform
  ResLane L;
  L.k = (int)(t - w * RES_NZ);
  L.li = (int)(w / RES_BJ);
  L.lj = (int)(w - (uint32_t)L.li * RES_BJ);
  L.i = (int)(bi * RES_BI) + L.li;
  L.j = (int)(bj * RES_BJ) + L.lj;
  L.c = ((uint32_t)L.i * (uint32_t)p.ny + (uint32_t)L.j) * RES_NZ + (uint32_t)L.k;
  return L;
}

// Is the cell at (li + di, lj + dj) of the block's rows in the block?  (A
// wave-uniform test: li, lj are the wave's row.)
FSI_HD bool res_in_block(int li, int lj, int di, int dj) {
  return (unsigned)(li + di) < (unsigned)RES_BI && (unsigned)(lj + dj) < (unsigned)RES_BJ;
}

// The same test for a row known only by its edge class in each axis (0: the
// block's first row, 1: an inner row, 2: its last row), which decides it
// because the block spans more than 2 rows (offsets are -1, 0, +1).
FSI_HD constexpr bool res_class_in(int ci, int cj, int di, int dj) {
  return (ci == 0 ? di >= 0 : (ci == 2 ? di <= 0 : true)) && (cj == 0 ? dj >= 0 : (cj == 2 ? dj <= 0 : true));
}
FSI_HD int res_row_class(int l, int rows) { return l == 0 ? 0 : (l == rows - 1 ? 2 : 1); }
static_assert(RES_BI > 2 && RES_BJ > 2, "res_class_in needs blocks of at least 3 rows");

// Block slot of the cell at (li + di, lj + dj, k + dk) of the block (k wraps
// within the row, as nbr_of's periodic z offsets do for nz = RES_NZ).
FSI_HD int res_slot(int li, int lj, int k, int di, int dj, int dk) {
  return ((li + di) * RES_BJ + (lj + dj)) * RES_NZ + ((k + dk) & (RES_NZ - 1));
}

// The block's rows and one row around them in i and j (periodic), whole
// k-rows: the solid velocities a block's cells next to a moving wall read,
// held for the subcycle (they are fixed during it).
constexpr int RES_RI = RES_BI + 2, RES_RJ = RES_BJ + 2;
constexpr int RES_RING = RES_RI * RES_RJ * RES_NZ;

// Ring entry of the cell at (li + di, lj + dj, k + dk) of the block.
FSI_HD int res_ring_slot(int li, int lj, int k, int di, int dj, int dk) {
  return ((li + 1 + di) * RES_RJ + (lj + 1 + dj)) * RES_NZ + ((k + dk) & (RES_NZ - 1));
}

// Lattice cell of ring entry r of `block`.
FSI_HD uint32_t res_ring_cell(const FsiLbmParams& p, uint32_t block, int r) {
  const uint32_t nbj = (uint32_t)p.ny / RES_BJ;
  const int bi = (int)(block / nbj), bj = (int)(block - (uint32_t)bi * nbj);
  const int row = r / RES_NZ, k = r - row * RES_NZ;
  const int ri = row / RES_RJ, rj = row - ri * RES_RJ;
  const int i = wrap(bi * RES_BI + ri - 1, p.nx), j = wrap(bj * RES_BJ + rj - 1, p.ny);
  return ((uint32_t)i * (uint32_t)p.ny + (uint32_t)j) * RES_NZ + (uint32_t)k;
}

// Why the resident subcycle cannot run (nullptr: it can), apart from the
// device's capacity, which the GPU entry point checks.
inline const char* res_unsupported(const FsiLbmParams& p, const FsiLbmState& st, int cur, int n, uint32_t flags,
                                   const float* halo, const uint32_t* sync) {
  if (cur != 0 && cur != 1) return "cur must be 0 or 1";
  if (n < 1) return "n_substeps must be positive";
  if (!halo || !sync) return "halo and sync buffers are required";
  if (!(flags & FSI_SUB_MOVING_WALL)) return "the resident subcycle is the moving-wall subcycle (FSI_SUB_MOVING_WALL)";
  if (!st.link_c[0] || !st.link_c[1] || !st.link_corr || !st.hydro_accum)
    return "moving-wall link buffers and the accumulator are required";
  if (p.storage != FSI_STORE_F32) return "the resident subcycle covers FP32 populations";
  if (p.open_bc != FSI_OPEN_BC_EQUILIBRIUM) return "the resident subcycle covers the legacy x-face reset";
  if (!res_shape_ok(p.nx, p.ny, p.nz)) return "the resident subcycle covers nz = 64 with nx, ny multiples of 4";
  return nullptr;
}

// The flags of a batched instance's LBM_RESIDENT launch: those of the
// LBM_SUBSTEP_MB launches it replaces (stored moments at t = 0, the previous
// links gathered for t > 0, statistics at the last substep).
constexpr uint32_t BATCH_RESIDENT_FLAGS =
    FSI_SUB_MOVING_WALL | FSI_SUB_STORED_MOMENTS | FSI_SUB_GATHER_PREV | FSI_SUB_DIAG;

// Why an active batched instance marked resident cannot take LBM_RESIDENT
// (nullptr: it can, up to the device's capacity, which the launch checks).
// COUNTERS_CLEAR of one instance: `which` 0 projection, 1 penalty, 2
// reaction coupling counters, 3 the lattice's bounce-back counters.
FSI_HD void batch_clear_counters(const FsiBatchInstance& I, int which) {
  FsiCouplingCounters* cc = I.cc;
  if (which == 0) {
    cc->projected_mass = cc->projected_volume_raw = cc->projected_volume_clamped = 0.0;
    cc->max_phi_raw_bits = 0;
    cc->active_cell_count = 0;
  } else if (which == 1) {
    cc->active_force_cell_count = 0;
    cc->max_cell_force_bits = cc->max_hydro_force_bits = 0;
    for (int d = 0; d < 3; ++d) cc->net_cell_force[d] = cc->net_hydro_force[d] = 0.0;
  } else if (which == 2) {
    cc->active_reaction_particle_count = 0;
    cc->max_particle_reaction_bits = cc->max_grid_reaction_bits = 0;
    for (int d = 0; d < 3; ++d) cc->net_particle_reaction[d] = cc->net_grid_reaction[d] = 0.0;
  } else {
    FsiLbmCounters* k = I.ls.counters;
    k->bb_link_count = 0;
    k->bb_max_correction_bits = 0;
    for (int d = 0; d < 3; ++d) k->bb_net_solid_force[d] = 0.0;
  }
}

inline const char* batch_resident_invalid(const FsiBatchInstance& I) {
  if (!I.resident) return nullptr;
  if (I.mode != FSI_BATCH_MODE_MOVING_BOUNDARY || I.cell_force)
    return "a resident batched instance must be a moving-boundary instance without cell force";
  return res_unsupported(I.lp, I.ls, I.cur0, I.n_substeps, BATCH_RESIDENT_FLAGS, I.res_halo, I.res_sync);
}

// Phase A of fluid cell c from its post-stream populations F.  put(out)
// fills slot (s, c) with out[s] for s = 1..18; bounce(s, v) fills slot
// (LR[s], c + e_s) on a link s.
// dot(s): e_s . u_wall(c + e_s) of link s (moving_wall_dots), read only by a
// cell next to a moving wall.
// Returns the cell's f_0 in F[0] and the velocity the collision used in
// (ux, uy, uz) (a pressure outlet behind the cell resets to it).  On the
// subcycle's last substep (`last`) the cell leaves the link scalars of its
// moving-wall and BC-face links in link_c[parity] and, with FSI_SUB_DIAG, its
// corrections in link_corr, as the push form does, for the finalize.
template <typename Dot, typename Put, typename Bounce>
FSI_HD void res_collide_cell(const FsiLbmParams& p, const FsiLbmState& st, uint32_t flags, int parity, bool last,
                             uint32_t N, uint32_t c, uint32_t info, bool face_o, float* F, float& ux, float& uy,
                             float& uz, const Dot& dot, const Put& put, const Bounce& bounce) {
  float f[Q], rho;
  collide_cell(p, st, flags, N, c, F, f, rho, ux, uy, uz);
  const bool moving = (info & CI_MOVING) != 0;
  const bool keep = last && (moving || face_o);
  const bool diag = last && (flags & FSI_SUB_DIAG) != 0;
  float* link = st.link_c[parity];
  F[0] = f[0];
  // six links at a time (a cell next to a moving wall reads its wall
  // velocities for them), the next six scheduled only after
#pragma unroll
  for (int g0 = 1; g0 < Q; g0 += 6) {
    float cr[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) cr[q] = moving && (info & (1u << (g0 + q))) ? wall_corr(g0 + q, rho, dot(g0 + q)) : 0.0f;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      const int s = g0 + q;
      if (info & (1u << s)) {
        const float bounced = f[s] + cr[q];
        bounce(s, bounced);
        if (keep) at32(link + (size_t)(s - 1) * N, c) = f[s] + bounced;
        if (diag) at32(st.link_corr + (size_t)(s - 1) * N, c) = cr[q];
        f[s] = f[s] + bounced;
      }
    }
    sched_fence();
  }
  put(f);
}

// Phase B of cell L.c, a fluid cell or (when the substep gathers) a wall
// cell.  Both kinds take the same 18 slots, v[q] = slot (q, c - e_q)
// (res_pull_values): a fluid cell its populations 1..18 (F[0] is its own f_0
// from phase A), a wall cell the link scalars c_s of its incoming links s (its
// other slots hold nothing of this substep and are masked).  A fluid cell on a
// BC face resets all 19 populations to the boundary equilibrium instead
// (face_reset_state with its own velocity (ux, uy, uz) and, behind a right
// pressure face, (obx, oby, obz), that of cell (nx - 2, j, k), 0 if solid).  A
// wall cell sums e_s c_s in direction order with 0 on the directions without a
// link, as wall_links_accumulate does, and adds it to its accumulator record
// acc (x, y, z and the running max of the norm, updated in place).
template <typename Pull>
FSI_HD void res_pull_values(float* v, const Pull& pull) {
#pragma unroll
  for (int q = 1; q < Q; ++q) v[q] = pull(q);
}

FSI_HD void res_take_values(const FsiLbmParams& p, const FsiLbmState& st, const ResLane& L, uint32_t info,
                            bool face_o, float ux, float uy, float uz, float obx, float oby, float obz, float* F,
                            double* acc, const float* v) {
  if (!(info & CI_SOLID)) {
    if (face_o) {
      float brho, bx, by, bz;
      face_reset_state(p, st, L.i, L.j, L.k, ux, uy, uz, obx, oby, obz, brho, bx, by, bz);
#pragma unroll
      for (int s = 0; s < Q; ++s) F[s] = feq(s, brho, bx, by, bz);
      return;
    }
#pragma unroll
    for (int q = 1; q < Q; ++q) F[q] = v[q];
    return;
  }
  float fx = 0.0f, fy = 0.0f, fz = 0.0f;
#pragma unroll
  for (int s = 1; s < Q; ++s) {
    const float cs = (info & (1u << s)) ? v[s] : 0.0f;
    fx += (float)EX[s] * cs;
    fy += (float)EY[s] * cs;
    fz += (float)EZ[s] * cs;
  }
  wall_accumulate_into(acc, acc, fx, fy, fz);
}

template <typename Pull>
FSI_HD void res_pull_cell(const FsiLbmParams& p, const FsiLbmState& st, const ResLane& L, uint32_t info,
                          bool face_o, float ux, float uy, float uz, float obx, float oby, float obz, float* F,
                          double* acc, const Pull& pull) {
  float v[Q];
  res_pull_values(v, pull);
  res_take_values(p, st, L, info, face_o, ux, uy, uz, obx, oby, obz, F, acc, v);
}

// Link statistics of direction s for wall cell y after the last moving-wall
// substep (bb_*_by_dir, fluid.py:489-519): if y has an incoming link from
// x = y - e_s, add 1 link, the solid force e_s c_s (c_s = f_s + bounced, read
// back as the wall gathers it) and |correction| to acc = {links, fx, fy, fz,
// sum |corr|, max |corr|}.  s is a run-time direction here.
template <typename Pop>
FSI_HD void link_stats_cell(const FsiLbmParams& p, const FsiLbmState& st, int cur, int parity, int s, uint32_t y,
                            double* acc) {
  const uint32_t info = at32(st.cellinfo, y);
  if (!(info & CI_SOLID) || !(info & (1u << s))) return;
  int ex = 0, ey = 0, ez = 0, lr = 0;
#pragma unroll
  for (int q = 1; q < Q; ++q)
    if (q == s) {
      ex = EX[q];
      ey = EY[q];
      ez = EZ[q];
      lr = LR[q];
    }
  const uint32_t N = (uint32_t)p.nx * (uint32_t)p.ny * (uint32_t)p.nz;
  const uint32_t plane = (uint32_t)p.ny * (uint32_t)p.nz;
  const uint32_t i = y / plane, jk = y - i * plane;
  const uint32_t j = jk / (uint32_t)p.nz, k = jk - j * (uint32_t)p.nz;
  const Nbr nb = nbr_of(p, (int)i, (int)j, (int)k);
  const FaceFlags ff = face_flags(p, (int)i);
  // x = y - e_s = y + e_LR(s)
  const int32_t d = (ex < 0 ? nb.xp : (ex > 0 ? nb.xm : 0)) + (ey < 0 ? nb.yp : (ey > 0 ? nb.ym : 0)) +
                    (ez < 0 ? nb.zp : (ez > 0 ? nb.zm : 0));
  const uint32_t x = y + (uint32_t)d;
  const bool x_face = ex > 0 ? ff.m : (ex < 0 ? ff.p : ff.o);
  const Pop Fc = pop_of<Pop>(st.F[cur]);
  const float cs = ((info & CI_MOVING) || x_face) ? at32(st.link_c[parity] + (size_t)(s - 1) * N, x)
                                                   : 2.0f * Fc.ld(N, lr, x);
  const float ca = fabsf(at32(st.link_corr + (size_t)(s - 1) * N, x));
  acc[0] += 1.0;
  acc[1] += (double)((float)ex * cs);
  acc[2] += (double)((float)ey * cs);
  acc[3] += (double)((float)ez * cs);
  acc[4] += (double)ca;
  if ((double)ca > acc[5]) acc[5] = (double)ca;
}

// ---------------------------------------------------------------------------
// Boundary_condition on the x faces as a pass of its own (fluid.py:538-570)
// ---------------------------------------------------------------------------
constexpr int UNKNOWN_X_MIN[5] = {1, 7, 9, 11, 13};  // streamed in across x = 0 (fluid.py:26)
constexpr int UNKNOWN_X_MAX[5] = {2, 8, 10, 12, 14};  // streamed in across x = nx-1 (fluid.py:27)

// The moments of cell c that the reference's rho, v hold at Boundary_condition
// time (those of the previous streaming3): the stored ones, or the moments of
// the populations F_src the substep started from (rho0 and 0 on a solid cell).
template <typename Pop>
FSI_HD void prev_moments(const FsiLbmParams& p, const FsiLbmState& st, const Pop& Fs, uint32_t N, uint32_t flags,
                         uint32_t c, float& rho, float* u) {
  if (flags & FSI_SUB_STORED_MOMENTS) {
    rho = at32(st.rho, c);
    u[0] = at32(st.vel, c);
    u[1] = at32(st.vel + N, c);
    u[2] = at32(st.vel + 2 * (size_t)N, c);
    return;
  }
  if (at32(st.cellinfo, c) & CI_SOLID) {
    rho = p.rho0;
    u[0] = u[1] = u[2] = 0.0f;
    return;
  }
  float Fn[Q];
#pragma unroll
  for (int s = 0; s < Q; ++s) Fn[s] = Fs.ld(N, s, c);
  float gx = p.force[0], gy = p.force[1], gz = p.force[2];
  if (flags & FSI_SUB_CELL_FORCE) {
    gx += at32(st.cell_force, c);
    gy += at32(st.cell_force + N, c);
    gz += at32(st.cell_force + 2 * (size_t)N, c);
  }
  moments_of(Fn, gx, gy, gz, rho, u[0], u[1], u[2]);
}

// _regularized_population and _limited_regularized_population (fluid.py:
// 1027-1030, :1060-1104): feq(target) plus the neighbour's non-equilibrium
// part Fn_s - feq(rho_n, u_n); the limited form clips the target density and
// speed, the non-equilibrium part and the result, counting each clip in
// cnt = {rho, velocity, non-equilibrium, floor, reconstructed}.
FSI_HD float ob_regularized(const FsiLbmParams& p, bool limited, int s, float target_rho, const float* tu, float Fn_s,
                            float rho_n, const float* un, uint32_t* cnt) {
  float rl = target_rho;
  float ul[3] = {tu[0], tu[1], tu[2]};
  const bool lim = limited && p.ob_limiter;
  if (lim) {
    cnt[4] += 1;
    if (rl < p.ob_rho_min) {
      rl = p.ob_rho_min;
      cnt[0] += 1;
    }
    if (rl > p.ob_rho_max) {
      rl = p.ob_rho_max;
      cnt[0] += 1;
    }
    const float norm = sqrtf(ul[0] * ul[0] + ul[1] * ul[1] + ul[2] * ul[2]);
    if (norm > p.ob_u_max) {
      const float sc = p.ob_u_max / norm;
      ul[0] *= sc;
      ul[1] *= sc;
      ul[2] *= sc;
      cnt[1] += 1;
    }
  }
  float noneq = Fn_s - feq(s, rho_n, un[0], un[1], un[2]);
  if (lim) {
    if (noneq > p.ob_noneq_cap) {
      noneq = p.ob_noneq_cap;
      cnt[2] += 1;
    }
    if (noneq < -p.ob_noneq_cap) {
      noneq = -p.ob_noneq_cap;
      cnt[2] += 1;
    }
  }
  float out = feq(s, rl, ul[0], ul[1], ul[2]) + noneq;
  if (lim && p.ob_floor_enabled && out < p.ob_floor) {
    out = p.ob_floor;
    cnt[3] += 1;
  }
  return out;
}

// Boundary_condition for the face cell (j, k) of the left (right = false) or
// right x face, on F[buf] in place: the legacy reset of all populations
// (Boundary_condition_legacy, fluid.py:927-958) or the open modes' rewrite of
// the unknown populations (fluid.py:1349-1493).  The inner neighbour is the
// next cell inwards, or the face cell itself when that one is solid.  Each
// unknown population reads only its own slot before writing it, so the
// in-place rewrite is safe when the neighbour is the face cell.
template <typename Pop>
FSI_HD void lbm_open_bc_cell(const FsiLbmParams& p, const FsiLbmState& st, int buf, uint32_t flags, bool right, int j,
                             int k, uint32_t* cnt) {
  const int bc = right ? p.bc_x_right : p.bc_x_left;
  if (bc != FSI_BC_PRESSURE && bc != FSI_BC_VELOCITY) return;
  const uint32_t N = (uint32_t)p.nx * (uint32_t)p.ny * (uint32_t)p.nz;
  const uint32_t jk = (uint32_t)j * (uint32_t)p.nz + (uint32_t)k, plane = (uint32_t)p.ny * (uint32_t)p.nz;
  const uint32_t c0 = (right ? (uint32_t)(p.nx - 1) : 0u) * plane + jk;
  if (at32(st.cellinfo, c0) & CI_SOLID) return;
  const uint32_t c1 = (right ? (uint32_t)(p.nx - 2) : 1u) * plane + jk;
  const bool inner_solid = (at32(st.cellinfo, c1) & CI_SOLID) != 0;
  const Pop Fd = pop_of<Pop>(st.F[buf]), Fs = pop_of<Pop>(st.F[1 - buf]);
  if (p.open_bc == FSI_OPEN_BC_EQUILIBRIUM) {
    float brho = 1.0f, u[3];
    if (bc == FSI_BC_PRESSURE) {
      // the left face takes v[1] when cell 1 is solid and its own v
      // otherwise; the right face takes v[nx-2] (fluid.py:928-950)
      brho = right ? p.rho_bc_x_right : p.rho_bc_x_left;
      float r;
      prev_moments(p, st, Fs, N, flags, (right || inner_solid) ? c1 : c0, r, u);
    } else {
      const float* bv = right ? p.vel_bc_x_right : p.vel_bc_x_left;
      u[0] = bv[0];
      u[1] = bv[1];
      u[2] = bv[2];
    }
#pragma unroll
    for (int s = 0; s < Q; ++s) Fd.st(N, s, c0, feq(s, brho, u[0], u[1], u[2]));
    return;
  }
  const uint32_t cn = inner_solid ? c0 : c1;
  if (right && bc == FSI_BC_PRESSURE && p.open_bc == FSI_OPEN_BC_CONVECTIVE) {
    // _convective_outlet_population (fluid.py:1106-1111): linear
    // extrapolation from the two inner cells, then the population floor
    const uint32_t c2 = inner_solid ? c0 : (uint32_t)(p.nx - 3) * plane + jk;
    const bool c2_fluid = !(at32(st.cellinfo, c2) & CI_SOLID);
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      const int s = UNKNOWN_X_MAX[q];
      float x = Fd.ld(N, s, cn);
      if (c2_fluid) x = 2.0f * x - Fd.ld(N, s, c2);
      if (p.ob_limiter && p.ob_floor_enabled && x < p.ob_floor) x = p.ob_floor;
      Fd.st(N, s, c0, x);
    }
    return;
  }
  float rho_n, un[3];
  prev_moments(p, st, Fs, N, flags, cn, rho_n, un);
  float target_rho, tu[3];
  if (bc == FSI_BC_PRESSURE) {
    target_rho = right ? p.rho_bc_x_right : p.rho_bc_x_left;
    tu[0] = un[0];
    tu[1] = un[1];
    tu[2] = un[2];
  } else {
    target_rho = rho_n <= 1.0e-6f ? p.rho0 : rho_n;
    const float* bv = right ? p.vel_bc_x_right : p.vel_bc_x_left;
    tu[0] = bv[0];
    tu[1] = bv[1];
    tu[2] = bv[2];
  }
  const bool limited = p.open_bc == FSI_OPEN_BC_REGULARIZED_LIMITED;
#pragma unroll
  for (int q = 0; q < 5; ++q) {
    const int s = right ? UNKNOWN_X_MAX[q] : UNKNOWN_X_MIN[q];
    Fd.st(N, s, c0, ob_regularized(p, limited, s, target_rho, tu, Fd.ld(N, s, cn), rho_n, un, cnt));
  }
}

// streaming3 (fluid.py:1795-1814) on the populations in F[cur].
template <typename Pop>
FSI_HD void lbm_moments_cell(const FsiLbmParams& p, const FsiLbmState& st, int cur, uint32_t flags,
                             long long c) {
  const Grid g = make_grid(p);
  if (st.cellinfo[c] & CI_SOLID) {
    st.rho[c] = p.rho0;
    st.vel[c] = 0.0f;
    st.vel[g.N + c] = 0.0f;
    st.vel[2 * g.N + c] = 0.0f;
    return;
  }
  float F[Q];
  const Pop Fs = pop_of<Pop>(st.F[cur]);
#pragma unroll
  for (int s = 0; s < Q; ++s) F[s] = Fs.get(g.N, s, c);
  float gx = p.force[0], gy = p.force[1], gz = p.force[2];
  if (flags & FSI_SUB_CELL_FORCE) {
    gx += st.cell_force[c];
    gy += st.cell_force[g.N + c];
    gz += st.cell_force[2 * g.N + c];
  }
  float rho, ux, uy, uz;
  moments_of(F, gx, gy, gz, rho, ux, uy, uz);
  st.rho[c] = rho;
  st.vel[c] = ux;
  st.vel[g.N + c] = uy;
  st.vel[2 * g.N + c] = uz;
}

// Values one cell contributes to the finalize reductions.
struct FinalizeOut {
  int links;
  float max_corr;
  float mean_norm;
  float runmax;
  double hx, hy, hz;  // last substep's link forces on this cell, summed link by link in fp64 (net force)
  float fx, fy, fz;   // the same summed in fp32 in link order: the cell's hydro force
};

// finalize_moving_boundary_diagnostics / finalize_moving_boundary_force_accumulator
// (fluid.py:522-536, :476-486) for cell y after a substep that wrote its
// populations to F[cur] and its link scalars to link_c[parity]; writes hydro_force[y].
template <typename Pop>
FSI_HD FinalizeOut lbm_finalize_cell(const FsiLbmParams& p, const FsiLbmState& st, int cur, int parity,
                                     int n_samples, int subcycled, long long yl) {
  const Grid g = make_grid(p);
  const uint32_t N = (uint32_t)g.N, y = (uint32_t)yl;
  FinalizeOut o;
  o.links = 0;
  o.max_corr = 0.0f;
  o.mean_norm = 0.0f;
  o.runmax = 0.0f;
  o.hx = o.hy = o.hz = 0.0;
  o.fx = o.fy = o.fz = 0.0f;
  const uint32_t info = st.cellinfo[y];
  if ((info & CI_SOLID) && (info & CI_MASK)) {
    int i, j, k;
    coords_of(g, yl, i, j, k);
    const Nbr nb = nbr_of(p, i, j, k);
    const FaceFlags ff = face_flags(p, i);
    const Pop Fc = pop_of<Pop>(st.F[cur]);
    const float* link = st.link_c[parity];
    const bool wall_moving = (info & CI_MOVING) != 0;
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
#pragma unroll
    for (int s = 1; s < Q; ++s) {
      if (info & (1u << s)) {
        ++o.links;
        const uint32_t x = y + (uint32_t)nb.d(LR[s]);
        const float cs = (wall_moving || ff.of_source(s)) ? at32(link + (size_t)(s - 1) * N, x)
                                                            : 2.0f * Fc.ld(N, LR[s], x);
        fx += (float)EX[s] * cs;
        fy += (float)EY[s] * cs;
        fz += (float)EZ[s] * cs;
        // the net force sums link by link (the reference's bb_net_solid_force
        // is the sum of its per-direction link sums, fluid.py:522-536)
        o.hx += (double)((float)EX[s] * cs);
        o.hy += (double)((float)EY[s] * cs);
        o.hz += (double)((float)EZ[s] * cs);
        const float a = fabsf(at32(st.link_corr + (size_t)(s - 1) * N, x));
        if (a > o.max_corr) o.max_corr = a;
      }
    }
    o.fx = fx;
    o.fy = fy;
    o.fz = fz;
    if (subcycled) wall_accumulate(st.hydro_accum, y, fx, fy, fz);
  }
  if (subcycled) {
    const double inv = 1.0 / (double)(n_samples > 0 ? n_samples : 1);
    const double* a = st.hydro_accum + 4 * (size_t)y;
    const float mx = (float)(a[0] * inv);
    const float my = (float)(a[1] * inv);
    const float mz = (float)(a[2] * inv);
    st.hydro_force[y] = mx;
    st.hydro_force[g.N + y] = my;
    st.hydro_force[2 * g.N + y] = mz;
    o.mean_norm = norm3(mx, my, mz);
    o.runmax = (float)a[3];
  } else {
    st.hydro_force[y] = o.fx;
    st.hydro_force[g.N + y] = o.fy;
    st.hydro_force[2 * g.N + y] = o.fz;
  }
  return o;
}

// clear_moving_boundary_force_accumulator (fluid.py:461-466).
FSI_HD void lbm_mb_begin_cell(const FsiLbmParams& p, const FsiLbmState& st, long long c) {
  (void)p;
  double* a = st.hydro_accum + 4 * (size_t)c;
  a[0] = a[1] = a[2] = a[3] = 0.0;
}

// ---------------------------------------------------------------------------
// MPM (mpm/solid.py) per-particle / per-node bodies.
//
// From here to the diagnostics, no fma fusion: the elastic stress of strains
// near 1e-6 turns any rounding difference in the particle path into a
// visible velocity difference, so the particle and coupling arithmetic keeps
// the oracle's rounding step by step.  The lattice substep above keeps fma.
// ---------------------------------------------------------------------------
FSI_NO_CONTRACT_BEGIN
struct MpmGrid {
  int nx, ny, nz;
  long long G;
};
FSI_HD MpmGrid make_mgrid(const FsiMpmParams& p) {
  MpmGrid g;
  g.nx = p.gnx;
  g.ny = p.gny;
  g.nz = p.gnz;
  g.G = (long long)p.gnx * p.gny * p.gnz;
  return g;
}

// Quadratic B-spline weights about base = int(X - 0.5) (truncation, as the
// reference's ti.cast) with fx = X - base (solid.py:222-230).
FSI_HD void bspline(float X, int& base, float& fx, float* w) {
  base = (int)(X - 0.5f);
  fx = X - (float)base;
  const float a = 1.5f - fx, b = fx - 1.0f, d = fx - 0.5f;
  w[0] = 0.5f * (a * a);
  w[1] = 0.75f - b * b;
  w[2] = 0.5f * (d * d);
}

// A particle's quadratic B-spline stencil: base node and weights per axis.
struct Stencil {
  int base[3];
  float fx[3];
  float w[3][3];
};

FSI_HD Stencil particle_stencil(const float* x, long long P, long long p, float inv) {
  Stencil s;
#pragma unroll
  for (int d = 0; d < 3; ++d) bspline(x[d * P + p] * inv, s.base[d], s.fx[d], s.w[d]);
  return s;
}

// The stencil of a particle at position x (held by the caller).
FSI_HD Stencil stencil_at(const float* x, float inv) {
  Stencil s;
#pragma unroll
  for (int d = 0; d < 3; ++d) bspline(x[d] * inv, s.base[d], s.fx[d], s.w[d]);
  return s;
}

FSI_HD float det3(const float* A) {
  return A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
         A[2] * (A[3] * A[7] - A[4] * A[6]);
}

// Symmetric 3x3 Jacobi eigen-decomposition A = V diag(A_ii) V^T, cyclic
// sweeps (Golub & Van Loan, Matrix Computations, 4th ed., sec. 8.5).
FSI_HD void sym_eig3(float* A, float* V) {
#pragma unroll
  for (int e = 0; e < 9; ++e) V[e] = (e == 0 || e == 4 || e == 8) ? 1.0f : 0.0f;
#pragma unroll 1
  for (int sweep = 0; sweep < 6; ++sweep) {
#pragma unroll
    for (int pq = 0; pq < 3; ++pq) {
      const int pi = pq == 2 ? 1 : 0;
      const int qi = pq == 0 ? 1 : 2;
      const float apq = A[pi * 3 + qi];
      const float scale = fabsf(A[pi * 3 + pi]) + fabsf(A[qi * 3 + qi]);
      if (apq != 0.0f && fabsf(apq) > 1e-12f * scale) {
        const float theta = (A[qi * 3 + qi] - A[pi * 3 + pi]) / (2.0f * apq);
        const float t = (theta >= 0.0f ? 1.0f : -1.0f) / (fabsf(theta) + sqrtf(theta * theta + 1.0f));
        const float cs = 1.0f / sqrtf(t * t + 1.0f);
        const float sn = t * cs;
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          const float arp = A[r * 3 + pi], arq = A[r * 3 + qi];
          A[r * 3 + pi] = cs * arp - sn * arq;
          A[r * 3 + qi] = sn * arp + cs * arq;
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const float apc = A[pi * 3 + c], aqc = A[qi * 3 + c];
          A[pi * 3 + c] = cs * apc - sn * aqc;
          A[qi * 3 + c] = sn * apc + cs * aqc;
        }
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          const float vrp = V[r * 3 + pi], vrq = V[r * 3 + qi];
          V[r * 3 + pi] = cs * vrp - sn * vrq;
          V[r * 3 + qi] = sn * vrp + cs * vrq;
        }
      }
    }
  }
}

// Rotation R = U V^T of the reference's ti.svd(F) (det U = det V = +1, the
// sign of det F carried by the smallest singular value), returned together
// with D = F - R, the only part of R the fixed-corotated stress needs.
//
// det F > 0: R is the orthogonal polar factor, R = F S^{-1} with
// S = (F^T F)^{1/2}, so D = F (I - S^{-1}).  I - S^{-1} is computed from the
// Green-strain form F^T F - I = H + H^T + H^T H (H = F - I) so that D keeps
// full fp32 relative precision at small strain:
//   F^T F - I = V diag(m) V^T,  s = sqrt(1 + m),  I - S^{-1} = V diag(m / (s (s + 1))) V^T.
// det F <= 0: SVD from the eigenvectors of F^T F (u_i = F v_i / s_i for the
// two largest s_i, u_2 = u_0 x u_1), R = sum u_i v_i^T, D = F - R.
FSI_HD void rotation_and_deviation(const float* Fm, float* R, float* D) {
  const float detF = det3(Fm);
  const bool pos = detF > 0.0f;
  // the matrix to diagonalise: F^T F - I (det F > 0) or F^T F, one eigen
  // solve for both cases
  float A[9];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float hr0 = Fm[0 * 3 + r] - (r == 0 ? 1.0f : 0.0f), hc0 = Fm[0 * 3 + c] - (c == 0 ? 1.0f : 0.0f);
      const float hr1 = Fm[1 * 3 + r] - (r == 1 ? 1.0f : 0.0f), hc1 = Fm[1 * 3 + c] - (c == 1 ? 1.0f : 0.0f);
      const float hr2 = Fm[2 * 3 + r] - (r == 2 ? 1.0f : 0.0f), hc2 = Fm[2 * 3 + c] - (c == 2 ? 1.0f : 0.0f);
      const float hrc = Fm[r * 3 + c] - (r == c ? 1.0f : 0.0f), hcr = Fm[c * 3 + r] - (r == c ? 1.0f : 0.0f);
      A[r * 3 + c] = pos ? hrc + hcr + (hr0 * hc0 + hr1 * hc1 + hr2 * hc2)
                         : Fm[0 * 3 + r] * Fm[0 * 3 + c] + Fm[1 * 3 + r] * Fm[1 * 3 + c] + Fm[2 * 3 + r] * Fm[2 * 3 + c];
    }
  float V[9];
  sym_eig3(A, V);
  if (pos) {
    float d[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const float m = A[i * 4];
      const float s = sqrtf(1.0f + m);
      d[i] = m / (s * (s + 1.0f));
    }
    float K[9];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c)
        K[r * 3 + c] = V[r * 3 + 0] * d[0] * V[c * 3 + 0] + V[r * 3 + 1] * d[1] * V[c * 3 + 1] +
                       V[r * 3 + 2] * d[2] * V[c * 3 + 2];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        D[r * 3 + c] = Fm[r * 3 + 0] * K[0 * 3 + c] + Fm[r * 3 + 1] * K[1 * 3 + c] + Fm[r * 3 + 2] * K[2 * 3 + c];
        R[r * 3 + c] = Fm[r * 3 + c] - D[r * 3 + c];
      }
    return;
  }
  float lam[3] = {A[0], A[4], A[8]};
  int o0 = 0, o1 = 1, o2 = 2;
  if (lam[o0] < lam[o1]) { int t = o0; o0 = o1; o1 = t; }
  if (lam[o1] < lam[o2]) { int t = o1; o1 = o2; o2 = t; }
  if (lam[o0] < lam[o1]) { int t = o0; o0 = o1; o1 = t; }
  float v0[3] = {V[0 * 3 + o0], V[1 * 3 + o0], V[2 * 3 + o0]};
  float v1[3] = {V[0 * 3 + o1], V[1 * 3 + o1], V[2 * 3 + o1]};
  float v2[3] = {V[0 * 3 + o2], V[1 * 3 + o2], V[2 * 3 + o2]};
  const float dv = v0[0] * (v1[1] * v2[2] - v1[2] * v2[1]) - v0[1] * (v1[0] * v2[2] - v1[2] * v2[0]) +
                   v0[2] * (v1[0] * v2[1] - v1[1] * v2[0]);
  if (dv < 0.0f) { v2[0] = -v2[0]; v2[1] = -v2[1]; v2[2] = -v2[2]; }
  float u0[3], u1[3], u2[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    u0[r] = Fm[r * 3 + 0] * v0[0] + Fm[r * 3 + 1] * v0[1] + Fm[r * 3 + 2] * v0[2];
    u1[r] = Fm[r * 3 + 0] * v1[0] + Fm[r * 3 + 1] * v1[1] + Fm[r * 3 + 2] * v1[2];
  }
  float n0 = sqrtf(u0[0] * u0[0] + u0[1] * u0[1] + u0[2] * u0[2]);
  n0 = n0 > 1e-30f ? n0 : 1e-30f;
  u0[0] /= n0; u0[1] /= n0; u0[2] /= n0;
  const float d01 = u0[0] * u1[0] + u0[1] * u1[1] + u0[2] * u1[2];
  u1[0] -= d01 * u0[0]; u1[1] -= d01 * u0[1]; u1[2] -= d01 * u0[2];
  float n1 = sqrtf(u1[0] * u1[0] + u1[1] * u1[1] + u1[2] * u1[2]);
  n1 = n1 > 1e-30f ? n1 : 1e-30f;
  u1[0] /= n1; u1[1] /= n1; u1[2] /= n1;
  u2[0] = u0[1] * u1[2] - u0[2] * u1[1];
  u2[1] = u0[2] * u1[0] - u0[0] * u1[2];
  u2[2] = u0[0] * u1[1] - u0[1] * u1[0];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      R[r * 3 + c] = u0[r] * v0[c] + u1[r] * v1[c] + u2[r] * v2[c];
      D[r * 3 + c] = Fm[r * 3 + c] - R[r * 3 + c];
    }
}

// Three consecutive floats p[0..2] (a stencil row along k) as one load.
struct Row3 {
  float v[3];
};
FSI_HD Row3 load_row3(const float* p) {
  Row3 r;
  __builtin_memcpy(&r, p, sizeof(r));
  return r;
}

// Start k0 of the 3-float row load covering cells K0..K0+2 along an axis of
// n >= 3 cells, kept inside [0, n - 3].
FSI_HD int row_start(int K0, int n) { return K0 < 0 ? 0 : (K0 > n - 3 ? n - 3 : K0); }
// Cells K0, K0 + 1, K0 + 2 of a row loaded from k0 = K0 + d: cell K0 + c is
// element c - d, and a cell outside the loaded row (c - d outside 0..2)
// comes back as 0, to be masked off by the caller.  A bitwise select: no
// branch, and the values are the loaded ones bit for bit.
FSI_HD Row3 row_shift(const Row3& r, int d) {
  uint32_t u[3];
  __builtin_memcpy(u, r.v, sizeof(u));
  Row3 o;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int e = c - d;
    const uint32_t b = (u[0] & (e == 0 ? ~0u : 0u)) | (u[1] & (e == 1 ? ~0u : 0u)) | (u[2] & (e == 2 ? ~0u : 0u));
    __builtin_memcpy(&o.v[c], &b, sizeof(b));
  }
  return o;
}

FSI_HD void load_mat(const float* base, long long P, long long p, float* A) {
#pragma unroll
  for (int e = 0; e < 9; ++e) A[e] = base[e * P + p];
}
FSI_HD void store_mat(float* base, long long P, long long p, const float* A) {
#pragma unroll
  for (int e = 0; e < 9; ++e) base[e * P + p] = A[e];
}

FSI_HD bool inside3(int i, int j, int k, int nx, int ny, int nz) {
  return 0 <= i && i < nx && 0 <= j && j < ny && 0 <= k && k < nz;
}

// p2g (solid.py:220-249) for particle p; Jp[p] = det F.
// Scatter target of the particle-to-grid passes: add v to channel ch of grid
// node n = (I, J, K).  GlobalSink adds atomically into the grid arrays; the
// GPU kernels may also scatter into a workgroup-private LDS tile first.
struct GlobalSink {
  float* ch[5];
  FSI_HDM void add(int c, long long n, int, int, int, float v) const { atomic_add_f32(ch[c] + n, v); }
};

FSI_HD GlobalSink p2g_sink(const FsiMpmState& ms, long long G) {
  GlobalSink k;
  k.ch[0] = ms.grid_v;
  k.ch[1] = ms.grid_v + G;
  k.ch[2] = ms.grid_v + 2 * G;
  k.ch[3] = ms.grid_m;
  k.ch[4] = nullptr;
  return k;
}

FSI_HD GlobalSink reaction_sink(const FsiMpmState& ms, long long G) {
  GlobalSink k;
  k.ch[0] = ms.grid_f;
  k.ch[1] = ms.grid_f + G;
  k.ch[2] = ms.grid_f + 2 * G;
  k.ch[3] = k.ch[4] = nullptr;
  return k;
}

// Per-particle terms of p2g (solid.py:220-249).
struct P2gTerms {
  float aff[9];
  float mv[3];
  float mass;
  float J;  // det F
};

// The particle state p2g reads.
struct P2gIn {
  float F[9], C[9], v[3], mass, vol0;
};

FSI_HD P2gIn p2g_load(const FsiMpmState& ms, long long P, long long p) {
  P2gIn in;
  load_mat(ms.F, P, p, in.F);
  load_mat(ms.C, P, p, in.C);
  in.v[0] = ms.v[p];
  in.v[1] = ms.v[P + p];
  in.v[2] = ms.v[2 * P + p];
  in.mass = ms.mass[p];
  in.vol0 = ms.vol0[p];
  return in;
}

FSI_HD P2gTerms p2g_terms_of(const FsiMpmParams& mp, const P2gIn& in) {
  P2gTerms t;
  const float* Fm = in.F;
  const float* Cm = in.C;
  float R[9], Dv[9];
  rotation_and_deviation(Fm, R, Dv);
  const float J = det3(Fm);
  t.J = J;
  // F^{-T} = cofactor(F) / det F
  float FiT[9];
  FiT[0] = (Fm[4] * Fm[8] - Fm[5] * Fm[7]) / J;
  FiT[1] = (Fm[5] * Fm[6] - Fm[3] * Fm[8]) / J;
  FiT[2] = (Fm[3] * Fm[7] - Fm[4] * Fm[6]) / J;
  FiT[3] = (Fm[2] * Fm[7] - Fm[1] * Fm[8]) / J;
  FiT[4] = (Fm[0] * Fm[8] - Fm[2] * Fm[6]) / J;
  FiT[5] = (Fm[1] * Fm[6] - Fm[0] * Fm[7]) / J;
  FiT[6] = (Fm[1] * Fm[5] - Fm[2] * Fm[4]) / J;
  FiT[7] = (Fm[2] * Fm[3] - Fm[0] * Fm[5]) / J;
  FiT[8] = (Fm[0] * Fm[4] - Fm[1] * Fm[3]) / J;
  float Pk[9];
  const float lj = mp.la * (J - 1.0f) * J;
#pragma unroll
  for (int e = 0; e < 9; ++e) Pk[e] = 2.0f * mp.mu * Dv[e] + lj * FiT[e];
  const float mass = in.mass;
  const float sc = -mp.dt * in.vol0;
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      // (P F^T)_{rc}
      const float pf = Pk[r * 3 + 0] * Fm[c * 3 + 0] + Pk[r * 3 + 1] * Fm[c * 3 + 1] + Pk[r * 3 + 2] * Fm[c * 3 + 2];
      t.aff[r * 3 + c] = sc * pf * mp.inv_dx * mp.inv_dx + mass * Cm[r * 3 + c];
    }
  t.mass = mass;
  t.mv[0] = mass * in.v[0];
  t.mv[1] = mass * in.v[1];
  t.mv[2] = mass * in.v[2];
  return t;
}

// Per-particle terms of p2g (solid.py:220-249): node momentum is
// wt * (mv + aff * (x_node - x_p)), node mass wt * mass.  Writes Jp[p] = det F.
FSI_HD P2gTerms p2g_terms(const FsiMpmParams& mp, const FsiMpmState& ms, long long p) {
  const P2gTerms t = p2g_terms_of(mp, p2g_load(ms, mp.n_particles, p));
  ms.Jp[p] = t.J;
  return t;
}

template <typename Sink>
FSI_HD void mpm_p2g_particle(const FsiMpmParams& mp, const FsiMpmState& ms, long long p, const Stencil& sn,
                             const Sink& sink) {
  const MpmGrid g = make_mgrid(mp);
  const int* base = sn.base;
  const float* fx = sn.fx;
  const float(&w)[3][3] = sn.w;
  const P2gTerms t = p2g_terms(mp, ms, p);
  const float* aff = t.aff;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      for (int cc = 0; cc < 3; ++cc) {
        const int I = base[0] + a, Jn = base[1] + b, K = base[2] + cc;
        if (!inside3(I, Jn, K, g.nx, g.ny, g.nz)) continue;
        const float dpx = ((float)a - fx[0]) * mp.dx;
        const float dpy = ((float)b - fx[1]) * mp.dx;
        const float dpz = ((float)cc - fx[2]) * mp.dx;
        const float wt = w[0][a] * w[1][b] * w[2][cc];
        const float mx = t.mv[0] + aff[0] * dpx + aff[1] * dpy + aff[2] * dpz;
        const float my = t.mv[1] + aff[3] * dpx + aff[4] * dpy + aff[5] * dpz;
        const float mz = t.mv[2] + aff[6] * dpx + aff[7] * dpy + aff[8] * dpz;
        const long long n = ((long long)I * g.ny + Jn) * g.nz + K;
        sink.add(0, n, I, Jn, K, wt * mx);
        sink.add(1, n, I, Jn, K, wt * my);
        sink.add(2, n, I, Jn, K, wt * mz);
        sink.add(3, n, I, Jn, K, wt * t.mass);
      }
}

template <typename Sink>
FSI_HD void mpm_p2g_particle(const FsiMpmParams& mp, const FsiMpmState& ms, long long p, const Sink& sink) {
  mpm_p2g_particle(mp, ms, p, particle_stencil(ms.x, mp.n_particles, p, mp.inv_dx), sink);
}

// grid_update + apply_boundary_at_node (solid.py:252-259, :97-111) of a
// node (I, J, K) of mass m > 0: its velocity from momentum mom and force f.
FSI_HD void grid_node_velocity(const FsiMpmParams& mp, float m, const float* mom, const float* f, int I, int J, int K,
                               float* v) {
  const int idx[3] = {I, J, K};
  const int ext[3] = {mp.gnx, mp.gny, mp.gnz};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    float vd = mom[d] / m;
    vd += mp.dt * mp.gravity[d];
    vd += mp.dt * f[d] / m;
    v[d] = vd;
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (idx[d] < mp.bound && v[d] < 0.0f) v[d] = 0.0f;
    if (idx[d] > ext[d] - mp.bound && v[d] > 0.0f) v[d] = 0.0f;
  }
}

FSI_HD void mpm_grid_node(const FsiMpmParams& mp, const FsiMpmState& ms, long long n) {
  const MpmGrid g = make_mgrid(mp);
  const float m = ms.grid_m[n];
  if (!(m > 0.0f)) return;
  const int K = (int)(n % g.nz);
  const int Jn = (int)((n / g.nz) % g.ny);
  const int I = (int)(n / ((long long)g.nz * g.ny));
  const float mom[3] = {ms.grid_v[n], ms.grid_v[g.G + n], ms.grid_v[2 * g.G + n]};
  const float f[3] = {ms.grid_f[n], ms.grid_f[g.G + n], ms.grid_f[2 * g.G + n]};
  float v[3];
  grid_node_velocity(mp, m, mom, f, I, Jn, K, v);
#pragma unroll
  for (int d = 0; d < 3; ++d) ms.grid_v[d * g.G + n] = v[d];
}

// The per-particle part of g2p + clamp + dynamics controls + fixed pin
// (solid.py:262-315, :114-124): from the node velocities gv of particle p's
// stencil (node k = 9a + 3b + c; nodes outside the grid cleared in inm), its
// position x0 and deformation gradient Fm, the new x, v, C, F and det F.
FSI_HD void g2p_update(const FsiMpmParams& mp, const FsiMpmState& ms, long long p, const float* fx,
                       const float (&w)[3][3], const float (&gv)[27][3], uint32_t inm, const float* x0,
                       const float* Fm, float* x, float* v, float* nC, float* Fn, float& Jn) {
  FSI_NO_CONTRACT
  const long long P = mp.n_particles;
  float nv[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int e = 0; e < 9; ++e) nC[e] = 0.0f;
#pragma unroll
  for (int k = 0; k < 27; ++k) {
    if (!((inm >> k) & 1u)) continue;
    const int a = k / 9, b = (k / 3) % 3, cc = k % 3;
    const float dp[3] = {((float)a - fx[0]) * mp.dx, ((float)b - fx[1]) * mp.dx, ((float)cc - fx[2]) * mp.dx};
    const float wt = w[0][a] * w[1][b] * w[2][cc];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      nv[r] += wt * gv[k][r];
#pragma unroll
      for (int c = 0; c < 3; ++c) nC[r * 3 + c] += 4.0f * wt * gv[k][r] * dp[c] * mp.inv_dx * mp.inv_dx;
    }
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    v[d] = nv[d];
    x[d] = x0[d] + mp.dt * nv[d];
  }
  const float eps = mp.dx;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (x[d] < eps) {
      x[d] = eps;
      if (v[d] < 0.0f) v[d] = 0.0f;
    }
    if (x[d] > mp.domain_max[d] - eps) {
      x[d] = mp.domain_max[d] - eps;
      if (v[d] > 0.0f) v[d] = 0.0f;
    }
  }
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float a0 = (r == 0 ? 1.0f : 0.0f) + mp.dt * nC[r * 3 + 0];
      const float a1 = (r == 1 ? 1.0f : 0.0f) + mp.dt * nC[r * 3 + 1];
      const float a2 = (r == 2 ? 1.0f : 0.0f) + mp.dt * nC[r * 3 + 2];
      Fn[r * 3 + c] = a0 * Fm[0 * 3 + c] + a1 * Fm[1 * 3 + c] + a2 * Fm[2 * 3 + c];
    }
  Jn = det3(Fn);
  if (mp.damping_enabled) {
    float df = 1.0f - mp.velocity_damping * mp.dt;
    df = df > 0.0f ? df : 0.0f;
    v[0] *= df; v[1] *= df; v[2] *= df;
  }
  if (mp.lock_z_enabled) {
    x[2] = ms.fixed_x[2 * P + p];
    v[2] = 0.0f;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      nC[2 * 3 + d] = 0.0f;
      nC[d * 3 + 2] = 0.0f;
      Fn[2 * 3 + d] = 0.0f;
      Fn[d * 3 + 2] = 0.0f;
    }
    Fn[8] = 1.0f;
    Jn = det3(Fn);
  }
  if (ms.fixed_mask[p] != 0) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      x[d] = ms.fixed_x[d * P + p];
      v[d] = 0.0f;
    }
#pragma unroll
    for (int e = 0; e < 9; ++e) {
      nC[e] = 0.0f;
      Fn[e] = (e == 0 || e == 4 || e == 8) ? 1.0f : 0.0f;
    }
    Jn = 1.0f;
  }
}

// g2p + clamp + dynamics controls + fixed pin (solid.py:262-315, :114-124).
FSI_HD void mpm_g2p_particle(const FsiMpmParams& mp, const FsiMpmState& ms, long long p) {
  FSI_NO_CONTRACT
  const MpmGrid g = make_mgrid(mp);
  const long long P = mp.n_particles;
  int base[3];
  float fx[3], w[3][3];
#pragma unroll
  for (int d = 0; d < 3; ++d) bspline(ms.x[d * P + p] * mp.inv_dx, base[d], fx[d], w[d]);
  float Fm[9];
  load_mat(ms.F, P, p, Fm);
  // Loads first, then the sums in node order: 9 rows of 3 nodes along k, one
  // 3-float load per row and channel from a row start kept inside the grid
  // (a row outside in i or j reads row 0), nodes outside flagged off.
  const int k0 = row_start(base[2], g.nz);
  const int dk = k0 - base[2];
  Row3 rv[9][3];
  bool row_ok[9];
#pragma unroll
  for (int r = 0; r < 9; ++r) {
    const int I = base[0] + r / 3, Jn = base[1] + r % 3;
    row_ok[r] = 0 <= I && I < g.nx && 0 <= Jn && Jn < g.ny;
    const long long n0 = row_ok[r] ? ((long long)I * g.ny + Jn) * g.nz + k0 : 0;
    rv[r][0] = load_row3(ms.grid_v + n0);
    rv[r][1] = load_row3(ms.grid_v + g.G + n0);
    rv[r][2] = load_row3(ms.grid_v + 2 * g.G + n0);
  }
  float gv[27][3];
  uint32_t inm = 0u;
#pragma unroll
  for (int r = 0; r < 9; ++r) {
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
      const Row3 v = row_shift(rv[r][ch], dk);
#pragma unroll
      for (int c = 0; c < 3; ++c) gv[3 * r + c][ch] = v.v[c];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int K = base[2] + c;
      inm |= (row_ok[r] && 0 <= K && K < g.nz) ? (1u << (3 * r + c)) : 0u;
    }
  }
  const float x0[3] = {ms.x[p], ms.x[P + p], ms.x[2 * P + p]};
  float x[3], v[3], nC[9], Fn[9], Jn;
  g2p_update(mp, ms, p, fx, w, gv, inm, x0, Fm, x, v, nC, Fn, Jn);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    ms.x[d * P + p] = x[d];
    ms.v[d * P + p] = v[d];
  }
  store_mat(ms.C, P, p, nC);
  store_mat(ms.F, P, p, Fn);
  ms.Jp[p] = Jn;
}

// ---------------------------------------------------------------------------
// Coupling (coupling/projection.py, penalty.py, moving_boundary.py)
// ---------------------------------------------------------------------------

// project_particles (projection.py:38-64): scatter phi, mass, mass*velocity.
// Projection channels on the lattice: solid_phi, solid_mass, solid_vel x/y/z.
FSI_HD GlobalSink proj_sink(const FsiLbmState& ls, long long N) {
  GlobalSink k;
  k.ch[0] = ls.solid_phi;
  k.ch[1] = ls.solid_mass;
  k.ch[2] = ls.solid_vel;
  k.ch[3] = ls.solid_vel + N;
  k.ch[4] = ls.solid_vel + 2 * N;
  return k;
}

template <typename Sink>
FSI_HD void project_particle(const FsiLbmParams& lp, const FsiLbmState& ls, const FsiMpmParams& mp,
                             const FsiMpmState& ms, const FsiCouplingParams& cp, long long p, const Stencil& sn,
                             const Sink& sink) {
  const Grid g = make_grid(lp);
  const long long P = mp.n_particles;
  const int* base = sn.base;
  const float* fx = sn.fx;
  const float(&w)[3][3] = sn.w;
  float J = ms.Jp[p];
  J = J > 0.0f ? J : 0.0f;
  const float vol = ms.vol0[p] * J;
  const float mass = ms.mass[p];
  const float vx = ms.v[p] * cp.vel_scale_norm_to_lbm;
  const float vy = ms.v[P + p] * cp.vel_scale_norm_to_lbm;
  const float vz = ms.v[2 * P + p] * cp.vel_scale_norm_to_lbm;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      for (int cc = 0; cc < 3; ++cc) {
        const int I = base[0] + a, Jn = base[1] + b, K = base[2] + cc;
        if (!inside3(I, Jn, K, g.nx, g.ny, g.nz)) continue;
        const float wt = w[0][a] * w[1][b] * w[2][cc];
        const float mc = wt * mass;
        const float vf = wt * vol / cp.cell_volume_norm;
        const long long c = cell_of(g, I, Jn, K);
        sink.add(0, c, I, Jn, K, vf);
        sink.add(1, c, I, Jn, K, mc);
        sink.add(2, c, I, Jn, K, mc * vx);
        sink.add(3, c, I, Jn, K, mc * vy);
        sink.add(4, c, I, Jn, K, mc * vz);
      }
}

template <typename Sink>
FSI_HD void project_particle(const FsiLbmParams& lp, const FsiLbmState& ls, const FsiMpmParams& mp,
                             const FsiMpmState& ms, const FsiCouplingParams& cp, long long p, const Sink& sink) {
  project_particle(lp, ls, mp, ms, cp, p, particle_stencil(ms.x, mp.n_particles, p, cp.inv_dx_norm), sink);
}

struct ProjOut {
  float phi_raw, phi_clamped, mass;
  int active;
};

// normalize_projection (projection.py:66-91) for cell c.
FSI_HD ProjOut normalize_cell(const FsiLbmParams& lp, const FsiLbmState& ls, const FsiCouplingParams& cp,
                              long long c) {
  const Grid g = make_grid(lp);
  ProjOut o;
  o.phi_raw = ls.solid_phi[c];
  o.mass = ls.solid_mass[c];
  if (o.mass > cp.eps_mass) {
    ls.solid_vel[c] = ls.solid_vel[c] / o.mass;
    ls.solid_vel[g.N + c] = ls.solid_vel[g.N + c] / o.mass;
    ls.solid_vel[2 * g.N + c] = ls.solid_vel[2 * g.N + c] / o.mass;
  } else {
    ls.solid_vel[c] = 0.0f;
    ls.solid_vel[g.N + c] = 0.0f;
    ls.solid_vel[2 * g.N + c] = 0.0f;
  }
  float pc = o.phi_raw > 0.0f ? o.phi_raw : 0.0f;
  pc = pc < 1.0f ? pc : 1.0f;
  ls.solid_phi[c] = pc;
  o.phi_clamped = pc;
  o.active = pc > 1.0e-8f ? 1 : 0;
  return o;
}

// build_penalty_force (penalty.py:59-88) for cell c; returns the force.
FSI_HD void penalty_cell(const FsiLbmParams& lp, const FsiLbmState& ls, const FsiCouplingParams& cp,
                         long long c, float* force, int& active) {
  const Grid g = make_grid(lp);
  force[0] = force[1] = force[2] = 0.0f;
  active = 0;
  if (ls.solid[c] == 0) {
    const float phi = ls.solid_phi[c];
    if (phi > cp.penalty_phi_min) {
      const float k = cp.beta_lbm * phi * ls.rho[c];
#pragma unroll
      for (int d = 0; d < 3; ++d) force[d] = k * (ls.solid_vel[d * g.N + c] - ls.vel[d * g.N + c]);
      const float n = norm3(force[0], force[1], force[2]);
      if (n > cp.force_cap_lbm) {
        const float s = cp.force_cap_lbm / n;
        force[0] *= s; force[1] *= s; force[2] *= s;
      }
      active = 1;
    }
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    ls.cell_force[d * g.N + c] = force[d];
    ls.hydro_force[d * g.N + c] = -force[d];
  }
}

struct ReactOut {
  int active;
  float particle_norm;
  float max_contrib;
  float f[3];  // the particle's reaction force
  float g[3];  // what its scatter adds to the grid nodes inside the grid: f times the in-grid weight
};

// Sum of a stencil's weights over the nodes inside an nx x ny x nz grid: the
// product of the per-axis sums, as a node is inside iff it is along every axis.
FSI_HD float inside_weight(const Stencil& sn, int nx, int ny, int nz) {
  const int n[3] = {nx, ny, nz};
  float prod = 1.0f;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    float s = 0.0f;
#pragma unroll
    for (int o = 0; o < 3; ++o) {
      const int I = sn.base[d] + o;
      if (0 <= I && I < n[d]) s += sn.w[d][o];
    }
    prod *= s;
  }
  return prod;
}

// Lattice values the reaction force reads at the 27 stencil cells: 9 rows
// (a-major, then b) of 3 cells along k, each row one 3-float load per array
// (hydro_force x, y, z and, for mode 2, solid_phi) from a start kept inside
// the lattice; cell K0 + c of a row is element c of row_shift(row, d).  `in`
// flags the cells inside the lattice (bit 3 row + c).
struct HydroGather {
  Row3 h[9][4];
  int d;
  uint32_t in;
  bool shifted;  // rows loaded from a start moved inside the lattice (d != 0 possible)
};

// Issues the loads and returns at once: no loaded value is used here, so the
// loads stay in flight while the caller works until reaction_from_gather.
// checked = false: the caller knows every stencil cell is inside the lattice.
FSI_HD void hydro_gather_load(const FsiLbmParams& lp, const FsiLbmState& ls, int mode, const Stencil& sn,
                              bool checked, HydroGather& G) {
  const Grid g = make_grid(lp);
  const int* base = sn.base;
  const int k0 = checked ? row_start(base[2], g.nz) : base[2];
  G.d = k0 - base[2];
  G.in = 0u;
  G.shifted = checked;
#pragma unroll
  for (int r = 0; r < 9; ++r) {
    const int I = base[0] + r / 3, Jn = base[1] + r % 3;
    const bool row_ok = !checked || (0 <= I && I < g.nx && 0 <= Jn && Jn < g.ny);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int K = base[2] + c;
      G.in |= (row_ok && (!checked || (0 <= K && K < g.nz))) ? (1u << (3 * r + c)) : 0u;
    }
    const long long c0 = row_ok ? cell_of(g, I, Jn, k0) : 0;  // a row outside reads row 0
    G.h[r][0] = load_row3(ls.hydro_force + c0);
    G.h[r][1] = load_row3(ls.hydro_force + g.N + c0);
    G.h[r][2] = load_row3(ls.hydro_force + 2 * g.N + c0);
    if (mode == 2) G.h[r][3] = load_row3(ls.solid_phi + c0);
    else G.h[r][3].v[0] = G.h[r][3].v[1] = G.h[r][3].v[2] = 0.0f;
  }
}

// Reaction force on a particle of deformation J and rest volume vol0 from its
// gathered lattice values: mode 1 = penalty (penalty.py:90-129), mode 2 =
// moving boundary (moving_boundary.py:58-111).  o.f is the particle's
// reaction force, zero for a mode-2 particle outside the solid.
FSI_HD ReactOut reaction_from_gather(const FsiCouplingParams& cp, int mode, float J, float vol0, const Stencil& sn,
                                     const HydroGather& G) {
  ReactOut o;
  o.active = 0;
  o.particle_norm = 0.0f;
  o.max_contrib = 0.0f;
  o.f[0] = o.f[1] = o.f[2] = 0.0f;
  o.g[0] = o.g[1] = o.g[2] = 0.0f;
  const float(&w)[3][3] = sn.w;
  float hx = 0.0f, hy = 0.0f, hz = 0.0f, sphi = 0.0f;
#pragma unroll
  for (int r = 0; r < 9; ++r) {
    // unchecked rows start at the stencil: row_shift by 0 would return them bit for bit
    const Row3 rx = G.shifted ? row_shift(G.h[r][0], G.d) : G.h[r][0];
    const Row3 ry = G.shifted ? row_shift(G.h[r][1], G.d) : G.h[r][1];
    const Row3 rz = G.shifted ? row_shift(G.h[r][2], G.d) : G.h[r][2];
    const Row3 rp = G.shifted ? row_shift(G.h[r][3], G.d) : G.h[r][3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int k = 3 * r + c;
      if (!((G.in >> k) & 1u)) continue;
      const float wt = w[0][k / 9] * w[1][(k / 3) % 3] * w[2][k % 3];
      hx += wt * rx.v[c];
      hy += wt * ry.v[c];
      hz += wt * rz.v[c];
      if (mode == 2) sphi += wt * rp.v[c];
    }
  }
  if (mode == 2 && !(sphi > cp.mb_phi_min)) return o;
  J = J > 0.0f ? J : 0.0f;
  const float vol = vol0 * J;
  float pf[3] = {hx * cp.force_density_scale * vol * cp.reaction_scale,
                 hy * cp.force_density_scale * vol * cp.reaction_scale,
                 hz * cp.force_density_scale * vol * cp.reaction_scale};
  if (mode == 2) {
    // the link-area transfer's global factor, last as in link_area.py:150-157
    // (1 for the engineering transfer, an exact no-op)
    pf[0] *= cp.area_scale;
    pf[1] *= cp.area_scale;
    pf[2] *= cp.area_scale;
    float n = norm3(pf[0], pf[1], pf[2]);
    if (n > cp.mb_force_cap_norm) {
      const float s = cp.mb_force_cap_norm / n;
      pf[0] *= s; pf[1] *= s; pf[2] *= s;
      n = cp.mb_force_cap_norm;
    }
    o.particle_norm = n;
    o.active = n > 0.0f ? 1 : 0;
  }
  o.f[0] = pf[0];
  o.f[1] = pf[1];
  o.f[2] = pf[2];
  return o;
}

// Reaction force on particle p (its Jp as p2g left it).  checked = false: the
// caller knows every stencil cell is inside the lattice.
FSI_HD ReactOut reaction_force(const FsiLbmParams& lp, const FsiLbmState& ls, const FsiMpmState& ms,
                               const FsiCouplingParams& cp, int mode, long long p, const Stencil& sn,
                               bool checked = true) {
  HydroGather G;
  hydro_gather_load(lp, ls, mode, sn, checked, G);
  return reaction_from_gather(cp, mode, ms.Jp[p], ms.vol0[p], sn, G);
}

// Reaction transfer for particle p: the force, then its scatter to the MPM
// grid nodes of the stencil (o.max_contrib = the largest node contribution).
template <typename Sink>
FSI_HD ReactOut reaction_particle(const FsiLbmParams& lp, const FsiLbmState& ls, const FsiMpmParams& mp,
                                  const FsiMpmState& ms, const FsiCouplingParams& cp, int mode,
                                  long long p, const Stencil& sn, const Sink& sink) {
  const MpmGrid mg = make_mgrid(mp);
  ReactOut o = reaction_force(lp, ls, ms, cp, mode, p, sn);  // zero force: adds of zero
  const int* base = sn.base;
  const float(&w)[3][3] = sn.w;
  const float pf[3] = {o.f[0], o.f[1], o.f[2]};
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      for (int cc = 0; cc < 3; ++cc) {
        const int I = base[0] + a, Jn = base[1] + b, K = base[2] + cc;
        if (!inside3(I, Jn, K, mg.nx, mg.ny, mg.nz)) continue;
        const float wt = w[0][a] * w[1][b] * w[2][cc];
        const long long n = ((long long)I * mg.ny + Jn) * mg.nz + K;
        const float cx = wt * pf[0], cy = wt * pf[1], cz = wt * pf[2];
        sink.add(0, n, I, Jn, K, cx);
        sink.add(1, n, I, Jn, K, cy);
        sink.add(2, n, I, Jn, K, cz);
        const float cn = norm3(cx, cy, cz);
        if (cn > o.max_contrib) o.max_contrib = cn;
      }
  const float win = inside_weight(sn, mg.nx, mg.ny, mg.nz);
  o.g[0] = pf[0] * win;
  o.g[1] = pf[1] * win;
  o.g[2] = pf[2] * win;
  return o;
}


template <typename Sink>
FSI_HD ReactOut reaction_particle(const FsiLbmParams& lp, const FsiLbmState& ls, const FsiMpmParams& mp,
                                  const FsiMpmState& ms, const FsiCouplingParams& cp, int mode,
                                  long long p, const Sink& sink) {
  return reaction_particle(lp, ls, mp, ms, cp, mode, p, particle_stencil(ms.x, mp.n_particles, p, cp.inv_dx_norm),
                           sink);
}

// ---------------------------------------------------------------------------
// Particle binning and particle-ordered transfers.
//
// A particle's bin is its quadratic B-spline stencil base, base_d =
// trunc(x_d * inv_dx - 0.5) per axis (solid.py:222-224): the node where its
// 3 x 3 x 3 stencil starts.  Bins cover bases -2 .. n per axis, i.e. every base
// whose stencil reaches a node inside an n-node axis, offset by 2; one overflow
// bin (index nb) takes any other particle.  bin_start (nb + 2 entries) and
// bin_order (P entries) list each bin's particles in increasing index.  The
// contributors of node (I, J, K) are the particles of the 27 bins
// (I - a, J - b, K - c), a, b, c in {0, 1, 2}; walking those 27 lists merged
// in increasing particle index visits them in the order of a sequential
// scatter over the particles (the oracle's), whatever the thread schedule, so
// the node sums are deterministic and in that order.
// ---------------------------------------------------------------------------
struct BinGrid {
  int nx, ny, nz;  // node grid
  int by, bz;      // bins along y and z (n + 3)
  long long nb;    // bins without the overflow bin
};

FSI_HD BinGrid make_bins(int nx, int ny, int nz) {
  BinGrid b;
  b.nx = nx;
  b.ny = ny;
  b.nz = nz;
  b.by = ny + 3;
  b.bz = nz + 3;
  b.nb = (long long)(nx + 3) * b.by * b.bz;
  return b;
}

// Bin of a stencil base (the overflow bin nb when its stencil misses the grid).
FSI_HD int32_t bin_of_base(const BinGrid& b, const int* base) {
  if (base[0] < -2 || base[0] > b.nx || base[1] < -2 || base[1] > b.ny || base[2] < -2 || base[2] > b.nz)
    return (int32_t)b.nb;
  return (int32_t)(((long long)(base[0] + 2) * b.by + (base[1] + 2)) * b.bz + (base[2] + 2));
}

// Bin of particle p: its stencil base on spacing 1/inv, as bspline() takes it.
FSI_HD int32_t particle_bin(const BinGrid& b, const float* x, long long P, long long p, float inv) {
  int base[3];
  float fx, w[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) bspline(x[d * P + p] * inv, base[d], fx, w);
  return bin_of_base(b, base);
}

// Calls fn(p, a, b, c) for every particle p whose stencil covers node (I, J, K)
// at offset (a, b, c), in increasing p.  start/order: the binning above.
template <typename Fn>
FSI_HD void for_each_contributor(const BinGrid& bg, const int32_t* start, const int32_t* order, int I, int J, int K,
                                 Fn&& fn) {
  constexpr int32_t kNone = 0x7fffffff;
  int32_t cur[27], end[27], head[27];
#pragma unroll
  for (int o = 0; o < 27; ++o) {
    const int a = o / 9, b = (o / 3) % 3, c = o % 3;
    const long long bin = ((long long)(I - a + 2) * bg.by + (J - b + 2)) * bg.bz + (K - c + 2);
    cur[o] = start[bin];
    end[o] = start[bin + 1];
    head[o] = cur[o] < end[o] ? order[cur[o]] : kNone;
  }
  for (;;) {
    int32_t best = kNone;
    int bo = 0;
#pragma unroll
    for (int o = 0; o < 27; ++o)
      if (head[o] < best) {
        best = head[o];
        bo = o;
      }
    if (best == kNone) break;
    fn(best, bo / 9, (bo / 3) % 3, bo % 3);
#pragma unroll
    for (int o = 0; o < 27; ++o)
      if (o == bo) {
        ++cur[o];
        head[o] = cur[o] < end[o] ? order[cur[o]] : kNone;
      }
  }
}

// Per-particle record of the ordered MPM substep (floats): the affine momentum
// terms of p2g, the reaction force, and the stencil (fx, weights).
constexpr int OT_AFF = 0, OT_MV = 9, OT_MASS = 12, OT_PF = 13, OT_FX = 16, OT_W = 19;
constexpr int ORDERED_TERMS = 28;
static_assert(ORDERED_TERMS == FSI_ORDERED_TERMS_FLOATS, "ordered-substep record size");

// The particle pass of the ordered substep: p2g's terms (writes Jp = det F,
// solid.py:226-240) and, for mode 1 / 2, the reaction force of
// add_lbm_reaction_to_mpm_grid / add_moving_boundary_reaction_to_mpm_grid,
// sampled on the same stencil (the lattice spacing is the MPM grid's).
// o.max_contrib is the largest in-grid node contribution, o.g the in-grid sum.
FSI_HD ReactOut mpm_ordered_terms_particle(const FsiLbmParams& lp, const FsiLbmState& ls, const FsiMpmParams& mp,
                                           const FsiMpmState& ms, const FsiCouplingParams& cp, int mode, long long p,
                                           float* terms) {
  const Stencil sn = particle_stencil(ms.x, mp.n_particles, p, mp.inv_dx);
  const P2gTerms t = p2g_terms(mp, ms, p);
  ReactOut o;
  o.active = 0;
  o.particle_norm = o.max_contrib = 0.0f;
  o.f[0] = o.f[1] = o.f[2] = 0.0f;
  o.g[0] = o.g[1] = o.g[2] = 0.0f;
  if (mode) {
    o = reaction_force(lp, ls, ms, cp, mode, p, sn);
    const int n[3] = {mp.gnx, mp.gny, mp.gnz};
    float wmax = 1.0f;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      float m = 0.0f;
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const int I = sn.base[d] + q;
        if (0 <= I && I < n[d] && sn.w[d][q] > m) m = sn.w[d][q];
      }
      wmax *= m;
    }
    const float win = inside_weight(sn, mp.gnx, mp.gny, mp.gnz);
#pragma unroll
    for (int d = 0; d < 3; ++d) o.g[d] = o.f[d] * win;
    if (wmax > 0.0f) o.max_contrib = norm3(wmax * o.f[0], wmax * o.f[1], wmax * o.f[2]);
  }
  float* T = terms + (long long)ORDERED_TERMS * p;
#pragma unroll
  for (int e = 0; e < 9; ++e) T[OT_AFF + e] = t.aff[e];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    T[OT_MV + d] = t.mv[d];
    T[OT_PF + d] = o.f[d];
    T[OT_FX + d] = sn.fx[d];
#pragma unroll
    for (int q = 0; q < 3; ++q) T[OT_W + 3 * d + q] = sn.w[d][q];
  }
  T[OT_MASS] = t.mass;
  return o;
}

// Node n of the ordered substep: clear_grid, p2g and the reaction scatter
// (solid.py:213-249, penalty.py:111-129, moving_boundary.py:100-111) summed in
// particle order with the oracle's operation order, then grid_update
// (solid.py:252-259).  Every node is written, so no clear pass is needed.
FSI_HD void mpm_ordered_node(const FsiMpmParams& mp, const FsiMpmState& ms, const BinGrid& bg, const int32_t* start,
                             const int32_t* order, const float* terms, bool reaction, long long n) {
  FSI_NO_CONTRACT
  const MpmGrid g = make_mgrid(mp);
  const int K = (int)(n % g.nz);
  const int Jn = (int)((n / g.nz) % g.ny);
  const int I = (int)(n / ((long long)g.nz * g.ny));
  float v0 = 0.0f, v1 = 0.0f, v2 = 0.0f, m = 0.0f, f0 = 0.0f, f1 = 0.0f, f2 = 0.0f;
  for_each_contributor(bg, start, order, I, Jn, K, [&](int32_t p, int a, int b, int c) {
    const float* T = terms + (long long)ORDERED_TERMS * p;
    const float wt = T[OT_W + a] * T[OT_W + 3 + b] * T[OT_W + 6 + c];
    const float dp0 = ((float)a - T[OT_FX]) * mp.dx;
    const float dp1 = ((float)b - T[OT_FX + 1]) * mp.dx;
    const float dp2 = ((float)c - T[OT_FX + 2]) * mp.dx;
    const float* A = T + OT_AFF;
    v0 += wt * (T[OT_MV] + (A[0] * dp0 + A[1] * dp1 + A[2] * dp2));
    v1 += wt * (T[OT_MV + 1] + (A[3] * dp0 + A[4] * dp1 + A[5] * dp2));
    v2 += wt * (T[OT_MV + 2] + (A[6] * dp0 + A[7] * dp1 + A[8] * dp2));
    m += wt * T[OT_MASS];
    if (reaction) {
      f0 += wt * T[OT_PF];
      f1 += wt * T[OT_PF + 1];
      f2 += wt * T[OT_PF + 2];
    }
  });
  float v[3] = {v0, v1, v2};
  const float f[3] = {f0, f1, f2};
  if (m > 0.0f) {
    const int idx[3] = {I, Jn, K};
    const int ext[3] = {g.nx, g.ny, g.nz};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      float vd = v[d] / m;
      vd += mp.dt * mp.gravity[d];
      vd += mp.dt * f[d] / m;
      v[d] = vd;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (idx[d] < mp.bound && v[d] < 0.0f) v[d] = 0.0f;
      if (idx[d] > ext[d] - mp.bound && v[d] > 0.0f) v[d] = 0.0f;
    }
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    ms.grid_v[d * g.G + n] = v[d];
    ms.grid_f[d * g.G + n] = f[d];
  }
  ms.grid_m[n] = m;
}

// Lattice cell c of the ordered projection (projection.py:38-64): phi, mass
// and mass-weighted velocity summed in particle order with the oracle's
// operation order; normalize_cell follows.  Bins on the lattice spacing.
FSI_HD void project_ordered_cell(const FsiLbmParams& lp, const FsiLbmState& ls, const FsiMpmParams& mp,
                                 const FsiMpmState& ms, const FsiCouplingParams& cp, const BinGrid& bg,
                                 const int32_t* start, const int32_t* order, long long c) {
  FSI_NO_CONTRACT
  const Grid g = make_grid(lp);
  int I, Jn, K;
  coords_of(g, c, I, Jn, K);
  const long long P = mp.n_particles;
  float phi = 0.0f, sm = 0.0f, s0 = 0.0f, s1 = 0.0f, s2 = 0.0f;
  for_each_contributor(bg, start, order, I, Jn, K, [&](int32_t p, int a, int b, int cc) {
    int base[3];
    float fx[3], w[3][3];
#pragma unroll
    for (int d = 0; d < 3; ++d) bspline(ms.x[d * P + p] * cp.inv_dx_norm, base[d], fx[d], w[d]);
    float J = ms.Jp[p];
    J = J > 0.0f ? J : 0.0f;
    const float vol = ms.vol0[p] * J;
    const float mass = ms.mass[p];
    const float wt = w[0][a] * w[1][b] * w[2][cc];
    const float mc = wt * mass;
    phi += wt * vol / cp.cell_volume_norm;
    sm += mc;
    s0 += mc * (ms.v[p] * cp.vel_scale_norm_to_lbm);
    s1 += mc * (ms.v[P + p] * cp.vel_scale_norm_to_lbm);
    s2 += mc * (ms.v[2 * P + p] * cp.vel_scale_norm_to_lbm);
  });
  ls.solid_phi[c] = phi;
  ls.solid_mass[c] = sm;
  ls.solid_vel[c] = s0;
  ls.solid_vel[g.N + c] = s1;
  ls.solid_vel[2 * g.N + c] = s2;
}

FSI_NO_CONTRACT_END

// ---------------------------------------------------------------------------
// Diagnostics reductions.  The stats structs are arrays of doubles; each field
// has a combine op.  Per-cell / per-particle contributions are defined here so
// the GPU two-pass reduction and the serial emulation agree on semantics.
// ---------------------------------------------------------------------------
enum RedOp { RED_SUM = 0, RED_MIN = 1, RED_MAX = 2 };
constexpr int LBM_STATS_N = (int)(sizeof(FsiLbmStats) / sizeof(double));
constexpr int MPM_STATS_N = (int)(sizeof(FsiMpmStats) / sizeof(double));

FSI_HD int lbm_stats_op(int f) {
  // fluid_count, rho_min, rho_max, mass_total, max_v, force_norm_max, fluid_sum_v[3],
  // phi_weight_sum, phi_weighted_v[3], far_count, far_sum_v[3], max_cell_force_norm_all,
  // max_hydro_force_norm_all, net_cell_force[3], net_hydro_force[3], mass_rho_all
  return f == 1 ? RED_MIN : ((f == 2 || f == 4 || f == 5 || f == 17 || f == 18) ? RED_MAX : RED_SUM);
}
FSI_HD int mpm_stats_op(int f) {
  // min_x[3], max_x[3], max_speed, min_J, max_J, total_mass, mass_weighted_v[3]
  return (f <= 2 || f == 7) ? RED_MIN : ((f <= 6 || f == 8) ? RED_MAX : RED_SUM);
}
FSI_HD double red_identity(int op) {
  return op == RED_SUM ? 0.0 : (op == RED_MIN ? 1.0e300 : -1.0e300);
}
FSI_HD double red_combine(int op, double a, double b) {
  return op == RED_SUM ? a + b : (op == RED_MIN ? (a < b ? a : b) : (a > b ? a : b));
}

// One lattice cell's contribution (fluid.py:1931-1968; fsi_diagnostics.py:8-97).
FSI_HD void lbm_stats_cell(const FsiLbmParams& p, const FsiLbmState& st, long long c, double* a) {
  const Grid g = make_grid(p);
  const bool fluid = st.solid[c] == 0;
  const float vx = st.vel[c], vy = st.vel[g.N + c], vz = st.vel[2 * g.N + c];
  const float cfx = st.cell_force[c], cfy = st.cell_force[g.N + c], cfz = st.cell_force[2 * g.N + c];
  const float hfx = st.hydro_force[c], hfy = st.hydro_force[g.N + c], hfz = st.hydro_force[2 * g.N + c];
  const float cfn = norm3(cfx, cfy, cfz);
  const float phi = st.solid_phi[c];
  if (fluid) {
    const double r = st.rho[c];
    a[0] += 1.0;
    a[1] = r < a[1] ? r : a[1];
    a[2] = r > a[2] ? r : a[2];
    a[3] += r;
    const double vn = norm3(vx, vy, vz);
    a[4] = vn > a[4] ? vn : a[4];
    a[5] = (double)cfn > a[5] ? (double)cfn : a[5];
    a[6] += vx;
    a[7] += vy;
    a[8] += vz;
    if (phi <= 1.0e-6f) {
      a[13] += 1.0;
      a[14] += vx;
      a[15] += vy;
      a[16] += vz;
    }
  }
  if (phi > 1.0e-6f) {
    a[9] += phi;
    a[10] += (double)phi * vx;
    a[11] += (double)phi * vy;
    a[12] += (double)phi * vz;
  }
  a[17] = (double)cfn > a[17] ? (double)cfn : a[17];
  const double hn = norm3(hfx, hfy, hfz);
  a[18] = hn > a[18] ? hn : a[18];
  a[19] += cfx;
  a[20] += cfy;
  a[21] += cfz;
  a[22] += hfx;
  a[23] += hfy;
  a[24] += hfz;
  a[25] += st.rho[c];
}

// One particle's contribution (solid.py:342-357; fsi_diagnostics.py:117-124).
FSI_HD void mpm_stats_particle(const FsiMpmParams& mp, const FsiMpmState& ms, long long p, double* a) {
  const long long P = mp.n_particles;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double x = ms.x[d * P + p];
    a[d] = x < a[d] ? x : a[d];
    a[3 + d] = x > a[3 + d] ? x : a[3 + d];
  }
  const float vx = ms.v[p], vy = ms.v[P + p], vz = ms.v[2 * P + p];
  const double sp = norm3(vx, vy, vz);
  a[6] = sp > a[6] ? sp : a[6];
  const double J = ms.Jp[p];
  a[7] = J < a[7] ? J : a[7];
  a[8] = J > a[8] ? J : a[8];
  const double m = ms.mass[p];
  a[9] += m;
  a[10] += m * vx;
  a[11] += m * vy;
  a[12] += m * vz;
}


// ---------------------------------------------------------------------------
// Batched form: per-instance substep semantics shared by the GPU launch and
// the serial emulation.
// ---------------------------------------------------------------------------
// Workgroup size of a batch domain's launches (particle ops use the smaller
// particle workgroups) and the number of workgroups an instance needs there.
FSI_HD int batch_block_size(int domain) { return domain == FSI_BATCH_DOMAIN_PARTICLES ? FSI_PARTICLE_BLOCK : 256; }

FSI_HD long long batch_domain_blocks(const FsiBatchInstance& I, int domain) {
  if (domain == FSI_BATCH_DOMAIN_INSTANCES) return 1;
  if (domain == FSI_BATCH_DOMAIN_BOX) {
    const long long cb = ((long long)I.lp.nx * I.lp.ny * I.lp.nz + 255) / 256;
    return cb < FSI_BATCH_BOX_BLOCKS ? cb : FSI_BATCH_BOX_BLOCKS;
  }
  if (domain == FSI_BATCH_DOMAIN_PLANES)  // one i-plane per run of tiles, as the single-instance substep
    return I.resident ? 0 : (long long)I.lp.nx * (((long long)I.lp.ny * I.lp.nz + 255) / 256);
  long long n = domain == FSI_BATCH_DOMAIN_CELLS ? (long long)I.lp.nx * I.lp.ny * I.lp.nz
              : domain == FSI_BATCH_DOMAIN_NODES ? (long long)I.mp.gnx * I.mp.gny * I.mp.gnz
                                                 : (long long)I.mp.n_particles;
  const int bs = batch_block_size(domain);
  return (n + bs - 1) / bs;
}

FSI_HD long long batch_domain_size(const FsiBatchInstance& I, int domain) {
  if (domain == FSI_BATCH_DOMAIN_CELLS) return (long long)I.lp.nx * I.lp.ny * I.lp.nz;
  if (domain == FSI_BATCH_DOMAIN_NODES) return (long long)I.mp.gnx * I.mp.gny * I.mp.gnz;
  return (long long)I.mp.n_particles;
}

// Flags of instance I at substep t of the exchange: moving-boundary instances
// collide with stored moments at t = 0 and gather the previous substep's wall
// forces after it, recording link diagnostics on the last substep; penalty /
// none instances refresh their moments after every substep.
FSI_HD uint32_t batch_substep_flags(const FsiBatchInstance& I, int t) {
  uint32_t f = I.cell_force ? FSI_SUB_CELL_FORCE : 0u;
  if (I.mode == FSI_BATCH_MODE_MOVING_BOUNDARY) {
    f |= FSI_SUB_MOVING_WALL;
    f |= t == 0 ? FSI_SUB_STORED_MOMENTS : FSI_SUB_GATHER_PREV;
    if (t == I.n_substeps - 1) f |= FSI_SUB_DIAG;
  } else if (t == 0) {
    f |= FSI_SUB_STORED_MOMENTS;
  }
  return f;
}

FSI_HD int batch_cur(const FsiBatchInstance& I, int t) { return I.cur0 ^ (t & 1); }

// Workgroups are handed to the 8 XCDs round-robin by dispatch order; the
// logical tile of dispatch slot b of an n-workgroup launch is chosen so that
// XCD x receives one contiguous run of tiles (and keeps their L2 lines).
constexpr unsigned NXCD = 8;

FSI_HD unsigned xcd_tile(unsigned b, unsigned n) {
  const unsigned per = n / NXCD, rem = n % NXCD, x = b % NXCD;
  return x * per + (x < rem ? x : rem) + b / NXCD;
}

// Bounding boxes of the batched coupling passes (FSI_BATCH_BOX_*): `b` points
// at one box record, [0] valid, [1..6] cur (lo, hi), [7..12] prev (lo, hi).
constexpr int BOX_CUR = 1, BOX_PREV = 7;

FSI_HD void box_roll(int32_t* b, int nx, int ny, int nz) {
  const int ext[3] = {nx, ny, nz};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    b[BOX_PREV + d] = b[0] ? b[BOX_CUR + d] : 0;
    b[BOX_PREV + 3 + d] = b[0] ? b[BOX_CUR + 3 + d] : ext[d] - 1;
    b[BOX_CUR + d] = 0x7fffffff;
    b[BOX_CUR + 3 + d] = -0x7fffffff - 1;
  }
  b[0] = 1;
}

// Whether (i, j, k) lies in the box lo = r[0..2], hi = r[3..5].
FSI_HD bool in_box(const int32_t* r, int i, int j, int k) {
  return r[0] <= i && i <= r[3] && r[1] <= j && j <= r[4] && r[2] <= k && k <= r[5];
}

// The smallest box holding boxes r and q.
FSI_HD void box_hull(const int32_t* r, const int32_t* q, int32_t* out) {
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    out[d] = r[d] < q[d] ? r[d] : q[d];
    out[3 + d] = r[3 + d] > q[3 + d] ? r[3 + d] : q[3 + d];
  }
}

// Whether the i-planes [i0, i1] meet box r.
FSI_HD bool planes_meet_box(const int32_t* r, long long i0, long long i1) { return i1 >= r[0] && i0 <= r[3]; }

// Where a penalty / none instance's substep t >= 1 takes rho and u from:
// the batched exchange runs no moments pass between its LBM steps.  A
// penalty instance's PENALTY pass has stored the moments (with the previous
// step's cell force, as the pass after the previous step would have) in the
// box it covers, the hull of the projection boxes, and in the plane the
// pressure outlet reads (i = nx - 2); everywhere else the cell force of both
// steps is zero and the moments of F_src are the same values.  A none
// instance takes the moments of F_src everywhere.  Whether penalty cell
// (i, j, k) of substep t >= 1 takes the stored moments:
FSI_HD bool batch_penalty_stored(const FsiBatchInstance& I, int i, int j, int k) {
  // the hull's bounds one at a time (uniform values; no array to keep)
  const int32_t* c = I.boxes + FSI_BATCH_BOX_PROJ + BOX_CUR;
  const int32_t* q = I.boxes + FSI_BATCH_BOX_PROJ + BOX_PREV;
  const int lo0 = c[0] < q[0] ? c[0] : q[0], hi0 = c[3] > q[3] ? c[3] : q[3];
  const int lo1 = c[1] < q[1] ? c[1] : q[1], hi1 = c[4] > q[4] ? c[4] : q[4];
  const int lo2 = c[2] < q[2] ? c[2] : q[2], hi2 = c[5] > q[5] ? c[5] : q[5];
  return lo0 <= i && i <= hi0 && lo1 <= j && j <= hi1 && lo2 <= k && k <= hi2;
}

// The flags of cell (i, j, k) of instance I at substep t and the flags its
// pressure-outlet reset reads the cell behind it with.
FSI_HD void batch_moment_source(const FsiBatchInstance& I, int t, uint32_t& flags, uint32_t& outlet_flags, int i,
                                int j, int k) {
  outlet_flags = flags;
  if (t == 0 || I.mode != FSI_BATCH_MODE_PENALTY) return;
  if (batch_penalty_stored(I, i, j, k)) flags |= FSI_SUB_STORED_MOMENTS;
  outlet_flags |= FSI_SUB_STORED_MOMENTS;
}

}  // namespace fsi
