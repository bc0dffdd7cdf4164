// Device helpers shared by the pair-sum and Krylov kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include "lh_types.h"

namespace lh {

// |d|^2 without fused multiply-add. Every kernel (and the host near-list
// builder) that decides whether a pair belongs to the near list evaluates
// exactly this expression, so the far-field exclusion and the near list can
// never disagree.
__host__ __device__ __forceinline__ double r2_exact(double dx, double dy, double dz) {
#pragma clang fp contract(off)
  double a = dx * dx;
  double b = dy * dy;
  double c = dz * dz;
  return (a + b) + c;
}

// The same for FP32 positions: the exclusion test of every far-field
// kernel and of the host near-list builder. IEEE single-precision subtract,
// multiply and add give identical bits on the host and on the GPU.
__host__ __device__ __forceinline__ float r2f_exact(float dx, float dy, float dz) {
#pragma clang fp contract(off)
  float a = dx * dx;
  float b = dy * dy;
  float c = dz * dz;
  return (a + b) + c;
}

// sin and cos of an FP32 phase, |x| < 1e4: Cody-Waite reduction by pi/2
// with a two-part constant (hi = pi/2 rounded to FP32, lo = pi/2 - hi),
// then the Taylor series of sin (to r^9) and cos (to r^8) on
// [-pi/4, pi/4], whose truncation error is below 2e-9, and the quadrant
// swap/sign of sin(q pi/2 + r), cos(q pi/2 + r).
__device__ __forceinline__ void sincos_f32(float x, float& s, float& c) {
  const float q = rintf(x * 0.63661977236758134f);
  float r = fmaf(-q, 1.57079637050628662f, x);
  r = fmaf(-q, -4.37113883e-8f, r);
  const float r2 = r * r;
  float ps = fmaf(r2, 1.0f / 362880.0f, -1.0f / 5040.0f);
  ps = fmaf(r2, ps, 1.0f / 120.0f);
  ps = fmaf(r2, ps, -1.0f / 6.0f);
  const float sr = fmaf(r * r2, ps, r);
  float pc = fmaf(r2, 1.0f / 40320.0f, -1.0f / 720.0f);
  pc = fmaf(r2, pc, 1.0f / 24.0f);
  pc = fmaf(r2, pc, -0.5f);
  const float cr = fmaf(r2, pc, 1.0f);
  const int qi = (int)q;
  const float ss = (qi & 1) ? cr : sr;
  const float cc = (qi & 1) ? sr : cr;
  s = (qi & 2) ? -ss : ss;
  c = ((qi + 1) & 2) ? -cc : cc;
}

// Sum over the 64 lanes of a wave with DPP moves; the total ends in lane 63
// (quad swaps, row shifts by 4 and 8 within each 16-lane row, then the row
// broadcasts 15 and 31 that chain the four rows).
template <int CTRL, int ROW_MASK = 0xf, int BANK_MASK = 0xf>
__device__ __forceinline__ float dpp_f(float v) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, ROW_MASK,
                                         BANK_MASK, false));
}
template <int CTRL, int ROW_MASK = 0xf, int BANK_MASK = 0xf>
__device__ __forceinline__ double dpp_d(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), CTRL, ROW_MASK,
                                             BANK_MASK, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROW_MASK, BANK_MASK,
                                             false);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ float wave_sum63(float v) {
  v += dpp_f<0xb1>(v);
  v += dpp_f<0x4e>(v);
  v += dpp_f<0x114, 0xf, 0xe>(v);
  v += dpp_f<0x118, 0xf, 0xc>(v);
  v += dpp_f<0x142, 0xa>(v);
  v += dpp_f<0x143, 0xc>(v);
  return v;
}
__device__ __forceinline__ double wave_sum63(double v) {
  v += dpp_d<0xb1>(v);
  v += dpp_d<0x4e>(v);
  v += dpp_d<0x114, 0xf, 0xe>(v);
  v += dpp_d<0x118, 0xf, 0xc>(v);
  v += dpp_d<0x142, 0xa>(v);
  v += dpp_d<0x143, 0xc>(v);
  return v;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Sum of v over a block of LH_VB threads; result valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double* lds /* >= 4 */) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) lds[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) s = (lds[0] + lds[1]) + (lds[2] + lds[3]);
  __syncthreads();
  return s;
}

// Point dyadic Green's tensor G(d) = pf [a I + b d d^T / r^2] with
// pf = k^2 e^{ikr} / (4 pi r), a = 1 + (ikr - 1)/(kr)^2,
// b = (3 - 3ikr - (kr)^2)/(kr)^2 (greens.rs:34-71), returned as
// G = A I + B d d^T. c0 = k^2 / (4 pi).
__device__ __forceinline__ void g_coeffs64(double r2, double k, double inv_k, double c0,
                                          double& Ar, double& Ai, double& Br, double& Bi) {
  const double inv_r = rsqrt(r2);
  const double r = r2 * inv_r;
  double s, c;
  sincos(k * r, &s, &c);
  const double u = inv_r * inv_k;  // 1 / (k r)
  const double u2 = u * u;
  const double pf = c0 * inv_r;
  const double pr = pf * c, pi = pf * s;
  const double ar = 1.0 - u2, ai = u;
  const double br = 3.0 * u2 - 1.0, bi = -3.0 * u;
  const double ir2 = inv_r * inv_r;
  Ar = pr * ar - pi * ai;
  Ai = pr * ai + pi * ar;
  Br = (pr * br - pi * bi) * ir2;
  Bi = (pr * bi + pi * br) * ir2;
}

// The six independent entries (xx yy zz xy xz yz) of the point tensor.
__device__ __forceinline__ void g6_point64(double dx, double dy, double dz, double k,
                                           double inv_k, double c0, double t[12]) {
  double Ar, Ai, Br, Bi;
  g_coeffs64(r2_exact(dx, dy, dz), k, inv_k, c0, Ar, Ai, Br, Bi);
  const double d[3] = {dx, dy, dz};
  t[0] = Ar + Br * dx * dx; t[1] = Ai + Bi * dx * dx;
  t[2] = Ar + Br * dy * dy; t[3] = Ai + Bi * dy * dy;
  t[4] = Ar + Br * dz * dz; t[5] = Ai + Bi * dz * dz;
  t[6] = Br * dx * dy;      t[7] = Bi * dx * dy;
  t[8] = Br * dx * dz;      t[9] = Bi * dx * dz;
  t[10] = Br * dy * dz;     t[11] = Bi * dy * dz;
  (void)d;
}

}  // namespace lh
