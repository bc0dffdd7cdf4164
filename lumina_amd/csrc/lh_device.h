// Device helpers shared by the pair-sum and Krylov kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include "lh_types.h"
#include "lh_math.h"


namespace lh {

// Whether solve s takes part in a launch (the kWhich codes of lh_types.h).
__device__ __forceinline__ bool takes_part(const GState* st, int s, int which) {
  if (which == kWhichAll) return true;
  const GState& g = st[s];
  if (which == kWhichInner) return g.inner != 0;
  if (which == kWhichStarted) return g.started != 0;
  return g.nflag != 0;
}

// One far-field partial (re, im) at complex index e of the partial slices:
// FP64, or rounded to FP32 when p32 (the slices of an FP32 matvec, which
// then occupy the first half of the buffer as float2 and are read back so).
__device__ __forceinline__ void put_part(double2* part, long long e, double re, double im,
                                         int p32) {
  if (p32) reinterpret_cast<float2*>(part)[e] = make_float2((float)re, (float)im);
  else part[e] = make_double2(re, im);
}
__device__ __forceinline__ double2 get_part(const double2* part, long long e, int p32) {
  if (p32) {
    const float2 v = reinterpret_cast<const float2*>(part)[e];
    return make_double2((double)v.x, (double)v.y);
  }
  return part[e];
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

// 1/sqrt(x) for normal x > 0: the hardware estimate refined by one
// second-order Newton-Raphson step, y (1 + e/2 + 3e^2/8) with e = 1 - x y^2,
// which takes a 2^-26-accurate estimate to full double precision. Zero,
// infinite and denormal arguments are not handled (pair distances are
// normal numbers; callers select away r = 0).
__device__ __forceinline__ double rsqrt_f64(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double e = fma(-x * y, y, 1.0);
  return fma(y * e, fma(e, 0.375, 0.5), y);
}

// Point dyadic Green's tensor G(d) = pf [a I + b d d^T / r^2] with
// pf = k^2 e^{ikr} / (4 pi r), a = 1 + (ikr - 1)/(kr)^2,
// b = (3 - 3ikr - (kr)^2)/(kr)^2 (greens.rs:34-71), returned as
// G = A I + B d d^T. c0 = k^2 / (4 pi). skip = true gives A = B = 0 (1/r
// is taken as 0, r2 must be finite).
__device__ __forceinline__ void g_coeffs64(double r2, double k, double inv_k, double c0,
                                          double& Ar, double& Ai, double& Br, double& Bi,
                                          bool skip = false) {
  const double inv_r = skip ? 0.0 : rsqrt_f64(r2);
  const double r = r2 * inv_r;
  double s, c;
  sincos_f64(k * r, s, c);
  const double u = inv_r * inv_k;  // 1 / (k r)
  const double pf = c0 * inv_r;
  const double pr = pf * c, pi = pf * s;  // g = pf e^{ikr}
  const double ar = fma(-u, u, 1.0);      // a = (1 - u^2) + iu
  Ar = pr * ar - pi * u;
  Ai = pr * u + pi * ar;
  // b = (3u^2 - 1) - 3iu = 2 - 3a, so B = g b / r^2 = (g - 1.5 A)(2 / r^2)
  const double h = 2.0 * (inv_r * inv_r);
  Br = fma(-1.5, Ar, pr) * h;
  Bi = fma(-1.5, Ai, pi) * h;
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

// FCD cell average of the point tensor (greens.rs:120-164): nodes
// d - (cell/2) n with n in {-sqrt(3/5), 0, sqrt(3/5)}^3, weights
// w_x w_y w_z from {5/9, 8/9, 5/9}, normalised by the total weight of the
// nodes kept (a node closer than 1e-10 is left out).
__device__ __forceinline__ void fcd_average64(double dx, double dy, double dz, double cell,
                                              double k, double inv_k, double c0, double t[12]) {
  const double half_d = cell / 2.0;
  const double nodes[3] = {-sqrt(3.0 / 5.0), 0.0, sqrt(3.0 / 5.0)};
  const double wts[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
  double acc[12];
#pragma unroll
  for (int q = 0; q < 12; ++q) acc[q] = 0.0;
  double tw = 0.0;
  for (int ix = 0; ix < 3; ++ix)
    for (int iy = 0; iy < 3; ++iy)
      for (int iz = 0; iz < 3; ++iz) {
        const double ex = dx - half_d * nodes[ix];
        const double ey = dy - half_d * nodes[iy];
        const double ez = dz - half_d * nodes[iz];
        if (r2_exact(ex, ey, ez) < 1e-20) continue;
        const double w = wts[ix] * wts[iy] * wts[iz];
        double g[12];
        g6_point64(ex, ey, ez, k, inv_k, c0, g);
#pragma unroll
        for (int q = 0; q < 12; ++q) acc[q] += w * g[q];
        tw += w;
      }
  const double inv_w = (tw > 1e-30) ? 1.0 / tw : 1.0;
#pragma unroll
  for (int q = 0; q < 12; ++q) t[q] = acc[q] * inv_w;
}

}  // namespace lh
