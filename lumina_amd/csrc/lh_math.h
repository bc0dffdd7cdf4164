// Scalar math shared by the device kernels and the host (plain C++, no HIP
// headers, so the CPU tests compile it with g++ as well).
#pragma once
#include <cmath>

#if defined(__HIPCC__)
#define LH_HD __host__ __device__
#else
#define LH_HD
#endif

namespace lh {

// |d|^2 without fused multiply-add. Every kernel and the host near-list
// builder that decide whether a pair belongs to the near list evaluate
// exactly this expression, so they can never disagree.
LH_HD inline double r2_exact(double dx, double dy, double dz) {
#pragma clang fp contract(off)
  double a = dx * dx;
  double b = dy * dy;
  double c = dz * dz;
  return (a + b) + c;
}

// The same for FP32 positions: the exclusion test of every far-field
// kernel and of the host near-list builder. IEEE single-precision subtract,
// multiply and add give identical bits on the host and on the GPU.
LH_HD inline float r2f_exact(float dx, float dy, float dz) {
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
LH_HD inline void sincos_f32(float x, float& s, float& c) {
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

// sin and cos of an FP64 phase, |x| < 1e6: Cody-Waite reduction by pi/2
// with a three-part constant (P1, P2: pi/2 and its remainder rounded to 33
// significant bits, so q*P1 and q*P2 are exact for |q| < 2^20; P3: the rest
// rounded to double), then the Taylor series of sin (to r^17) and cos (to
// r^16) on [-pi/4, pi/4] (truncation below 3e-18), and the quadrant
// swap/sign. The coefficients are 1/n!.
LH_HD inline void sincos_f64(double x, double& s, double& c) {
  const double q = rint(x * 0.6366197723675814);
  double r = fma(-q, 1.5707963267341256, x);
  r = fma(-q, 6.077100506303966e-11, r);
  r = fma(-q, 2.0222662487959506e-21, r);
  const double r2 = r * r;
  double ps = fma(r2, 2.8114572543455206e-15, -7.647163731819816e-13);
  ps = fma(r2, ps, 1.6059043836821613e-10);
  ps = fma(r2, ps, -2.505210838544172e-08);
  ps = fma(r2, ps, 2.7557319223985893e-06);
  ps = fma(r2, ps, -1.984126984126984e-04);
  ps = fma(r2, ps, 8.333333333333333e-03);
  ps = fma(r2, ps, -1.6666666666666666e-01);
  const double sr = fma(r * r2, ps, r);
  double pc = fma(r2, 4.779477332387385e-14, -1.1470745597729725e-11);
  pc = fma(r2, pc, 2.08767569878681e-09);
  pc = fma(r2, pc, -2.755731922398589e-07);
  pc = fma(r2, pc, 2.48015873015873e-05);
  pc = fma(r2, pc, -1.388888888888889e-03);
  pc = fma(r2, pc, 4.1666666666666664e-02);
  pc = fma(r2, pc, -0.5);
  const double cr = fma(r2, pc, 1.0);
  const int qi = (int)q;   // |q| < 2^31 for |x| < 1e9
  const double ss = (qi & 1) ? cr : sr;
  const double cc = (qi & 1) ? sr : cr;
  s = (qi & 2) ? -ss : ss;
  c = ((qi + 1) & 2) ? -cc : cc;
}

}  // namespace lh
