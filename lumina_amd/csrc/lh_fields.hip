// Fields of the solved dipoles (fields.rs): the total field at observation
// points (near-field maps) and the far-field radiation pattern.
//
//  * field_at_point (fields.rs:18-48): E(r) = E_inc(r) + sum_i G(r, r_i) p_i,
//    a dipole closer than |d|^2 < 1e-20 to the point left out; the point
//    tensor of greens.rs:34-71 (no FCD). k_field_part sums one 1024-dipole
//    chunk for a block of 256 points (one point per lane, dipole records read
//    through the scalar cache), k_field_reduce adds the chunks in order and
//    the incident field and forms |E|^2 (fields.rs:87).
//  * compute_far_field (fields.rs:156-221): F(rhat) = sum_i p_i
//    e^{-ik rhat.r_i} per direction, then |(I - rhat rhat) F|^2; the same
//    chunked layout (k_ff_part, k_ff_reduce).
#include "lh_device.h"

namespace lh {

// Source records of solve `sd` from its positions and the moments in slot
// `slot`.
__global__ __launch_bounds__(256) void k_field_pack(int n, const double* __restrict__ px,
                                                    const double* __restrict__ py,
                                                    const double* __restrict__ pz,
                                                    const double2* __restrict__ mom,
                                                    FieldSrc* __restrict__ out) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= n) return;
  FieldSrc r;
  r.x = px[j];
  r.y = py[j];
  r.z = pz[j];
  r.pad = 0.0;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double2 v = mom[3LL * j + c];
    r.v[2 * c] = v.x;
    r.v[2 * c + 1] = v.y;
  }
  out[j] = r;
}

__global__ __launch_bounds__(256) void k_field_part(int m, const double* __restrict__ pts,
                                                    const FieldSrc* __restrict__ src, int n,
                                                    int chunk, double k, double inv_k, double c0,
                                                    double2* __restrict__ part) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  const bool valid = q < m;
  const int qq = valid ? q : 0;
  const double ox = pts[3 * qq], oy = pts[3 * qq + 1], oz = pts[3 * qq + 2];
  const int j0 = blockIdx.y * chunk, j1 = min(n, j0 + chunk);
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int j = j0; j < j1; ++j) {
    const FieldSrc& r = src[j];
    const double dx = ox - r.x, dy = oy - r.y, dz = oz - r.z;
    const double r2 = dx * dx + dy * dy + dz * dz;
    double Ar, Ai, Br, Bi;
    g_coeffs64(r2, k, inv_k, c0, Ar, Ai, Br, Bi, r2 < 1e-20);
    const double dr = fma(dz, r.v[4], fma(dy, r.v[2], dx * r.v[0]));
    const double di = fma(dz, r.v[5], fma(dy, r.v[3], dx * r.v[1]));
    const double wr = fma(-Bi, di, Br * dr), wi = fma(Bi, dr, Br * di);
    const double d[3] = {dx, dy, dz};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      acc[2 * c] = fma(d[c], wr, fma(-Ai, r.v[2 * c + 1], fma(Ar, r.v[2 * c], acc[2 * c])));
      acc[2 * c + 1] = fma(d[c], wi, fma(Ai, r.v[2 * c], fma(Ar, r.v[2 * c + 1], acc[2 * c + 1])));
    }
  }
  if (valid) {
    double2* o = part + ((long long)blockIdx.y * m + q) * 3;
#pragma unroll
    for (int c = 0; c < 3; ++c) o[c] = make_double2(acc[2 * c], acc[2 * c + 1]);
  }
}

// The incident polarisation (re, im per component), passed by value.
struct Pol6 {
  double v[6];
};

// field and intensity may be null (not written).
__global__ __launch_bounds__(256) void k_field_reduce(int m, int nchunk,
                                                      const double* __restrict__ pts,
                                                      const double2* __restrict__ part, double k,
                                                      double dx, double dy, double dz,
                                                      double amp, Pol6 pol,
                                                      double2* __restrict__ field,
                                                      double* __restrict__ intensity) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= m) return;
  double e[6];
  // types.rs:60-74: E0 e e^{ik dir.r}
  const double kdotr = k * (dx * pts[3 * q] + dy * pts[3 * q + 1] + dz * pts[3 * q + 2]);
  double sn, cs;
  sincos_f64(kdotr, sn, cs);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double er = amp * pol.v[2 * c], ei = amp * pol.v[2 * c + 1];
    e[2 * c] = er * cs - ei * sn;
    e[2 * c + 1] = er * sn + ei * cs;
  }
  for (int ch = 0; ch < nchunk; ++ch) {
    const double2* p = part + ((long long)ch * m + q) * 3;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      e[2 * c] += p[c].x;
      e[2 * c + 1] += p[c].y;
    }
  }
  double in = 0.0;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (field) field[3LL * q + c] = make_double2(e[2 * c], e[2 * c + 1]);
    in += e[2 * c] * e[2 * c] + e[2 * c + 1] * e[2 * c + 1];
  }
  if (intensity) intensity[q] = in;
}

__global__ __launch_bounds__(256) void k_ff_part(int m, const double* __restrict__ rhat,
                                                 const FieldSrc* __restrict__ src, int n, int chunk,
                                                 double k, double2* __restrict__ part) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  const bool valid = q < m;
  const int qq = valid ? q : 0;
  const double rx = rhat[3 * qq], ry = rhat[3 * qq + 1], rz = rhat[3 * qq + 2];
  const int j0 = blockIdx.y * chunk, j1 = min(n, j0 + chunk);
  double f[6] = {0, 0, 0, 0, 0, 0};
  for (int j = j0; j < j1; ++j) {
    const FieldSrc& r = src[j];
    const double kdotr = k * (rx * r.x + ry * r.y + rz * r.z);
    double sn, cs;
    sincos_f64(kdotr, sn, cs);
    // phase e^{-i k rhat.r} = (cos, -sin)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      f[2 * c] = fma(r.v[2 * c], cs, fma(r.v[2 * c + 1], sn, f[2 * c]));
      f[2 * c + 1] = fma(r.v[2 * c + 1], cs, fma(-r.v[2 * c], sn, f[2 * c + 1]));
    }
  }
  if (valid) {
    double2* o = part + ((long long)blockIdx.y * m + q) * 3;
#pragma unroll
    for (int c = 0; c < 3; ++c) o[c] = make_double2(f[2 * c], f[2 * c + 1]);
  }
}

__global__ __launch_bounds__(256) void k_ff_reduce(int m, int nchunk,
                                                   const double* __restrict__ rhat,
                                                   const double2* __restrict__ part,
                                                   double* __restrict__ intensity) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= m) return;
  double f[6] = {0, 0, 0, 0, 0, 0};
  for (int ch = 0; ch < nchunk; ++ch) {
    const double2* p = part + ((long long)ch * m + q) * 3;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      f[2 * c] += p[c].x;
      f[2 * c + 1] += p[c].y;
    }
  }
  const double rh[3] = {rhat[3 * q], rhat[3 * q + 1], rhat[3 * q + 2]};
  const double dr = rh[0] * f[0] + rh[1] * f[2] + rh[2] * f[4];
  const double di = rh[0] * f[1] + rh[1] * f[3] + rh[2] * f[5];
  double in = 0.0;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double tr = f[2 * c] - dr * rh[c], ti = f[2 * c + 1] - di * rh[c];
    in += tr * tr + ti * ti;
  }
  intensity[q] = in;
}

// ---------------------------------------------------------------------------
// Launchers (host).
// ---------------------------------------------------------------------------

void launch_field_pack(hipStream_t st, int n, const double* px, const double* py,
                       const double* pz, const double2* mom, FieldSrc* out) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_field_pack, dim3((n + 255) / 256), dim3(256), 0, st, n, px, py, pz, mom,
                     out);
}

void launch_field_part(hipStream_t st, int m, const double* pts, const FieldSrc* src, int n,
                       int chunk, double k, double2* part) {
  if (m <= 0 || n <= 0) return;
  const dim3 grid((m + 255) / 256, (n + chunk - 1) / chunk);
  hipLaunchKernelGGL(k_field_part, grid, dim3(256), 0, st, m, pts, src, n, chunk, k, 1.0 / k,
                     k * k / (4.0 * M_PI), part);
}

void launch_field_reduce(hipStream_t st, int m, int nchunk, const double* pts,
                         const double2* part, double k, const double dir[3], double amp,
                         const double pol[6], double2* field, double* intensity) {
  if (m <= 0) return;
  Pol6 p6;
  for (int q = 0; q < 6; ++q) p6.v[q] = pol[q];
  hipLaunchKernelGGL(k_field_reduce, dim3((m + 255) / 256), dim3(256), 0, st, m, nchunk, pts, part,
                     k, dir[0], dir[1], dir[2], amp, p6, field, intensity);
}

void launch_ff_part(hipStream_t st, int m, const double* rhat, const FieldSrc* src, int n, int chunk,
                    double k, double2* part) {
  if (m <= 0 || n <= 0) return;
  const dim3 grid((m + 255) / 256, (n + chunk - 1) / chunk);
  hipLaunchKernelGGL(k_ff_part, grid, dim3(256), 0, st, m, rhat, src, n, chunk, k, part);
}

void launch_ff_reduce(hipStream_t st, int m, int nchunk, const double* rhat, const double2* part,
                      double* intensity) {
  if (m <= 0) return;
  hipLaunchKernelGGL(k_ff_reduce, dim3((m + 255) / 256), dim3(256), 0, st, m, nchunk, rhat, part,
                     intensity);
}

}  // namespace lh
