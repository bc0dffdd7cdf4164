// On-the-fly CDA matvec y = A x for a batch of solves (gfx950).
//
//   y_i = alpha_i^-1 x_i - sum_{j != i} G(r_i, r_j) x_j
//                        - sum_j (m o G(r_i, mirror r_j)) x_j        (substrate)
//
// (assembly.rs:209-318 matvec_on_the_fly). The pair sum is split in two:
//
//  * far field (k_far64 / k_far32): every pair outside the exclusion radius
//    (and j != i) evaluates the point Green's tensor on the fly
//    (greens.rs:34-71) in the y += A x + B d (d . x) form -- no tensor is
//    materialised. One target dipole per lane; the source records stream
//    through the scalar cache (uniform loads into SGPRs), so a source is
//    fetched once per wave and consumed by 64 targets without LDS traffic.
//    Work items are (solve, 256-target tile, source chunk); each writes its
//    own partial sum, so no atomics are needed and results are
//    reproducible.
//  * near field (k_near_tensors + k_reduce): pairs inside the exclusion
//    radius use tensors precomputed once per solve in FP64: the FCD 27-node
//    cell average (greens.rs:94-167) where dist <= 2 cell_size, the point
//    tensor otherwise. k_reduce adds alpha^-1 x, the far partials and the
//    near tensors and writes y (or the residual rhs - y and its norm).
//
// The image term (assembly.rs:272-295, Fresnel branch) is a second point
// tensor per pair at the mirrored source, applied to (-D, -D, +D) o x_j,
// for every pair including j == i, skipped when |d|^2 < 1e-20.
#include "lh_device.h"

namespace lh {

// ---------------------------------------------------------------------------
// Near-field tensors (once per solve).
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_near_tensors(
    const SolveDev* __restrict__ solves, const int* __restrict__ item_s,
    const long long* __restrict__ item_e, int nitems, const double* __restrict__ px,
    const double* __restrict__ py, const double* __restrict__ pz,
    const int* __restrict__ near_row, const int* __restrict__ near_col,
    const unsigned char* __restrict__ near_fcd, double* __restrict__ near_t, long long t_stride, const long long* __restrict__ e_end) {
  const int b = blockIdx.x;
  if (b >= nitems) return;
  const int s = item_s[b];
  const SolveDev S = solves[s];
  const long long e = item_e[b] + threadIdx.x;
  if (e >= e_end[b]) return;
  const int i = near_row[e];
  const int j = near_col[e];
  const double xi = px[S.pos_off + i], yi = py[S.pos_off + i], zi = pz[S.pos_off + i];
  const double xj = px[S.pos_off + j], yj = py[S.pos_off + j], zj = pz[S.pos_off + j];
  double t[12];
  if (near_fcd[e]) {
    // greens.rs:120-164: nodes r_j + (d/2) n, n in {-sqrt(3/5), 0, sqrt(3/5)}^3,
    // weights w_x w_y w_z from {5/9, 8/9, 5/9}, normalised by the total weight.
    const double cs = S.cell;
    const double half_d = cs / 2.0;
    const double nodes[3] = {-sqrt(3.0 / 5.0), 0.0, sqrt(3.0 / 5.0)};
    const double wts[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
    double acc[12];
#pragma unroll
    for (int q = 0; q < 12; ++q) acc[q] = 0.0;
    double tw = 0.0;
    for (int ix = 0; ix < 3; ++ix)
      for (int iy = 0; iy < 3; ++iy)
        for (int iz = 0; iz < 3; ++iz) {
          const double dx = xi - (xj + half_d * nodes[ix]);
          const double dy = yi - (yj + half_d * nodes[iy]);
          const double dz = zi - (zj + half_d * nodes[iz]);
          if (r2_exact(dx, dy, dz) < 1e-20) continue;
          const double w = wts[ix] * wts[iy] * wts[iz];
          double g[12];
          g6_point64(dx, dy, dz, S.k, S.inv_k, S.c0, g);
#pragma unroll
          for (int q = 0; q < 12; ++q) acc[q] += w * g[q];
          tw += w;
        }
    const double inv_w = (tw > 1e-30) ? 1.0 / tw : 1.0;
#pragma unroll
    for (int q = 0; q < 12; ++q) t[q] = acc[q] * inv_w;
  } else {
    g6_point64(xi - xj, yi - yj, zi - zj, S.k, S.inv_k, S.c0, t);
  }
  const long long te = S.t_off + (e - S.e0);
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    near_t[(2 * q) * t_stride + te] = t[2 * q];
    near_t[(2 * q + 1) * t_stride + te] = t[2 * q + 1];
  }
}

// ---------------------------------------------------------------------------
// Per-solve participation flags (which phase of the GMRES cycle uses a
// kernel): 0 = in the Arnoldi loop, 1 = in a restart cycle, 2 = normalise
// w -> v_{j+1}, 3 = always (single matvec).
// ---------------------------------------------------------------------------

__device__ __forceinline__ bool takes_part(const GState* st, int s, int which) {
  if (which == 3) return true;
  const GState& g = st[s];
  if (which == 0) return g.inner != 0;
  if (which == 1) return g.started != 0;
  return g.nflag != 0;
}

// ---------------------------------------------------------------------------
// Source records for one matvec: position + input vector (+ image-scaled
// input). x is read from the Krylov slot `slot` of each solve.
// ---------------------------------------------------------------------------

template <typename SRC>
__global__ __launch_bounds__(LH_VB) void k_prep(const VecItem* __restrict__ items,
                                               const SolveDev* __restrict__ solves,
                                               const GState* __restrict__ st, int which,
                                               const double* __restrict__ px,
                                               const double* __restrict__ py,
                                               const double* __restrict__ pz,
                                               const double2* __restrict__ vec, int slot,
                                               SRC* __restrict__ src) {
  const VecItem it = items[blockIdx.x];
  if (!takes_part(st, it.s, which)) return;
  const SolveDev S = solves[it.s];
  const int j = it.t0 + threadIdx.x;
  if (j >= S.n) return;
  const long long base = S.vec_off + (long long)slot * 3 * S.n + 3LL * j;
  double v[6];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double2 a = vec[base + c];
    v[2 * c] = a.x;
    v[2 * c + 1] = a.y;
  }
  // (-D, -D, +D) o x_j (assembly.rs:288-292: m_b scales source component b)
  double w[6];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double sg = (c == 2) ? 1.0 : -1.0;
    const double mr = sg * S.dre, mi = sg * S.dim;
    w[2 * c] = mr * v[2 * c] - mi * v[2 * c + 1];
    w[2 * c + 1] = mr * v[2 * c + 1] + mi * v[2 * c];
  }
  SRC r;
  r.x = px[S.pos_off + j];
  r.y = py[S.pos_off + j];
  r.z = pz[S.pos_off + j];
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    r.v[q] = static_cast<typename SRC::value_type>(v[q]);
    r.w[q] = static_cast<typename SRC::value_type>(w[q]);
  }
  src[S.src_off + j] = r;
}

// ---------------------------------------------------------------------------
// Far field, FP64.
// ---------------------------------------------------------------------------

__device__ __forceinline__ void far_pair64(double dx, double dy, double dz, double r2,
                                           double scale, const double* __restrict__ v,
                                           double k, double inv_k, double c0,
                                           double acc[6]) {
  double Ar, Ai, Br, Bi;
  g_coeffs64(r2, k, inv_k, c0, Ar, Ai, Br, Bi);
  Ar *= scale; Ai *= scale; Br *= scale; Bi *= scale;
  // d . x_j
  const double dr = dx * v[0] + dy * v[2] + dz * v[4];
  const double di = dx * v[1] + dy * v[3] + dz * v[5];
  const double wr = Br * dr - Bi * di;
  const double wi = Br * di + Bi * dr;
  acc[0] += Ar * v[0] - Ai * v[1] + dx * wr;
  acc[1] += Ar * v[1] + Ai * v[0] + dx * wi;
  acc[2] += Ar * v[2] - Ai * v[3] + dy * wr;
  acc[3] += Ar * v[3] + Ai * v[2] + dy * wi;
  acc[4] += Ar * v[4] - Ai * v[5] + dz * wr;
  acc[5] += Ar * v[5] + Ai * v[4] + dz * wi;
}

template <bool SUB>
__global__ __launch_bounds__(256) void k_far64(const FarItem* __restrict__ items,
                                               const SolveDev* __restrict__ solves,
                                               const GState* __restrict__ st, int which,
                                               const double* __restrict__ px,
                                               const double* __restrict__ py,
                                               const double* __restrict__ pz,
                                               const Src64* __restrict__ src,
                                               double2* __restrict__ part) {
  const FarItem it = items[blockIdx.x];
  if (!takes_part(st, it.s, which)) return;
  const SolveDev S = solves[it.s];
  if (SUB != (S.has_sub != 0)) return;
  const int i = it.t0 + threadIdx.x;
  const bool valid = i < S.n;
  const int ii = valid ? i : 0;
  const double xi = px[S.pos_off + ii], yi = py[S.pos_off + ii], zi = pz[S.pos_off + ii];
  const double k = S.k, inv_k = S.inv_k, c0 = S.c0, rskip2 = S.rskip2, z2 = S.z2;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  const Src64* __restrict__ sp = src + S.src_off;
  for (int j = it.c0; j < it.c1; ++j) {
    const Src64& r = sp[j];
    const double dx = xi - r.x, dy = yi - r.y, dz = zi - r.z;
    const double d2 = r2_exact(dx, dy, dz);
    const bool skip = d2 <= rskip2;
    far_pair64(dx, dy, dz, skip ? 1.0 : d2, skip ? 0.0 : 1.0, r.v, k, inv_k, c0, acc);
    if (SUB) {
      const double ez = zi - (z2 - r.z);
      const double e2 = r2_exact(dx, dy, ez);
      const bool eskip = e2 < 1e-20;
      far_pair64(dx, dy, ez, eskip ? 1.0 : e2, eskip ? 0.0 : 1.0, r.w, k, inv_k, c0, acc);
    }
  }
  if (valid) {
    double2* o = part + S.part_off + (long long)it.chunk * 3 * S.n + 3LL * i;
    o[0] = make_double2(acc[0], acc[1]);
    o[1] = make_double2(acc[2], acc[3]);
    o[2] = make_double2(acc[4], acc[5]);
  }
}

// ---------------------------------------------------------------------------
// Far field, FP32 evaluation with compensated accumulation. The exclusion
// test and the displacement are FP64 (the same test as the FP64 kernel and
// the near list); G and the products are FP32. Terms are summed in FP32
// over runs of LH_RUN sources, and the run sums are folded into a float-float
// (hi, lo) accumulator by an error-free TwoSum.
// ---------------------------------------------------------------------------

#define LH_RUN 32

__device__ __forceinline__ void two_sum(float& hi, float& lo, float v) {
  const float s = hi + v;
  const float bb = s - hi;
  const float err = (hi - (s - bb)) + (v - bb);
  hi = s;
  lo += err;
}

__device__ __forceinline__ void far_pair32(float dx, float dy, float dz, float r2, float scale,
                                           const float* __restrict__ v, float k, float inv_k,
                                           float c0, float acc[6]) {
  const float inv_r = rsqrtf(r2);
  const float r = r2 * inv_r;
  float s, c;
  sincosf(k * r, &s, &c);
  const float u = inv_r * inv_k;
  const float u2 = u * u;
  const float pf = c0 * inv_r * scale;
  const float pr = pf * c, pi = pf * s;
  const float ar = 1.0f - u2, ai = u;
  const float br = 3.0f * u2 - 1.0f, bi = -3.0f * u;
  const float ir2 = inv_r * inv_r;
  const float Ar = pr * ar - pi * ai;
  const float Ai = pr * ai + pi * ar;
  const float Br = (pr * br - pi * bi) * ir2;
  const float Bi = (pr * bi + pi * br) * ir2;
  const float dr = dx * v[0] + dy * v[2] + dz * v[4];
  const float di = dx * v[1] + dy * v[3] + dz * v[5];
  const float wr = Br * dr - Bi * di;
  const float wi = Br * di + Bi * dr;
  acc[0] += Ar * v[0] - Ai * v[1] + dx * wr;
  acc[1] += Ar * v[1] + Ai * v[0] + dx * wi;
  acc[2] += Ar * v[2] - Ai * v[3] + dy * wr;
  acc[3] += Ar * v[3] + Ai * v[2] + dy * wi;
  acc[4] += Ar * v[4] - Ai * v[5] + dz * wr;
  acc[5] += Ar * v[5] + Ai * v[4] + dz * wi;
}

template <bool SUB>
__global__ __launch_bounds__(256) void k_far32(const FarItem* __restrict__ items,
                                               const SolveDev* __restrict__ solves,
                                               const GState* __restrict__ st, int which,
                                               const double* __restrict__ px,
                                               const double* __restrict__ py,
                                               const double* __restrict__ pz,
                                               const Src32* __restrict__ src,
                                               double2* __restrict__ part) {
  const FarItem it = items[blockIdx.x];
  if (!takes_part(st, it.s, which)) return;
  const SolveDev S = solves[it.s];
  if (SUB != (S.has_sub != 0)) return;
  const int i = it.t0 + threadIdx.x;
  const bool valid = i < S.n;
  const int ii = valid ? i : 0;
  const double xi = px[S.pos_off + ii], yi = py[S.pos_off + ii], zi = pz[S.pos_off + ii];
  const float k = (float)S.k, inv_k = (float)S.inv_k, c0 = (float)S.c0;
  const double rskip2 = S.rskip2, z2 = S.z2;
  float hi[6] = {0, 0, 0, 0, 0, 0}, lo[6] = {0, 0, 0, 0, 0, 0};
  const Src32* __restrict__ sp = src + S.src_off;
  for (int j0 = it.c0; j0 < it.c1; j0 += LH_RUN) {
    const int j1 = min(j0 + LH_RUN, it.c1);
    float acc[6] = {0, 0, 0, 0, 0, 0};
    for (int j = j0; j < j1; ++j) {
      const Src32& r = sp[j];
      const double dx = xi - r.x, dy = yi - r.y, dz = zi - r.z;
      const double d2 = r2_exact(dx, dy, dz);
      const bool skip = d2 <= rskip2;
      far_pair32((float)dx, (float)dy, (float)dz, skip ? 1.0f : (float)d2, skip ? 0.0f : 1.0f,
                 r.v, k, inv_k, c0, acc);
      if (SUB) {
        const double ez = zi - (z2 - r.z);
        const double e2 = r2_exact(dx, dy, ez);
        const bool eskip = e2 < 1e-20;
        far_pair32((float)dx, (float)dy, (float)ez, eskip ? 1.0f : (float)e2,
                   eskip ? 0.0f : 1.0f, r.w, k, inv_k, c0, acc);
      }
    }
#pragma unroll
    for (int q = 0; q < 6; ++q) two_sum(hi[q], lo[q], acc[q]);
  }
  if (valid) {
    double2* o = part + S.part_off + (long long)it.chunk * 3 * S.n + 3LL * i;
#pragma unroll
    for (int c = 0; c < 3; ++c)
      o[c] = make_double2((double)hi[2 * c] + (double)lo[2 * c],
                          (double)hi[2 * c + 1] + (double)lo[2 * c + 1]);
  }
}

// ---------------------------------------------------------------------------
// Reduction: y_i = alpha_i^-1 x_i - sum_chunks part - sum_near T_ij x_j.
// mode 0 writes y to slot `out_slot`; mode 1 writes r = rhs - y to the
// residual slot and a per-item partial of |r|^2 to norm_part.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(LH_VB) void k_reduce(
    const VecItem* __restrict__ items, const SolveDev* __restrict__ solves,
    const GState* __restrict__ st, int which, const long long* __restrict__ near_rp,
    const int* __restrict__ near_col, const double* __restrict__ near_t, long long t_stride,
    const double2* __restrict__ ainv, double2* __restrict__ vec, int in_slot, int out_slot,
    const double2* __restrict__ part, int mode, double* __restrict__ norm_part) {
  __shared__ double lds[4];
  const VecItem it = items[blockIdx.x];
  if (!takes_part(st, it.s, which)) return;
  const SolveDev S = solves[it.s];
  const int i = it.t0 + threadIdx.x;
  double nrm = 0.0;
  if (i < S.n) {
    const long long n3 = 3LL * S.n;
    const double2* xin = vec + S.vec_off + (long long)in_slot * n3;
    double xr[3], xim[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double2 a = xin[3LL * i + c];
      xr[c] = a.x;
      xim[c] = a.y;
    }
    const double2* A = ainv + S.ainv_off + (S.ainv_uniform ? 0 : 9LL * i);
    double yr[3], yi[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      double sr = 0.0, si = 0.0;
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        const double2 m = A[3 * a + b];
        sr += m.x * xr[b] - m.y * xim[b];
        si += m.x * xim[b] + m.y * xr[b];
      }
      yr[a] = sr;
      yi[a] = si;
    }
    for (int ch = 0; ch < S.nsc; ++ch) {
      const double2* p = part + S.part_off + (long long)ch * n3 + 3LL * i;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const double2 q = p[c];
        yr[c] -= q.x;
        yi[c] -= q.y;
      }
    }
    const long long e_lo = near_rp[S.rp_off + i], e_hi = near_rp[S.rp_off + i + 1];
    for (long long e = e_lo; e < e_hi; ++e) {
      const int j = near_col[e];
      const long long te = S.t_off + (e - S.e0);
      double t[12];
#pragma unroll
      for (int q = 0; q < 12; ++q) t[q] = near_t[q * t_stride + te];
      double vr[3], vi[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const double2 a = xin[3LL * j + c];
        vr[c] = a.x;
        vi[c] = a.y;
      }
      // symmetric tensor: (xx yy zz xy xz yz) at t[0..11] as (re, im)
      const int idx[3][3] = {{0, 3, 4}, {3, 1, 5}, {4, 5, 2}};
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        double sr = 0.0, si = 0.0;
#pragma unroll
        for (int b = 0; b < 3; ++b) {
          const double gr = t[2 * idx[a][b]], gi = t[2 * idx[a][b] + 1];
          sr += gr * vr[b] - gi * vi[b];
          si += gr * vi[b] + gi * vr[b];
        }
        yr[a] -= sr;
        yi[a] -= si;
      }
    }
    if (mode == 0) {
      double2* o = vec + S.vec_off + (long long)out_slot * n3 + 3LL * i;
#pragma unroll
      for (int c = 0; c < 3; ++c) o[c] = make_double2(yr[c], yi[c]);
    } else {
      const double2* rhs = vec + S.vec_off + (long long)LH_SLOT_RHS * n3 + 3LL * i;
      double2* o = vec + S.vec_off + (long long)LH_SLOT_R * n3 + 3LL * i;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const double2 b = rhs[c];
        const double rr = b.x - yr[c], ri = b.y - yi[c];
        o[c] = make_double2(rr, ri);
        nrm += rr * rr + ri * ri;
      }
    }
  }
  if (mode != 0) {
    const double tot = block_sum(nrm, lds);
    if (threadIdx.x == 0) norm_part[S.vi_off + it.b] = tot;
  }
}

// ---------------------------------------------------------------------------
// Launchers (host).
// ---------------------------------------------------------------------------

void launch_near_tensors(hipStream_t st, int nitems, const SolveDev* solves, const int* item_s,
                         const long long* item_e, const long long* e_end, const double* px,
                         const double* py, const double* pz, const int* near_row,
                         const int* near_col, const unsigned char* near_fcd, double* near_t,
                         long long t_stride) {
  if (nitems <= 0) return;
  hipLaunchKernelGGL(k_near_tensors, dim3(nitems), dim3(256), 0, st, solves, item_s, item_e,
                     nitems, px, py, pz, near_row, near_col, near_fcd, near_t, t_stride,
                     e_end);
}

void launch_prep64(hipStream_t st, int nitems, const VecItem* items, const SolveDev* solves,
                   const GState* gs, int which, const double* px, const double* py,
                   const double* pz, const double2* vec, int slot, Src64* src) {
  hipLaunchKernelGGL(k_prep<Src64>, dim3(nitems), dim3(LH_VB), 0, st, items, solves, gs, which,
                     px, py, pz, vec, slot, src);
}

void launch_prep32(hipStream_t st, int nitems, const VecItem* items, const SolveDev* solves,
                   const GState* gs, int which, const double* px, const double* py,
                   const double* pz, const double2* vec, int slot, Src32* src) {
  hipLaunchKernelGGL(k_prep<Src32>, dim3(nitems), dim3(LH_VB), 0, st, items, solves, gs, which,
                     px, py, pz, vec, slot, src);
}

void launch_far64(hipStream_t st, bool sub, int nitems, const FarItem* items,
                  const SolveDev* solves, const GState* gs, int which, const double* px,
                  const double* py, const double* pz, const Src64* src, double2* part) {
  if (nitems <= 0) return;
  if (sub)
    hipLaunchKernelGGL(k_far64<true>, dim3(nitems), dim3(256), 0, st, items, solves, gs, which,
                       px, py, pz, src, part);
  else
    hipLaunchKernelGGL(k_far64<false>, dim3(nitems), dim3(256), 0, st, items, solves, gs,
                       which, px, py, pz, src, part);
}

void launch_far32(hipStream_t st, bool sub, int nitems, const FarItem* items,
                  const SolveDev* solves, const GState* gs, int which, const double* px,
                  const double* py, const double* pz, const Src32* src, double2* part) {
  if (nitems <= 0) return;
  if (sub)
    hipLaunchKernelGGL(k_far32<true>, dim3(nitems), dim3(256), 0, st, items, solves, gs, which,
                       px, py, pz, src, part);
  else
    hipLaunchKernelGGL(k_far32<false>, dim3(nitems), dim3(256), 0, st, items, solves, gs,
                       which, px, py, pz, src, part);
}

void launch_reduce(hipStream_t st, int nitems, const VecItem* items, const SolveDev* solves,
                   const GState* gs, int which, const long long* near_rp, const int* near_col,
                   const double* near_t, long long t_stride, const double2* ainv, double2* vec,
                   int in_slot, int out_slot, const double2* part, int mode, double* norm_part) {
  hipLaunchKernelGGL(k_reduce, dim3(nitems), dim3(LH_VB), 0, st, items, solves, gs, which,
                     near_rp, near_col, near_t, t_stride, ainv, vec, in_slot, out_slot, part,
                     mode, norm_part);
}

}  // namespace lh
