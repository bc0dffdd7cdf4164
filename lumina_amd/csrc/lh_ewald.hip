// 2D-periodic (Ewald) pair sums for lattice solves.
//
// With a lattice, the reference replaces every direct tensor, the i == j
// self block included, by the Bloch-periodic Ewald tensor
// G_per(r_i - r_j) (assembly.rs:241-262), and the Fresnel image tensor by
// G_per at the mirrored displacement (assembly.rs:278-295); no FCD filter is
// applied. G_per = real-space sum + spectral sum (ewald.rs:91-106):
//   real space (ewald.rs:119-174): sum over lattice images R of
//     erfc(eta |s|) G_free(s) e^{i k_par . R}, s = r - R, the R = 0 image
//     left out when |r|^2 < 1e-24 and any |s|^2 < 1e-24 left out;
//   spectral (ewald.rs:186-273): (i / 2A) sum over Q = k_par + G of
//     e^{i Q . rho} e^{-q |z|} / q (k^2 delta_ab - kt_a kt_b),
//     q = csqrt+(|Q|^2 - k^2), kt = (Qx, Qy, i q sgn z), sgn 0 = +1.
// The lattice images, their Bloch phases and the r-independent parts of the
// spectral terms (Q, q, (i / 2A) / q) are prepared on the host per solve.
//
// Two evaluation forms:
//  * grid tables (lat = 1): when the dipoles sit on a regular grid, every
//    displacement r_i - r_j is a grid offset, and every image displacement
//    (x_i - x_j, y_i - y_j, z_i + z_j - 2 z0) is one too. G_per is tabulated
//    once per solve for every offset (k_ewald_table) and the matvec reads
//    the table (k_ewt_shift), FP64 or FP32. The table is symmetric in its
//    3x3 indices (both sums are), so 6 complex entries per offset.
//  * per pair (lat = 2): any other position set evaluates G_per for every
//    pair inside the matvec (k_ewp_row, FP64).
#include <type_traits>

#include "lh_device.h"

namespace lh {

// erfc by Abramowitz & Stegun 7.1.26 (ewald.rs:316-327), |error| < 1.5e-7:
// the reference's own approximation, kept so that the tensors agree.
__device__ __forceinline__ double erfc_as(double x) {
  const bool neg = x < 0.0;
  const double ax = neg ? -x : x;
  const double t = 1.0 / (1.0 + 0.3275911 * ax);
  const double poly =
      t * (0.254829592 +
           t * (-0.284496736 + t * (1.421413741 + t * (-1.453152027 + t * 1.061405429))));
  const double v = poly * exp(-ax * ax);
  return neg ? 2.0 - v : v;
}

// G_per(r) into t[12] (xx yy zz xy xz yz as re, im).
__device__ void ewald_tensor(const SolveDev& S, const EwReal* __restrict__ rl,
                             const EwRecip* __restrict__ rq, double rx, double ry, double rz,
                             double t[12]) {
#pragma unroll
  for (int q = 0; q < 12; ++q) t[q] = 0.0;
  const bool r_zero = r2_exact(rx, ry, rz) < 1e-24;
  for (int m = 0; m < S.ew_nreal; ++m) {
    const EwReal R = rl[m];
    if (r_zero && R.zero != 0.0) continue;
    const double sx = rx - R.x, sy = ry - R.y, sz = rz - R.z;
    const double s2 = r2_exact(sx, sy, sz);
    if (s2 < 1e-24) continue;
    const double damp = erfc_as(S.ew_eta * sqrt(s2));
    double Ar, Ai, Br, Bi;
    g_coeffs64(s2, S.k, S.inv_k, S.c0, Ar, Ai, Br, Bi);
    const double fr = R.br * damp, fi = R.bi * damp;
    const double g[12] = {Ar + Br * sx * sx, Ai + Bi * sx * sx, Ar + Br * sy * sy,
                          Ai + Bi * sy * sy, Ar + Br * sz * sz, Ai + Bi * sz * sz,
                          Br * sx * sy,      Bi * sx * sy,      Br * sx * sz,
                          Bi * sx * sz,      Br * sy * sz,      Bi * sy * sz};
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      t[2 * q] += fr * g[2 * q] - fi * g[2 * q + 1];
      t[2 * q + 1] += fr * g[2 * q + 1] + fi * g[2 * q];
    }
  }
  const double az = fabs(rz);
  const double sg = (rz >= 0.0) ? 1.0 : -1.0;
  const double k2 = S.k * S.k;
  for (int m = 0; m < S.ew_nrecip; ++m) {
    const EwRecip Q = rq[m];
    double bs, bc;
    sincos(Q.qx * rx + Q.qy * ry, &bs, &bc);
    double es, ec;
    sincos(Q.qi * az, &es, &ec);
    const double ea = exp(-Q.qr * az);
    // e^{-q |z|} = e^{-qr |z|} (cos(qi |z|) - i sin(qi |z|))
    const double er = ea * ec, ei = -ea * es;
    // coeff = c e^{i Q.rho} e^{-q|z|}
    const double pr = Q.cr * bc - Q.ci * bs, pi = Q.cr * bs + Q.ci * bc;
    const double cr = pr * er - pi * ei, ci = pr * ei + pi * er;
    // kt_z = i q sgn
    const double tzr = -Q.qi * sg, tzi = Q.qr * sg;
    const double m_re[6] = {k2 - Q.qx * Q.qx, k2 - Q.qy * Q.qy, k2 - (tzr * tzr - tzi * tzi),
                            -Q.qx * Q.qy,     -Q.qx * tzr,      -Q.qy * tzr};
    const double m_im[6] = {0.0, 0.0, -2.0 * tzr * tzi, 0.0, -Q.qx * tzi, -Q.qy * tzi};
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      t[2 * q] += cr * m_re[q] - ci * m_im[q];
      t[2 * q + 1] += cr * m_im[q] + ci * m_re[q];
    }
  }
}

// out += T v for a symmetric complex 3x3 T stored as (xx yy zz xy xz yz).
template <typename R>
__device__ __forceinline__ void apply_sym(const R* __restrict__ t, const R* v, R out[6]) {
  const R xxr = t[0], xxi = t[1], yyr = t[2], yyi = t[3], zzr = t[4], zzi = t[5];
  const R xyr = t[6], xyi = t[7], xzr = t[8], xzi = t[9], yzr = t[10], yzi = t[11];
  out[0] = fma(xxr, v[0], fma(-xxi, v[1], fma(xyr, v[2], fma(-xyi, v[3],
           fma(xzr, v[4], fma(-xzi, v[5], out[0]))))));
  out[1] = fma(xxr, v[1], fma(xxi, v[0], fma(xyr, v[3], fma(xyi, v[2],
           fma(xzr, v[5], fma(xzi, v[4], out[1]))))));
  out[2] = fma(xyr, v[0], fma(-xyi, v[1], fma(yyr, v[2], fma(-yyi, v[3],
           fma(yzr, v[4], fma(-yzi, v[5], out[2]))))));
  out[3] = fma(xyr, v[1], fma(xyi, v[0], fma(yyr, v[3], fma(yyi, v[2],
           fma(yzr, v[5], fma(yzi, v[4], out[3]))))));
  out[4] = fma(xzr, v[0], fma(-xzi, v[1], fma(yzr, v[2], fma(-yzi, v[3],
           fma(zzr, v[4], fma(-zzi, v[5], out[4]))))));
  out[5] = fma(xzr, v[1], fma(xzi, v[0], fma(yzr, v[3], fma(yzi, v[2],
           fma(zzr, v[5], fma(zzi, v[4], out[5]))))));
}

// ---------------------------------------------------------------------------
// Tables: one thread per grid offset; entries [0, ewn) direct, [ewn, 2 ewn)
// image (substrate solves only).
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_ewald_table(const SolveDev* __restrict__ solves,
                                                    const int* __restrict__ item_s,
                                                    const long long* __restrict__ item_e,
                                                    const double* __restrict__ ewl,
                                                    double* __restrict__ t64,
                                                    float* __restrict__ t32) {
  const SolveDev S = solves[item_s[blockIdx.x]];
  const long long e = item_e[blockIdx.x] + threadIdx.x;
  const long long total = S.ewn * (S.has_sub ? 2 : 1);
  if (e >= total) return;
  const int kind = e >= S.ewn ? 1 : 0;
  const long long rem = e - (long long)kind * S.ewn;
  const long long yz = (long long)S.ew_nyd * S.ew_nzd;
  const int ix = (int)(rem / yz);
  const int iy = (int)((rem / S.ew_nzd) % S.ew_nyd);
  const int iz = (int)(rem % S.ew_nzd);
  const int nxd = (int)(S.ewn / yz);
  const double rx = (double)(ix - (nxd - 1) / 2) * S.ew_h[0];
  const double ry = (double)(iy - (S.ew_nyd - 1) / 2) * S.ew_h[1];
  double rz;
  if (kind == 0) rz = (double)(iz - S.ew_nz1) * S.ew_h[2];
  else rz = S.ew_zimg0 + (double)iz * S.ew_h[2];
  double t[12];
  if (kind == 1 && r2_exact(rx, ry, rz) < 1e-20) {
    // assembly.rs:281: the image pair is skipped when |d|^2 < 1e-20
#pragma unroll
    for (int q = 0; q < 12; ++q) t[q] = 0.0;
  } else {
    const EwReal* rl = reinterpret_cast<const EwReal*>(ewl) + S.ewl_off;
    const EwRecip* rq = reinterpret_cast<const EwRecip*>(rl + S.ew_nreal);
    ewald_tensor(S, rl, rq, rx, ry, rz, t);
  }
  double* o64 = t64 + 12 * (S.ew_off + e);
  float* o32 = t32 + 12 * (S.ew_off + e);
#pragma unroll
  for (int q = 0; q < 12; ++q) {
    o64[q] = t[q];
    o32[q] = (float)t[q];
  }
}

// ---------------------------------------------------------------------------
// Matvec partials from the tables: targets per lane, sources uniform over the
// work item (the same source tiles and partial slices as the free-space far
// field).
// ---------------------------------------------------------------------------

// Table form for one wave of 64 consecutive targets, using the Toeplitz
// structure of the tables: when source j is the z-neighbour of source j - 1
// (pack_j = pack_{j-1} + 1), the direct entry lane l needs is the one lane
// l - 1 used for j - 1 if target l is the z-neighbour of target l - 1, and
// the image entry (index + 2 iz_j) is the one lane l + 1 used if target l + 1
// is the z-neighbour of target l. Those entries move between lanes by DPP
// shifts; only the lanes at the ends of z-runs (and every lane when the
// source run breaks) read the tables.
// (Table indices are linear in the packed coordinates, so the direct shift
// holds for any pack difference of one; the image shift needs iz_j =
// iz_{j-1} + 1 as well.)
// DPP wave_shr:1 (0x138): lane l takes lane l - 1, lane 0 keeps its value;
// DPP wave_shl:1 (0x130): lane l takes lane l + 1, lane 63 keeps its value.
template <int CTRL>
__device__ __forceinline__ float wave_shift(float v) {
  const int b = __builtin_bit_cast(int, v);
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(b, b, CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double wave_shift(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = (int)(b & 0xffffffffLL), hi = (int)(b >> 32);
  const int l2 = __builtin_amdgcn_update_dpp(lo, lo, CTRL, 0xF, 0xF, false);
  const int h2 = __builtin_amdgcn_update_dpp(hi, hi, CTRL, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((long long)h2 << 32) | (long long)(unsigned)l2);
}

template <typename R, typename SRC>
__global__ __launch_bounds__(64) void k_ewt_shift(const FarItem* __restrict__ items,
                                                 const SolveDev* __restrict__ solves,
                                                 const GState* __restrict__ st, int which,
                                                 const SRC* __restrict__ src,
                                                 const int* __restrict__ gpack,
                                                 const int* __restrict__ gizj2,
                                                 const R* __restrict__ tab,
                                                 double2* __restrict__ part, int p32) {
  const FarItem it = items[blockIdx.x];
  if (!takes_part(st, it.s, which)) return;
  const SolveDev S = solves[it.s];
  const int lane = threadIdx.x;
  const int i = it.t0 + lane;
  const bool valid = i < S.n;
  const int* __restrict__ gp = gpack + S.pos_off;
  const int* __restrict__ gz = gizj2 + S.pos_off;
  const int pi = gp[valid ? i : it.t0] + S.ew_center;
  // z-neighbours among the targets (invalid lanes never qualify)
  const int pprev = __builtin_amdgcn_update_dpp(pi, pi, 0x138, 0xF, 0xF, false);
  const int pnext = __builtin_amdgcn_update_dpp(pi, pi, 0x130, 0xF, 0xF, false);
  const bool nb_prev = lane > 0 && valid && pi - pprev == 1;
  const bool nb_next = lane < 63 && (i + 1) < S.n && pnext - pi == 1;
  const SRC* __restrict__ sp = src + S.src_off;
  const R* __restrict__ td = tab + 12 * S.ew_off;
  const R* __restrict__ ti = td + 12 * S.ewn;
  const bool sub = S.has_sub != 0;
  R hi[6], lo[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) hi[q] = lo[q] = R(0);
  R t[12], u[12];
#pragma unroll
  for (int q = 0; q < 12; ++q) t[q] = u[q] = R(0);
  int qprev = -2, gprev = -1 << 30;
  for (int j0 = it.c0; j0 < it.c1; j0 += 32) {
    const int j1 = min(j0 + 32, it.c1);
    R acc[6] = {0, 0, 0, 0, 0, 0};
    for (int j = j0; j < j1; ++j) {
      const int qj = gp[j], gj = gz[j];
      // uniform: source j follows j - 1 along the table's fastest index; the
      // image index moves the other way only if iz also grew by one (always
      // the case with more than one z plane)
      const bool run = qj == qprev + 1;
      const bool run_img = run && gj == gprev + 2;
      qprev = qj;
      gprev = gj;
      const long long e = (long long)(pi - qj);
      if (run) {
#pragma unroll
        for (int q = 0; q < 12; ++q) t[q] = wave_shift<0x138>(t[q]);
      }
      if (!(run && nb_prev)) {
        const R* a = td + 12 * e;
#pragma unroll
        for (int q = 0; q < 12; ++q) t[q] = a[q];
      }
      R v[6];
#pragma unroll
      for (int q = 0; q < 6; ++q) v[q] = sp[j].v[q];
      apply_sym<R>(t, v, acc);
      if (sub) {
        if (run_img) {
#pragma unroll
          for (int q = 0; q < 12; ++q) u[q] = wave_shift<0x130>(u[q]);
        }
        if (!(run_img && nb_next)) {
          const R* b = ti + 12 * (e + gj);
#pragma unroll
          for (int q = 0; q < 12; ++q) u[q] = b[q];
        }
#pragma unroll
        for (int q = 0; q < 6; ++q) v[q] = sp[j].w[q];
        apply_sym<R>(u, v, acc);
      }
    }
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      if constexpr (std::is_same<R, float>::value) {
        const float s2 = hi[q] + acc[q];
        const float bb = s2 - hi[q];
        lo[q] += (hi[q] - (s2 - bb)) + (acc[q] - bb);
        hi[q] = s2;
      } else {
        hi[q] += acc[q];
      }
    }
  }
  if (valid) {
    const long long e = S.part_off + (long long)(it.c0 / LH_TILE) * 3LL * S.n + 3LL * i;
#pragma unroll
    for (int c = 0; c < 3; ++c)
      put_part(part, e + c, (double)hi[2 * c] + (double)lo[2 * c],
               (double)hi[2 * c + 1] + (double)lo[2 * c + 1], p32);
  }
}

// Per-pair form (positions off any regular grid), FP64.
__global__ __launch_bounds__(256) void k_ewp_row(const FarItem* __restrict__ items,
                                                 const SolveDev* __restrict__ solves,
                                                 const GState* __restrict__ st, int which,
                                                 const Src64* __restrict__ src,
                                                 const double* __restrict__ ewl,
                                                 double2* __restrict__ part, int p32) {
  const FarItem it = items[blockIdx.x];
  if (!takes_part(st, it.s, which)) return;
  const SolveDev S = solves[it.s];
  const int i = it.t0 + threadIdx.x;
  const bool valid = i < S.n;
  const Src64* __restrict__ sp = src + S.src_off;
  const Src64& me = sp[valid ? i : 0];
  const EwReal* rl = reinterpret_cast<const EwReal*>(ewl) + S.ewl_off;
  const EwRecip* rq = reinterpret_cast<const EwRecip*>(rl + S.ew_nreal);
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int j = it.c0; j < it.c1; ++j) {
    const Src64& r = sp[j];
    double t[12];
    const double dx = me.x - r.x, dy = me.y - r.y, dz = me.z - r.z;
    ewald_tensor(S, rl, rq, dx, dy, dz, t);
    apply_sym<double>(t, r.v, acc);
    if (S.has_sub) {
      // image of source j at (x_j, y_j, 2 z0 - z_j) (assembly.rs:278-295)
      const double ez = me.z - (S.z2 - r.z);
      if (!(r2_exact(dx, dy, ez) < 1e-20)) {
        ewald_tensor(S, rl, rq, dx, dy, ez, t);
        apply_sym<double>(t, r.w, acc);
      }
    }
  }
  if (valid) {
    const long long e = S.part_off + (long long)(it.c0 / LH_TILE) * 3LL * S.n + 3LL * i;
#pragma unroll
    for (int c = 0; c < 3; ++c) put_part(part, e + c, acc[2 * c], acc[2 * c + 1], p32);
  }
}

// ---------------------------------------------------------------------------
// Launchers (host).
// ---------------------------------------------------------------------------

void launch_ewald_table(hipStream_t st, int nitems, const SolveDev* solves, const int* item_s,
                        const long long* item_e, const double* ewl, double* t64, float* t32) {
  if (nitems <= 0) return;
  hipLaunchKernelGGL(k_ewald_table, dim3(nitems), dim3(256), 0, st, solves, item_s, item_e, ewl,
                     t64, t32);
}

void launch_ewt_shift32(hipStream_t st, int nitems, const FarItem* items, const SolveDev* solves,
                        const GState* gs, int which, const Src32* src, const int* gpack,
                        const int* gizj2, const float* tab, double2* part, int p32) {
  if (nitems <= 0) return;
  hipLaunchKernelGGL((k_ewt_shift<float, Src32>), dim3(nitems), dim3(64), 0, st, items, solves,
                     gs, which, src, gpack, gizj2, tab, part, p32);
}

void launch_ewt_shift64(hipStream_t st, int nitems, const FarItem* items, const SolveDev* solves,
                        const GState* gs, int which, const Src64* src, const int* gpack,
                        const int* gizj2, const double* tab, double2* part) {
  if (nitems <= 0) return;
  hipLaunchKernelGGL((k_ewt_shift<double, Src64>), dim3(nitems), dim3(64), 0, st, items, solves,
                     gs, which, src, gpack, gizj2, tab, part, 0);
}

void launch_ewp(hipStream_t st, int nitems, const FarItem* items, const SolveDev* solves,
                const GState* gs, int which, const Src64* src, const double* ewl, double2* part,
                int block, int p32) {
  if (nitems <= 0) return;
  hipLaunchKernelGGL(k_ewp_row, dim3(nitems), dim3(block), 0, st, items, solves, gs, which, src,
                     ewl, part, p32);
}

}  // namespace lh
