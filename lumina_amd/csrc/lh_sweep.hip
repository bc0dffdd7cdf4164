// Per-wavelength setup of a sweep on the device (runner.rs:150-186): for
// every solve s (one wavelength) the permittivity of every material from its
// spline (spline.rs:68-90, johnson_christy.rs:266-279), the polarisability
// of every optical class (Clausius-Mossotti with the radiative correction,
// types.rs:243-259, or the ellipsoid tensor, types.rs:286-317) and its
// inverse (assembly.rs:335-363), the substrate reflection factor
// (substrate.rs:148-197) and the incident field (types.rs:60-74), written
// straight into the engine's alpha^-1 blocks and RHS/E slots.
//
// Complex quotients are num-complex's (a conj(b)) / |b|^2, as the reference.
#include "lh_device.h"
#include "lh_engine.h"

namespace lh {

namespace {

__device__ __forceinline__ double2 c_mul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 c_sub(double2 a, double2 b) {
  return make_double2(a.x - b.x, a.y - b.y);
}
__device__ __forceinline__ double2 c_add(double2 a, double2 b) {
  return make_double2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ double2 c_div(double2 a, double2 b) {
  const double nsq = b.x * b.x + b.y * b.y;
  return make_double2((a.x * b.x + a.y * b.y) / nsq, (a.y * b.x - a.x * b.y) / nsq);
}

// Principal square root: Re >= 0, the sign of Im taken from the argument.
__device__ __forceinline__ double2 c_sqrt(double2 z) {
  const double m = sqrt(z.x * z.x + z.y * z.y);
  if (m == 0.0) return make_double2(0.0, 0.0);
  const double t = sqrt(0.5 * (m + fabs(z.x)));
  if (z.x >= 0.0) return make_double2(t, z.y / (2.0 * t));
  return make_double2(fabs(z.y) / (2.0 * t), copysign(t, z.y));
}

// spline.rs:68-90.
__device__ double spline_at(const double* xs, const double* ys, const double* y2, int n,
                            double x) {
  int lo = 0, hi = n - 1;
  while (hi - lo > 1) {
    const int mid = (lo + hi) / 2;
    if (xs[mid] > x) hi = mid;
    else lo = mid;
  }
  const double h = xs[hi] - xs[lo];
  const double a = (xs[hi] - x) / h;
  const double b = (x - xs[lo]) / h;
  return a * ys[lo] + b * ys[hi] + ((a * a * a - a) * y2[lo] + (b * b * b - b) * y2[hi]) * h * h / 6.0;
}

struct Splines {
  const int* off;
  const int* cnt;
  const double *xs, *yr, *yi, *y2r, *y2i;
};

__device__ double2 eps_at(const Splines& sp, int m, double lam) {
  const int o = sp.off[m], n = sp.cnt[m];
  return make_double2(spline_at(sp.xs + o, sp.yr + o, sp.y2r + o, n, lam),
                      spline_at(sp.xs + o, sp.yi + o, sp.y2i + o, n, lam));
}

// types.rs:256-259.
__device__ double2 radiative(double2 a_cm, double k) {
  const double2 corr = make_double2(0.0, (k * k * k) / (6.0 * M_PI));
  return c_div(a_cm, c_sub(make_double2(1.0, 0.0), c_mul(corr, a_cm)));
}

// assembly.rs:335-363; false for |det| <= 1e-30.
__device__ bool invert3_dev(const double2* m, double2* out) {
  const double2 a = m[0], b = m[1], c = m[2], d = m[3], e = m[4], f = m[5], g = m[6], h = m[7],
                k = m[8];
  const double2 det =
      c_add(c_sub(c_mul(a, c_sub(c_mul(e, k), c_mul(f, h))), c_mul(b, c_sub(c_mul(d, k), c_mul(f, g)))),
            c_mul(c, c_sub(c_mul(d, h), c_mul(e, g))));
  const double dn = hypot(det.x, det.y);
  if (!(dn > 1e-30)) return false;
  const double nsq = det.x * det.x + det.y * det.y;
  const double2 inv = make_double2(det.x / nsq, -det.y / nsq);
  out[0] = c_mul(inv, c_sub(c_mul(e, k), c_mul(f, h)));
  out[1] = c_mul(inv, c_sub(c_mul(c, h), c_mul(b, k)));
  out[2] = c_mul(inv, c_sub(c_mul(b, f), c_mul(c, e)));
  out[3] = c_mul(inv, c_sub(c_mul(f, g), c_mul(d, k)));
  out[4] = c_mul(inv, c_sub(c_mul(a, k), c_mul(c, g)));
  out[5] = c_mul(inv, c_sub(c_mul(c, d), c_mul(a, f)));
  out[6] = c_mul(inv, c_sub(c_mul(d, h), c_mul(e, g)));
  out[7] = c_mul(inv, c_sub(c_mul(b, g), c_mul(a, h)));
  out[8] = c_mul(inv, c_sub(c_mul(a, e), c_mul(b, d)));
  return true;
}

}  // namespace

// alpha^-1 of class c at solve s, [S][C][9]; bad[0] counts singular blocks.
__global__ __launch_bounds__(64) void k_sweep_class(int S, int C, const double* __restrict__ lam,
                                                    double n_env, const int* __restrict__ cmat,
                                                    const int* __restrict__ chint,
                                                    const double* __restrict__ cvol,
                                                    const double* __restrict__ hdepol,
                                                    const double* __restrict__ hrot, Splines sp,
                                                    double2* __restrict__ out,
                                                    int* __restrict__ bad) {
  const int q = blockIdx.x * 64 + threadIdx.x;
  if (q >= S * C) return;
  const int s = q / C, c = q % C;
  const double L = lam[s];
  const double k = 2.0 * M_PI * n_env / L;   // cda/mod.rs:169-171
  const double em = n_env * n_env;
  const double2 eps = eps_at(sp, cmat[c], L);
  const double V = cvol[c];
  double2 blk[9];
  const int h = chint[c];
  if (h < 0) {
    // types.rs:243-246 Clausius-Mossotti, then the radiative correction
    const double2 num = make_double2(3.0 * V * (eps.x - em), 3.0 * V * eps.y);
    const double2 den = make_double2(eps.x + 2.0 * em, eps.y);
    const double2 a = radiative(c_div(num, den), k);
#pragma unroll
    for (int u = 0; u < 9; ++u) blk[u] = make_double2(0.0, 0.0);
    blk[0] = blk[4] = blk[8] = a;
  } else {
    // types.rs:286-317: per-axis CM with depolarisation L_q, the radiative
    // correction, then R diag(alpha) R^T
    const double2 de = make_double2(eps.x - em, eps.y);
    double2 al[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double Lq = hdepol[3 * h + a];
      const double2 den = make_double2(em + Lq * de.x, Lq * de.y);
      al[a] = radiative(c_div(make_double2(V * de.x, V * de.y), den), k);
    }
    const double* R = hrot + 9 * h;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double2 acc = make_double2(0.0, 0.0);
        for (int a = 0; a < 3; ++a) {
          const double r1 = R[i * 3 + a], r2 = R[j * 3 + a];
          acc = c_add(acc, make_double2(al[a].x * r1 * r2, al[a].y * r1 * r2));
        }
        blk[i * 3 + j] = acc;
      }
  }
  double2 inv[9];
  if (!invert3_dev(blk, inv)) {
    atomicAdd(bad, 1);
#pragma unroll
    for (int u = 0; u < 9; ++u) inv[u] = make_double2(0.0, 0.0);
  }
  double2* o = out + 9LL * q;
#pragma unroll
  for (int u = 0; u < 9; ++u) o[u] = inv[u];
}

// The class blocks into each solve's alpha^-1 storage (one block per solve
// when its layout is uniform, else one per dipole).
__global__ __launch_bounds__(256) void k_sweep_ainv(const SolveDev* __restrict__ solves, int S,
                                                    int C, const int* __restrict__ cls_of,
                                                    const double2* __restrict__ cls_ainv,
                                                    double2* __restrict__ ainv) {
  const int s = blockIdx.y;
  const SolveDev D = solves[s];
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int nb = D.ainv_uniform ? 1 : D.n;
  if (i >= nb) return;
  const int c = (D.ainv_uniform || cls_of == nullptr) ? 0 : cls_of[i];
  const double2* src = cls_ainv + 9LL * ((long long)s * C + c);
  double2* dst = ainv + D.ainv_off + 9LL * i;
#pragma unroll
  for (int u = 0; u < 9; ++u) dst[u] = src[u];
}

// substrate_reflection_factor (substrate.rs:148-197) at every wavelength.
__global__ __launch_bounds__(64) void k_sweep_delta(int S, const double* __restrict__ lam,
                                                    double n_env, int mat, int retarded,
                                                    Splines sp, double2* __restrict__ delta) {
  const int s = blockIdx.x * 64 + threadIdx.x;
  if (s >= S) return;
  const double2 es = eps_at(sp, mat, lam[s]);
  const double ee = n_env * n_env;
  double2 r;
  if (retarded) {
    const double2 ns = c_sqrt(es);
    const double2 ne = make_double2(sqrt(ee), 0.0);
    r = c_div(c_sub(ns, ne), c_add(ns, ne));
  } else {
    r = c_div(make_double2(es.x - ee, es.y), make_double2(es.x + ee, es.y));
  }
  delta[s] = r;
}

// types.rs:60-74 E0 e exp(i k d.r) at every dipole of every solve, into the
// RHS and E slots.
__global__ __launch_bounds__(256) void k_sweep_incident(const SolveDev* __restrict__ solves,
                                                        const double* __restrict__ px,
                                                        const double* __restrict__ py,
                                                        const double* __restrict__ pz, double dx,
                                                        double dy, double dz, double amp,
                                                        const double* __restrict__ pol,
                                                        double2* __restrict__ vec) {
  const int s = blockIdx.y;
  const SolveDev D = solves[s];
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= D.n) return;
  const int p = D.pos_off + i;
  const double kdotr = D.k * (dx * px[p] + dy * py[p] + dz * pz[p]);
  double sn, cs;
  sincos_f64(kdotr, sn, cs);
  const long long n3 = 3LL * D.n;
  double2* rhs = vec + D.vec_off + (long long)LH_SLOT_RHS * n3 + 3LL * i;
  double2* e = vec + D.vec_off + (long long)LH_SLOT_E * n3 + 3LL * i;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double er = amp * pol[2 * c], ei = amp * pol[2 * c + 1];
    const double2 v = make_double2(er * cs - ei * sn, er * sn + ei * cs);
    rhs[c] = v;
    e[c] = v;
  }
}

// ---------------------------------------------------------------------------
// Launcher (host).
// ---------------------------------------------------------------------------

void launch_sweep_setup(hipStream_t st, const SweepDevArgs& a, const SolveDev* solves, int max_n,
                        double2* ainv) {
  const Splines sp{a.moff, a.mcnt, a.xs, a.yr, a.yi, a.y2r, a.y2i};
  const int SC = a.S * a.C;
  hipLaunchKernelGGL(k_sweep_class, dim3((SC + 63) / 64), dim3(64), 0, st, a.S, a.C, a.lam,
                     a.n_env, a.cmat, a.chint, a.cvol, a.hdepol, a.hrot, sp, a.cls_ainv, a.bad);
  hipLaunchKernelGGL(k_sweep_ainv, dim3((max_n + 255) / 256, a.S), dim3(256), 0, st, solves, a.S,
                     a.C, a.cls_of, a.cls_ainv, ainv);
  if (a.sub_material >= 0)
    hipLaunchKernelGGL(k_sweep_delta, dim3((a.S + 63) / 64), dim3(64), 0, st, a.S, a.lam, a.n_env,
                       a.sub_material, a.sub_retarded, sp, a.delta);
}

void launch_sweep_incident(hipStream_t st, const SolveDev* solves, int S, int max_n,
                           const double* px, const double* py, const double* pz,
                           const double dir[3], double amp, const double* pol, double2* vec) {
  hipLaunchKernelGGL(k_sweep_incident, dim3((max_n + 255) / 256, S), dim3(256), 0, st, solves, px,
                     py, pz, dir[0], dir[1], dir[2], amp, pol, vec);
}

}  // namespace lh
