// CPU FP64 restatement of the Lumina CDA solve hot path.
//
// TEST INFRASTRUCTURE ONLY. This file is the oracle: the parity reference the
// HIP path is checked against. Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load it, and only as the checker (or the
// timed CPU baseline), never as a product code path.
//
// Every function follows one reference function; the reference file:line is
// given above each (paths relative to the reference repository root). The
// arithmetic keeps the reference's operation order where that is cheap:
// complex values use the naive num-complex 0.4 formulas (product, quotient
// a*conj(b)/|b|^2, exp = from_polar(e^re, im)), sums run in the reference's
// loop order, and nothing is fused (the build uses -ffp-contract=off, as Rust
// never contracts).
//
// Parity status per function (how each is pinned) is written next to it.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>
#include <algorithm>
#include <atomic>

namespace {

struct C64 {
  double re, im;
};

inline C64 c(double r, double i = 0.0) { return C64{r, i}; }
inline C64 add(C64 a, C64 b) { return c(a.re + b.re, a.im + b.im); }
inline C64 sub(C64 a, C64 b) { return c(a.re - b.re, a.im - b.im); }
inline C64 mul(C64 a, C64 b) {
  return c(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
inline C64 muls(C64 a, double s) { return c(a.re * s, a.im * s); }
inline C64 divs(C64 a, double s) { return c(a.re / s, a.im / s); }
// num-complex 0.4 Div: (a * conj(b)) / |b|^2, without scaling.
inline C64 cdiv(C64 a, C64 b) {
  double nsq = b.re * b.re + b.im * b.im;
  double re = a.re * b.re + a.im * b.im;
  double im = a.im * b.re - a.re * b.im;
  return c(re / nsq, im / nsq);
}
inline C64 conj(C64 a) { return c(a.re, -a.im); }
inline C64 neg(C64 a) { return c(-a.re, -a.im); }
inline double norm(C64 a) { return std::hypot(a.re, a.im); }
inline double norm_sqr(C64 a) { return a.re * a.re + a.im * a.im; }
// num-complex exp: from_polar(exp(re), im)
inline C64 cexp(C64 a) {
  double r = std::exp(a.re);
  return c(r * std::cos(a.im), r * std::sin(a.im));
}
// num-complex 0.4 sqrt (special cases for purely real / imaginary input,
// polar form otherwise).
inline C64 csqrt(C64 z) {
  if (z.im == 0.0) {
    if (!std::signbit(z.re)) return c(std::sqrt(z.re), z.im);
    double im = std::sqrt(-z.re);
    return std::signbit(z.im) ? c(0.0, -im) : c(0.0, im);
  }
  if (z.re == 0.0) {
    double x = std::sqrt(std::fabs(z.im) / 2.0);
    return std::signbit(z.im) ? c(x, -x) : c(x, x);
  }
  double r = std::hypot(z.re, z.im);
  double th = std::atan2(z.im, z.re);
  double sr = std::sqrt(r);
  double h = th / 2.0;
  return c(sr * std::cos(h), sr * std::sin(h));
}

const double PI = 3.14159265358979323846;

typedef C64 T33[3][3];

// greens.rs:34-71 dyadic_greens_tensor.
// Pinned by: 50-digit mpmath golden tensors (tests/golden/greens_point.json),
// reciprocity greens.rs:175-190.
void greens_point(const double r1[3], const double r2[3], double k, T33 g) {
  double rx = r1[0] - r2[0];
  double ry = r1[1] - r2[1];
  double rz = r1[2] - r2[2];
  double r_sq = rx * rx + ry * ry + rz * rz;
  double r = std::sqrt(r_sq);
  double kr = k * r;
  double kr_sq = kr * kr;
  C64 ikr = c(0.0, kr);
  C64 exp_ikr = cexp(ikr);
  C64 prefactor = divs(muls(exp_ikr, k * k), 4.0 * PI * r);
  C64 a = add(c(1.0), cdiv(sub(ikr, c(1.0)), c(kr_sq)));
  C64 b = cdiv(sub(sub(c(3.0), mul(c(3.0), ikr)), c(kr_sq)), c(kr_sq));
  double rh[3] = {rx / r, ry / r, rz / r};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double d = (i == j) ? 1.0 : 0.0;
      // prefactor * (a * delta + b * r_hat_i * r_hat_j)
      C64 t = add(muls(a, d), muls(muls(b, rh[i]), rh[j]));
      g[i][j] = mul(prefactor, t);
    }
}

// ewald.rs:283-307 dyadic_greens_tensor_at (same as greens_point at
// displacement s).
void greens_at(const double s[3], double k, T33 g) {
  const double z[3] = {0.0, 0.0, 0.0};
  double r_sq = s[0] * s[0] + s[1] * s[1] + s[2] * s[2];
  double r = std::sqrt(r_sq);
  double kr = k * r;
  C64 ikr = c(0.0, kr);
  C64 exp_ikr = cexp(ikr);
  C64 prefactor = divs(muls(exp_ikr, k * k), 4.0 * PI * r);
  C64 a = add(c(1.0), cdiv(sub(ikr, c(1.0)), c(kr * kr)));
  C64 b = cdiv(sub(sub(c(3.0), mul(c(3.0), ikr)), c(kr * kr)), c(kr * kr));
  double rh[3] = {s[0] / r, s[1] / r, s[2] / r};
  (void)z;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double d = (i == j) ? 1.0 : 0.0;
      C64 t = add(muls(a, d), muls(muls(b, rh[i]), rh[j]));
      g[i][j] = mul(prefactor, t);
    }
}

// greens.rs:94-167 dyadic_greens_tensor_filtered (FCD / IGT, 27-node
// Gauss-Legendre average over the source cell when dist <= 2*cell_size).
// Pinned by: mpmath quadrature golden values (tests/golden/greens_fcd.json),
// greens.rs:201-261 (FCD -> point G at 10 cells; finite and different at R=d).
void greens_fcd_average(const double ro[3], const double rs[3], double k, double cell_size,
                        T33 out);
void greens_filtered(const double ro[3], const double rs[3], double k,
                     double cell_size, T33 out) {
  double dx = ro[0] - rs[0];
  double dy = ro[1] - rs[1];
  double dz = ro[2] - rs[2];
  double dist = std::sqrt(dx * dx + dy * dy + dz * dz);
  if (dist > 2.0 * cell_size) {
    greens_point(ro, rs, k, out);
    return;
  }
  greens_fcd_average(ro, rs, k, cell_size, out);
}

// The 27-node average of greens.rs:120-164 (the branch dist <= 2 cell_size).
void greens_fcd_average(const double ro[3], const double rs[3], double k, double cell_size,
                        T33 out) {
  const double nodes[3] = {-std::sqrt(3.0 / 5.0), 0.0, std::sqrt(3.0 / 5.0)};
  const double wts[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
  double half_d = cell_size / 2.0;
  T33 acc;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) acc[a][b] = c(0.0);
  double total_w = 0.0;
  for (int ix = 0; ix < 3; ++ix)
    for (int iy = 0; iy < 3; ++iy)
      for (int iz = 0; iz < 3; ++iz) {
        double rp[3] = {rs[0] + half_d * nodes[ix], rs[1] + half_d * nodes[iy],
                        rs[2] + half_d * nodes[iz]};
        double ddx = ro[0] - rp[0], ddy = ro[1] - rp[1], ddz = ro[2] - rp[2];
        double r_sq = ddx * ddx + ddy * ddy + ddz * ddz;
        if (r_sq < 1e-20) continue;
        double w = wts[ix] * wts[iy] * wts[iz];
        T33 g;
        greens_point(ro, rp, k, g);
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) acc[a][b] = add(acc[a][b], mul(c(w), g[a][b]));
        total_w += w;
      }
  if (total_w > 1e-30) {
    C64 inv_w = c(1.0 / total_w);
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) acc[a][b] = mul(acc[a][b], inv_w);
  }
  std::memcpy(out, acc, sizeof(T33));
}

// ewald.rs:316-327 erfc_real, Abramowitz & Stegun 7.1.26 (|err| < 1.5e-7).
// Pinned by: ewald.rs:816-824 known answers and the A&S error bound against
// scipy.special.erfc (tests/test_oracle.py).
double erfc_real(double x) {
  if (x < 0.0) return 2.0 - erfc_real(-x);
  double t = 1.0 / (1.0 + 0.3275911 * x);
  double poly =
      t * (0.254829592 +
           t * (-0.284496736 + t * (1.421413741 + t * (-1.453152027 + t * 1.061405429))));
  return poly * std::exp(-x * x);
}

// ewald.rs:334-338 csqrt_positive_imag.
// Pinned by: numpy's principal square root over every quadrant and both
// axes, flipped to Im >= 0 (4e-16), and ewald.rs:827-838 known answers.
C64 csqrt_pos_imag(C64 z) {
  C64 s = csqrt(z);
  return (s.im < 0.0) ? neg(s) : s;
}

struct Lattice2D {
  double a1[3], a2[3];
  double eta;     // <= 0 selects the automatic sqrt(pi / A)
  int n_real, n_recip;
};

// Lattice helpers (lattice.rs:75-233). Pinned by: lattice.rs:243-301 known
// answers (square lattice b_i = 2 pi / a, eta = sqrt(pi / A)), b_i . a_j =
// 2 pi delta_ij in the spectral-sum closed forms of tests/test_oracle_ewald.py.
// lattice.rs:75-89 cell_measure (2D branch).
double cell_measure(const Lattice2D& L) {
  double cx = L.a1[1] * L.a2[2] - L.a1[2] * L.a2[1];
  double cy = L.a1[2] * L.a2[0] - L.a1[0] * L.a2[2];
  double cz = L.a1[0] * L.a2[1] - L.a1[1] * L.a2[0];
  return std::sqrt(cx * cx + cy * cy + cz * cz);
}
// lattice.rs:96-102 eta.
double lattice_eta(const Lattice2D& L) {
  if (L.eta > 0.0) return L.eta;
  return std::sqrt(PI / cell_measure(L));
}
// lattice.rs:113-147 reciprocal_vectors.
void recip_vectors(const Lattice2D& L, double b1[3], double b2[3]) {
  const double two_pi = 2.0 * PI;
  double a2xz[2] = {L.a2[1], -L.a2[0]};
  double d1 = L.a1[0] * a2xz[0] + L.a1[1] * a2xz[1];
  b1[0] = two_pi * a2xz[0] / d1;
  b1[1] = two_pi * a2xz[1] / d1;
  b1[2] = 0.0;
  double zxa1[2] = {-L.a1[1], L.a1[0]};
  double d2 = L.a2[0] * zxa1[0] + L.a2[1] * zxa1[1];
  b2[0] = two_pi * zxa1[0] / d2;
  b2[1] = two_pi * zxa1[1] / d2;
  b2[2] = 0.0;
}
// lattice.rs:168-196 real_lattice_points (2D, m1 outer, m2 inner).
std::vector<double> real_points(const Lattice2D& L, int ns) {
  std::vector<double> p;
  for (int m1 = -ns; m1 <= ns; ++m1)
    for (int m2 = -ns; m2 <= ns; ++m2) {
      p.push_back((double)m1 * L.a1[0] + (double)m2 * L.a2[0]);
      p.push_back((double)m1 * L.a1[1] + (double)m2 * L.a2[1]);
      p.push_back((double)m1 * L.a1[2] + (double)m2 * L.a2[2]);
    }
  return p;
}
// lattice.rs:203-233 recip_lattice_points (2D).
std::vector<double> recip_points(const Lattice2D& L, int ns) {
  double b1[3], b2[3];
  recip_vectors(L, b1, b2);
  std::vector<double> p;
  for (int n1 = -ns; n1 <= ns; ++n1)
    for (int n2 = -ns; n2 <= ns; ++n2) {
      p.push_back((double)n1 * b1[0] + (double)n2 * b2[0]);
      p.push_back((double)n1 * b1[1] + (double)n2 * b2[1]);
      p.push_back(0.0);
    }
  return p;
}

// ewald.rs:119-174 real_space (2D branch: erfc damping).
// Pinned by: the one-shell sum against an independent damped point sum with
// scipy's erfc times the mpmath-pinned point tensor (to the A&S 7.1.26
// error), the A&S erfc against scipy, and the invariants of
// tests/test_oracle.py (split identity, reciprocity, square symmetry, shell
// convergence).
void ewald_real(const Lattice2D& L, const double r[3], const double kb[3], double k,
                T33 g) {
  double rn2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
  bool r_zero = rn2 < 1e-24;
  double eta = lattice_eta(L);
  std::vector<double> lp = real_points(L, L.n_real);
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) g[a][b] = c(0.0);
  for (size_t t = 0; t < lp.size() / 3; ++t) {
    const double* R = &lp[3 * t];
    bool lp_zero = std::fabs(R[0]) < 1e-12 && std::fabs(R[1]) < 1e-12 && std::fabs(R[2]) < 1e-12;
    if (r_zero && lp_zero) continue;
    double s[3] = {r[0] - R[0], r[1] - R[1], r[2] - R[2]};
    double s_sq = s[0] * s[0] + s[1] * s[1] + s[2] * s[2];
    if (s_sq < 1e-24) continue;
    double s_norm = std::sqrt(s_sq);
    double barg = kb[0] * R[0] + kb[1] * R[1] + kb[2] * R[2];
    C64 bph = c(std::cos(barg), std::sin(barg));
    double damp = erfc_real(eta * s_norm);
    T33 gs;
    greens_at(s, k, gs);
    C64 f = muls(bph, damp);
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) g[a][b] = add(g[a][b], mul(f, gs[a][b]));
  }
}

// ewald.rs:186-273 recip_space (2D spectral sum).
// Pinned by: the closed forms the documented sum reduces to
// (tests/test_oracle_ewald.py): far above the lattice only the (0,0) order
// survives, at normal and at oblique incidence; the first evanescent shell
// of a rectangular lattice; an independent numpy evaluation of the sum; and
// planted mistakes (prefactor, exponent sign, k_z, phase sign, basis) that
// each fail one of them.
void ewald_recip(const Lattice2D& L, const double r[3], const double kb[3], double k,
                 T33 g) {
  double area = cell_measure(L);
  std::vector<double> gp = recip_points(L, L.n_recip);
  C64 iu = c(0.0, 1.0);
  C64 pre = divs(iu, 2.0 * area);
  double rho_x = r[0], rho_y = r[1], z = r[2];
  double abs_z = std::fabs(z);
  double sign_z = (z >= 0.0) ? 1.0 : -1.0;
  C64 k2 = c(k * k);
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) g[a][b] = c(0.0);
  for (size_t t = 0; t < gp.size() / 3; ++t) {
    double qx = kb[0] + gp[3 * t + 0];
    double qy = kb[1] + gp[3 * t + 1];
    double q2r = qx * qx + qy * qy;
    C64 q = csqrt_pos_imag(sub(c(q2r), k2));
    if (norm(q) < 1e-10) continue;
    double parg = qx * rho_x + qy * rho_y;
    C64 bph = c(std::cos(parg), std::sin(parg));
    C64 eqz = cexp(muls(neg(q), abs_z));
    C64 coeff = cdiv(mul(mul(pre, bph), eqz), q);
    C64 kt[3] = {c(qx), c(qy), muls(mul(iu, q), sign_z)};
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        C64 delta = (a == b) ? k2 : c(0.0);
        C64 m_ab = sub(delta, mul(kt[a], kt[b]));
        g[a][b] = add(g[a][b], mul(coeff, m_ab));
      }
  }
}

// ewald.rs:91-106 evaluate = real + recip.
void ewald_eval(const Lattice2D& L, const double r[3], const double kb[3], double k,
                T33 out) {
  T33 gr, gg;
  ewald_real(L, r, kb, k, gr);
  ewald_recip(L, r, kb, k, gg);
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) out[a][b] = add(gr[a][b], gg[a][b]);
}

// assembly.rs:335-363 invert_3x3 (cofactor inverse).
// Pinned by: identity, scalar and random complex matrices against
// numpy.linalg.inv (tests/test_oracle.py).
void invert3(const C64 m[9], C64 out[9]) {
  C64 a = m[0], b = m[1], cc = m[2], d = m[3], e = m[4], f = m[5], g = m[6], h = m[7],
      k = m[8];
  C64 det = add(sub(mul(a, sub(mul(e, k), mul(f, h))), mul(b, sub(mul(d, k), mul(f, g)))),
                mul(cc, sub(mul(d, h), mul(e, g))));
  C64 inv = cdiv(c(1.0), det);
  out[0] = mul(inv, sub(mul(e, k), mul(f, h)));
  out[1] = mul(inv, sub(mul(cc, h), mul(b, k)));
  out[2] = mul(inv, sub(mul(b, f), mul(cc, e)));
  out[3] = mul(inv, sub(mul(f, g), mul(d, k)));
  out[4] = mul(inv, sub(mul(a, k), mul(cc, g)));
  out[5] = mul(inv, sub(mul(cc, d), mul(a, f)));
  out[6] = mul(inv, sub(mul(d, h), mul(e, g)));
  out[7] = mul(inv, sub(mul(b, g), mul(a, h)));
  out[8] = mul(inv, sub(mul(a, e), mul(b, d)));
}

// The matvec problem: exactly the arguments of matvec_on_the_fly
// (assembly.rs:209-218) with the Dipole list split into positions and
// polarisabilities.
struct Problem {
  int64_t n;
  const double* pos;      // [n][3]
  const C64* alpha;       // [n][9] row-major, or [9] when alpha_uniform
  int alpha_uniform;
  double k;
  int use_fcd;
  double cell_size;
  const Lattice2D* lattice;  // nullptr: free space
  double kb[3];
  int has_sub;
  double z_interface;
  C64 delta;
  // Finite-difference support (no reference counterpart): when set, the FCD
  // near/far decision of pair (i, j) is taken on these positions instead of
  // pos, so that a perturbed dipole keeps its unperturbed classification.
  const double* cls_pos = nullptr;
};

// The direct tensor of pair (i, j) on the free-space path: FCD filtered or
// point (assembly.rs:252-261), with the classification of cls_pos if set.
void direct_tensor(const Problem& P, int64_t i, int64_t j, T33 g) {
  const double* ri = P.pos + 3 * i;
  const double* rj = P.pos + 3 * j;
  if (!P.use_fcd) {
    greens_point(ri, rj, P.k, g);
    return;
  }
  if (!P.cls_pos) {
    greens_filtered(ri, rj, P.k, P.cell_size, g);
    return;
  }
  const double* ci = P.cls_pos + 3 * i;
  const double* cj = P.cls_pos + 3 * j;
  double dx = ci[0] - cj[0], dy = ci[1] - cj[1], dz = ci[2] - cj[2];
  double dist = std::sqrt(dx * dx + dy * dy + dz * dz);
  if (dist > 2.0 * P.cell_size) greens_point(ri, rj, P.k, g);
  else greens_fcd_average(ri, rj, P.k, P.cell_size, g);
}

// ---------------------------------------------------------------------------
// Adjoint gradients (no reference counterpart; the reference has no
// derivatives). Derived here independently of the device code: the product
// rule on G_ab = pf (a delta_ab + b rh_a rh_b) of greens.rs:34-71 with
// pf(r), a(r), b(r) and d rh_a / d r_c = (delta_ac - rh_a rh_c) / r.
// Pinned by: central finite differences of C_ext on this oracle's own
// solves (tests/test_oracle.py) and on the device solves.
// ---------------------------------------------------------------------------

typedef C64 T333[3][3][3];   // [c][a][b] = d G_ab / d (r_obs)_c

void greens_point_grad(const double ro[3], const double rs[3], double k, T333 dg) {
  double rx = ro[0] - rs[0], ry = ro[1] - rs[1], rz = ro[2] - rs[2];
  double r_sq = rx * rx + ry * ry + rz * rz;
  double r = std::sqrt(r_sq);
  double kr = k * r;
  double kr_sq = kr * kr;
  C64 ikr = c(0.0, kr);
  C64 pf = divs(muls(cexp(ikr), k * k), 4.0 * PI * r);
  C64 a = add(c(1.0), cdiv(sub(ikr, c(1.0)), c(kr_sq)));
  C64 b = cdiv(sub(sub(c(3.0), mul(c(3.0), ikr)), c(kr_sq)), c(kr_sq));
  C64 dpf = mul(pf, c(-1.0 / r, k));                                      // pf (ik - 1/r)
  C64 da = c(2.0 / (k * k * r * r * r), -1.0 / (k * r * r));             // d a / dr
  C64 db = c(-6.0 / (k * k * r * r * r), 3.0 / (k * r * r));             // d b / dr
  double rh[3] = {rx / r, ry / r, rz / r};
  C64 pfb = mul(pf, b);
  for (int cc = 0; cc < 3; ++cc)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double dij = (i == j) ? 1.0 : 0.0;
        double dic = (i == cc) ? 1.0 : 0.0;
        double djc = (j == cc) ? 1.0 : 0.0;
        double hh = rh[i] * rh[j];
        C64 radial = add(mul(dpf, add(muls(a, dij), muls(b, hh))),
                         mul(pf, add(muls(da, dij), muls(db, hh))));
        double ang = ((dic - rh[i] * rh[cc]) * rh[j] + rh[i] * (djc - rh[j] * rh[cc])) / r;
        dg[cc][i][j] = add(muls(radial, rh[cc]), muls(pfb, ang));
      }
}

// Gradient of the 27-node FCD average (greens.rs:120-164), with the same
// node skip and weight normalisation.
void greens_fcd_average_grad(const double ro[3], const double rs[3], double k, double cell_size,
                             T333 out) {
  const double nodes[3] = {-std::sqrt(3.0 / 5.0), 0.0, std::sqrt(3.0 / 5.0)};
  const double wts[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
  double half_d = cell_size / 2.0;
  T333 acc;
  for (int q = 0; q < 27; ++q) (&acc[0][0][0])[q] = c(0.0);
  double total_w = 0.0;
  for (int ix = 0; ix < 3; ++ix)
    for (int iy = 0; iy < 3; ++iy)
      for (int iz = 0; iz < 3; ++iz) {
        double rp[3] = {rs[0] + half_d * nodes[ix], rs[1] + half_d * nodes[iy],
                        rs[2] + half_d * nodes[iz]};
        double ddx = ro[0] - rp[0], ddy = ro[1] - rp[1], ddz = ro[2] - rp[2];
        if (ddx * ddx + ddy * ddy + ddz * ddz < 1e-20) continue;
        double w = wts[ix] * wts[iy] * wts[iz];
        T333 g;
        greens_point_grad(ro, rp, k, g);
        for (int q = 0; q < 27; ++q) (&acc[0][0][0])[q] = add((&acc[0][0][0])[q], muls((&g[0][0][0])[q], w));
        total_w += w;
      }
  for (int q = 0; q < 27; ++q)
    (&out[0][0][0])[q] = total_w > 1e-30 ? divs((&acc[0][0][0])[q], total_w) : (&acc[0][0][0])[q];
}

// Gradient of direct_tensor with respect to the observation position r_i.
void direct_tensor_grad(const Problem& P, int64_t i, int64_t j, T333 dg) {
  const double* ri = P.pos + 3 * i;
  const double* rj = P.pos + 3 * j;
  const double* ci = P.cls_pos ? P.cls_pos + 3 * i : ri;
  const double* cj = P.cls_pos ? P.cls_pos + 3 * j : rj;
  double dx = ci[0] - cj[0], dy = ci[1] - cj[1], dz = ci[2] - cj[2];
  double dist = std::sqrt(dx * dx + dy * dy + dz * dz);
  if (!P.use_fcd || dist > 2.0 * P.cell_size) greens_point_grad(ri, rj, P.k, dg);
  else greens_fcd_average_grad(ri, rj, P.k, P.cell_size, dg);
}

// assembly.rs:209-318 matvec_on_the_fly, one row block i.
// Pinned by: the dense assembled matrix times x (1e-15, FCD and image terms
// included) and a plain-Python 4-dipole sum (tests/test_oracle.py).
void matvec_row(const Problem& P, const C64* x, int64_t i, C64 yi[3]) {
  const double* ri = P.pos + 3 * i;
  C64 inv[9];
  invert3(P.alpha_uniform ? P.alpha : P.alpha + 9 * i, inv);
  yi[0] = yi[1] = yi[2] = c(0.0);
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) yi[a] = add(yi[a], mul(inv[3 * a + b], x[3 * i + b]));
  for (int64_t j = 0; j < P.n; ++j) {
    const double* rj = P.pos + 3 * j;
    bool skip_direct = (i == j) && (P.lattice == nullptr);
    if (!skip_direct) {
      T33 g;
      if (P.lattice) {
        double r[3] = {ri[0] - rj[0], ri[1] - rj[1], ri[2] - rj[2]};
        ewald_eval(*P.lattice, r, P.kb, P.k, g);
      } else {
        direct_tensor(P, i, j, g);
      }
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) yi[a] = sub(yi[a], mul(g[a][b], x[3 * j + b]));
    }
    if (P.has_sub) {
      double rimg[3] = {rj[0], rj[1], 2.0 * P.z_interface - rj[2]};
      double diff[3] = {ri[0] - rimg[0], ri[1] - rimg[1], ri[2] - rimg[2]};
      double dsq = diff[0] * diff[0] + diff[1] * diff[1] + diff[2] * diff[2];
      if (dsq >= 1e-20) {
        T33 gi;
        if (P.lattice) ewald_eval(*P.lattice, diff, P.kb, P.k, gi);
        else greens_point(ri, rimg, P.k, gi);
        C64 m[3] = {neg(P.delta), neg(P.delta), P.delta};
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b)
            yi[a] = sub(yi[a], mul(mul(m[b], gi[a][b]), x[3 * j + b]));
      }
    }
  }
}

// Row-parallel matvec over a pool of std::threads: the restatement of the
// rayon row split at assembly.rs:224-226 (rows are independent, so the
// result does not depend on the thread count).
void matvec(const Problem& P, const C64* x, C64* y, int nthreads) {
  if (nthreads < 1) nthreads = 1;
  auto work = [&](int64_t lo, int64_t hi) {
    for (int64_t i = lo; i < hi; ++i) {
      C64 yi[3];
      matvec_row(P, x, i, yi);
      y[3 * i] = yi[0];
      y[3 * i + 1] = yi[1];
      y[3 * i + 2] = yi[2];
    }
  };
  if (nthreads == 1 || P.n < 64) {
    work(0, P.n);
    return;
  }
  // interleaved chunks so each thread gets a mix of cheap/expensive rows
  std::vector<std::thread> th;
  const int64_t chunk = 16;
  std::atomic_int64_t counter(0);
  for (int t = 0; t < nthreads; ++t)
    th.emplace_back([&]() {
      for (;;) {
        int64_t lo = counter.fetch_add(chunk);
        if (lo >= P.n) break;
        work(lo, std::min(P.n, lo + chunk));
      }
    });
  for (auto& t : th) t.join();
}

// Vector kernels and Givens rotations of GMRES (iterative.rs:163-206).
// Pinned by: the GMRES known systems and LAPACK solutions below.
// iterative.rs:163-168 vector_dot: sum(conj(a_i) * b_i), sequential.
C64 vdot(const C64* a, const C64* b, int64_t n) {
  C64 s = c(0.0);
  for (int64_t i = 0; i < n; ++i) s = add(s, mul(conj(a[i]), b[i]));
  return s;
}
// iterative.rs:171-173 vector_norm.
double vnorm(const C64* v, int64_t n) {
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) s += norm_sqr(v[i]);
  return std::sqrt(s);
}
// iterative.rs:180-191 givens_rotation.
void givens(C64 a, C64 b, double* cs, C64* sn) {
  if (norm(b) < 1e-30) { *cs = 1.0; *sn = c(0.0); return; }
  if (norm(a) < 1e-30) { *cs = 0.0; *sn = cdiv(b, c(norm(b))); return; }
  double r = std::sqrt(norm_sqr(a) + norm_sqr(b));
  *cs = norm(a) / r;
  *sn = cdiv(mul(cdiv(a, c(norm(a))), conj(b)), c(r));
}

struct GmresInfo {
  int64_t iterations;   // Arnoldi steps (total_iters)
  int64_t cycles;       // restart cycles entered
  int64_t matvecs;      // matvec calls made
  double residual;      // final ||rhs - A x|| / ||rhs||
  int status;           // 0 ok, 5 convergence failure
};

// iterative.rs:33-160 solve_gmres with restart GMRES(min(30, n)), MGS
// Arnoldi, complex Givens rotations, explicit residual per cycle.
// Pinned by: iterative.rs:219-323 known systems, LAPACK on the assembled
// matrix (tests/test_oracle.py).
template <class MV>
int gmres(MV&& mv, const C64* rhs, int64_t n, double tol, int64_t max_iter, C64* xout,
          GmresInfo* info) {
  const int64_t m = std::min<int64_t>(30, n);
  std::vector<C64> x(n, c(0.0)), r(n), w(n), Ax(n);
  info->iterations = info->cycles = info->matvecs = 0;
  double rhs_norm = vnorm(rhs, n);
  if (rhs_norm < 1e-30) {
    std::copy(x.begin(), x.end(), xout);
    info->residual = 0.0;
    info->status = 0;
    return 0;
  }
  double abs_tol = tol * rhs_norm;
  int64_t total = 0;
  std::vector<std::vector<C64>> V;
  std::vector<C64> H((m + 1) * m), g(m + 1), sn;
  std::vector<double> cs;
  auto Hat = [&](int64_t i, int64_t j) -> C64& { return H[i * m + j]; };
  for (;;) {
    mv(x.data(), Ax.data());
    info->matvecs++;
    for (int64_t t = 0; t < n; ++t) r[t] = sub(rhs[t], Ax[t]);
    double beta = vnorm(r.data(), n);
    if (beta < abs_tol) {
      std::copy(x.begin(), x.end(), xout);
      info->iterations = total;
      info->residual = beta / rhs_norm;
      info->status = 0;
      return 0;
    }
    if (total >= max_iter) {
      std::copy(x.begin(), x.end(), xout);
      info->iterations = total;
      info->residual = beta / rhs_norm;
      info->status = 5;
      return 5;
    }
    info->cycles++;
    V.assign(1, std::vector<C64>(n));
    for (int64_t t = 0; t < n; ++t) V[0][t] = cdiv(r[t], c(beta));
    std::fill(H.begin(), H.end(), c(0.0));
    cs.clear();
    sn.clear();
    std::fill(g.begin(), g.end(), c(0.0));
    g[0] = c(beta);
    int64_t j = 0;
    while (j < m && total < max_iter) {
      mv(V[j].data(), w.data());
      info->matvecs++;
      for (int64_t i = 0; i <= j; ++i) {
        C64 hij = vdot(V[i].data(), w.data(), n);
        Hat(i, j) = hij;
        for (int64_t t = 0; t < n; ++t) w[t] = sub(w[t], mul(V[i][t], hij));
      }
      double hn = vnorm(w.data(), n);
      Hat(j + 1, j) = c(hn);
      V.emplace_back(n);
      if (hn > 1e-30) {
        for (int64_t t = 0; t < n; ++t) V[j + 1][t] = cdiv(w[t], c(hn));
      } else {
        std::fill(V[j + 1].begin(), V[j + 1].end(), c(0.0));
      }
      for (int64_t i = 0; i < j; ++i) {
        C64 temp = add(muls(Hat(i, j), cs[i]), mul(sn[i], Hat(i + 1, j)));
        Hat(i + 1, j) = add(mul(neg(conj(sn[i])), Hat(i, j)), mul(c(cs[i]), Hat(i + 1, j)));
        Hat(i, j) = temp;
      }
      double cc;
      C64 ss;
      givens(Hat(j, j), Hat(j + 1, j), &cc, &ss);
      cs.push_back(cc);
      sn.push_back(ss);
      C64 temp = add(mul(c(cc), Hat(j, j)), mul(ss, Hat(j + 1, j)));
      Hat(j + 1, j) = c(0.0);
      Hat(j, j) = temp;
      temp = add(mul(c(cc), g[j]), mul(ss, g[j + 1]));
      g[j + 1] = add(mul(neg(conj(ss)), g[j]), mul(c(cc), g[j + 1]));
      g[j] = temp;
      total++;
      j++;
      if (norm(g[j]) < abs_tol) break;
    }
    // iterative.rs:194-206 back_substitute
    std::vector<C64> y(j, c(0.0));
    for (int64_t i = j - 1; i >= 0; --i) {
      C64 s = g[i];
      for (int64_t l = i + 1; l < j; ++l) s = sub(s, mul(Hat(i, l), y[l]));
      if (norm(Hat(i, i)) > 1e-30) y[i] = cdiv(s, Hat(i, i));
    }
    for (int64_t i = 0; i < j; ++i)
      for (int64_t t = 0; t < n; ++t) x[t] = add(x[t], mul(V[i][t], y[i]));
    mv(x.data(), Ax.data());
    info->matvecs++;
    for (int64_t t = 0; t < n; ++t) r[t] = sub(rhs[t], Ax[t]);
    double res = vnorm(r.data(), n);
    if (res < abs_tol) {
      std::copy(x.begin(), x.end(), xout);
      info->iterations = total;
      info->residual = res / rhs_norm;
      info->status = 0;
      return 0;
    }
    if (total >= max_iter) {
      std::copy(x.begin(), x.end(), xout);
      info->iterations = total;
      info->residual = res / rhs_norm;
      info->status = 5;
      return 5;
    }
  }
}

}  // namespace

extern "C" {

// ---- building blocks (each follows the function named in its body) ----

void or_greens_point(const double* r1, const double* r2, double k, double* out18) {
  T33 g;
  greens_point(r1, r2, k, g);
  std::memcpy(out18, g, sizeof(g));
}

void or_greens_filtered(const double* ro, const double* rs, double k, double cell,
                        double* out18) {
  T33 g;
  greens_filtered(ro, rs, k, cell, g);
  std::memcpy(out18, g, sizeof(g));
}

double or_erfc_real(double x) { return erfc_real(x); }

void or_csqrt_pos_imag(const double* z, double* out) {
  C64 s = csqrt_pos_imag(c(z[0], z[1]));
  out[0] = s.re;
  out[1] = s.im;
}

static Lattice2D make_lat(const double* a1, const double* a2, double eta, int nr, int nq) {
  Lattice2D L;
  for (int t = 0; t < 3; ++t) { L.a1[t] = a1[t]; L.a2[t] = a2[t]; }
  L.eta = eta;
  L.n_real = nr;
  L.n_recip = nq;
  return L;
}

// which: 0 = evaluate, 1 = real_space, 2 = recip_space
void or_ewald(const double* a1, const double* a2, double eta, int n_real, int n_recip,
              const double* r, const double* kb, double k, int which, double* out18) {
  Lattice2D L = make_lat(a1, a2, eta, n_real, n_recip);
  T33 g;
  if (which == 1) ewald_real(L, r, kb, k, g);
  else if (which == 2) ewald_recip(L, r, kb, k, g);
  else ewald_eval(L, r, kb, k, g);
  std::memcpy(out18, g, sizeof(g));
}

double or_lattice_eta(const double* a1, const double* a2, double eta) {
  Lattice2D L = make_lat(a1, a2, eta, 5, 5);
  return lattice_eta(L);
}

void or_recip_vectors(const double* a1, const double* a2, double* b1, double* b2) {
  Lattice2D L = make_lat(a1, a2, 0.0, 5, 5);
  recip_vectors(L, b1, b2);
}

void or_invert3(const double* m18, double* out18) {
  C64 m[9], o[9];
  std::memcpy(m, m18, sizeof(m));
  invert3(m, o);
  std::memcpy(out18, o, sizeof(o));
}

// substrate.rs:181-185 fresnel_reflection_amplitude and :160-163
// fresnel_delta_eps (substrate_reflection_factor :191-197).
// Pinned by: substrate.rs:204-276 known answers (0.2, 1.25/3.25, 0).
void or_substrate_factor(const double* eps_sub, double eps_env, int retarded, double* out) {
  C64 es = c(eps_sub[0], eps_sub[1]);
  C64 r;
  if (retarded) {
    C64 ns = csqrt(es);
    C64 ne = c(std::sqrt(eps_env));
    r = cdiv(sub(ns, ne), add(ns, ne));
  } else {
    C64 env = c(eps_env);
    r = cdiv(sub(es, env), add(es, env));
  }
  out[0] = r.re;
  out[1] = r.im;
}

struct OrSpec {
  int64_t n;
  const double* pos;
  const double* alpha;  // complex interleaved
  int alpha_uniform;
  double k;
  int use_fcd;
  double cell_size;
  int has_lattice;
  double a1[3], a2[3];
  double eta;
  int n_real, n_recip;
  double kb[3];
  int has_sub;
  double z_interface;
  double delta[2];
  const double* cls_pos;   // FCD classification positions or null
};

static Problem to_problem(const OrSpec* s, Lattice2D* L) {
  Problem P;
  P.n = s->n;
  P.pos = s->pos;
  P.alpha = reinterpret_cast<const C64*>(s->alpha);
  P.alpha_uniform = s->alpha_uniform;
  P.k = s->k;
  P.use_fcd = s->use_fcd;
  P.cell_size = s->cell_size;
  if (s->has_lattice) {
    *L = make_lat(s->a1, s->a2, s->eta, s->n_real, s->n_recip);
    P.lattice = L;
  } else {
    P.lattice = nullptr;
  }
  for (int t = 0; t < 3; ++t) P.kb[t] = s->kb[t];
  P.has_sub = s->has_sub;
  P.z_interface = s->z_interface;
  P.delta = c(s->delta[0], s->delta[1]);
  P.cls_pos = s->cls_pos;
  return P;
}

// assembly.rs:209-318 matvec_on_the_fly
void or_matvec(const OrSpec* spec, const double* x, double* y, int nthreads) {
  Lattice2D L;
  Problem P = to_problem(spec, &L);
  matvec(P, reinterpret_cast<const C64*>(x), reinterpret_cast<C64*>(y), nthreads);
}

// Rows [row_lo, row_hi) only (the sampled-row CPU baseline and row-share checks).
void or_matvec_rows(const OrSpec* spec, const double* x, double* y, int64_t row_lo,
                    int64_t row_hi) {
  Lattice2D L;
  Problem P = to_problem(spec, &L);
  for (int64_t i = row_lo; i < row_hi; ++i) {
    C64 yi[3];
    matvec_row(P, reinterpret_cast<const C64*>(x), i, yi);
    std::memcpy(y + 6 * (i - row_lo), yi, sizeof(yi));
  }
}

// assembly.rs:45-175 assemble_interaction_matrix (dense, row-major 3N x 3N);
// used with LAPACK as an independent check of the on-the-fly matvec + GMRES.
// Pinned by: a 4-dipole matrix built in plain Python from the golden-checked
// G (tests/test_oracle.py), and A x equal to the on-the-fly matvec.
void or_assemble(const OrSpec* spec, double* out) {
  Lattice2D L;
  Problem P = to_problem(spec, &L);
  int64_t n = P.n, dim = 3 * n;
  C64* A = reinterpret_cast<C64*>(out);
  std::fill(A, A + dim * dim, c(0.0));
  for (int64_t i = 0; i < n; ++i) {
    C64 inv[9];
    invert3(P.alpha_uniform ? P.alpha : P.alpha + 9 * i, inv);
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) A[(3 * i + a) * dim + 3 * i + b] = inv[3 * a + b];
  }
  for (int64_t i = 0; i < n; ++i)
    for (int64_t j = 0; j < n; ++j) {
      const double* ri = P.pos + 3 * i;
      const double* rj = P.pos + 3 * j;
      if (!((i == j) && P.lattice == nullptr)) {
        T33 g;
        if (P.lattice) {
          double r[3] = {ri[0] - rj[0], ri[1] - rj[1], ri[2] - rj[2]};
          ewald_eval(*P.lattice, r, P.kb, P.k, g);
        } else {
          direct_tensor(P, i, j, g);
        }
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) {
            C64& e = A[(3 * i + a) * dim + 3 * j + b];
            e = sub(e, g[a][b]);
          }
      }
      if (P.has_sub) {
        double rimg[3] = {rj[0], rj[1], 2.0 * P.z_interface - rj[2]};
        double diff[3] = {ri[0] - rimg[0], ri[1] - rimg[1], ri[2] - rimg[2]};
        double dsq = diff[0] * diff[0] + diff[1] * diff[1] + diff[2] * diff[2];
        if (dsq >= 1e-20) {
          T33 gi;
          if (P.lattice) ewald_eval(*P.lattice, diff, P.kb, P.k, gi);
          else greens_point(ri, rimg, P.k, gi);
          C64 m[3] = {neg(P.delta), neg(P.delta), P.delta};
          for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
              C64& e = A[(3 * i + a) * dim + 3 * j + b];
              e = sub(e, mul(m[b], gi[a][b]));
            }
        }
      }
    }
}

// iterative.rs:33-160 on the on-the-fly matvec (cda/mod.rs:329-345).
// info_out: [iterations, cycles, matvecs, residual, status] as doubles.
int or_solve_gmres(const OrSpec* spec, const double* rhs, double tol, int64_t max_iter,
                   double* x, double* info_out, int nthreads) {
  Lattice2D L;
  Problem P = to_problem(spec, &L);
  GmresInfo info;
  auto mv = [&](const C64* in, C64* out) { matvec(P, in, out, nthreads); };
  int st = gmres(mv, reinterpret_cast<const C64*>(rhs), 3 * P.n, tol, max_iter,
                 reinterpret_cast<C64*>(x), &info);
  info_out[0] = (double)info.iterations;
  info_out[1] = (double)info.cycles;
  info_out[2] = (double)info.matvecs;
  info_out[3] = info.residual;
  info_out[4] = (double)info.status;
  return st;
}

// GMRES on an explicit dense matrix (iterative.rs tests :211-323 use
// matrix.dot as the closure).
int or_solve_gmres_dense(const double* Adense, int64_t n, const double* rhs, double tol,
                         int64_t max_iter, double* x, double* info_out) {
  const C64* A = reinterpret_cast<const C64*>(Adense);
  GmresInfo info;
  auto mv = [&](const C64* in, C64* out) {
    for (int64_t i = 0; i < n; ++i) {
      C64 s = c(0.0);
      for (int64_t j = 0; j < n; ++j) s = add(s, mul(A[i * n + j], in[j]));
      out[i] = s;
    }
  };
  int st = gmres(mv, reinterpret_cast<const C64*>(rhs), n, tol, max_iter,
                 reinterpret_cast<C64*>(x), &info);
  info_out[0] = (double)info.iterations;
  info_out[1] = (double)info.cycles;
  info_out[2] = (double)info.matvecs;
  info_out[3] = info.residual;
  info_out[4] = (double)info.status;
  return st;
}

// types.rs:60-74 IncidentField::at_position for every dipole
// Pinned by: the numpy plane wave at an oblique direction with a complex
// polarisation (tests/test_oracle.py).
// (assembly.rs:178-194 build_incident_field_vector). The polarisation is
// complex here so that circular polarisation (no reference counterpart)
// uses the same routine; a real polarisation gives the reference's values.
void or_incident(int64_t n, const double* pos, const double* dir, const double* pol6,
                 double amplitude, double k, double* out) {
  C64* E = reinterpret_cast<C64*>(out);
  for (int64_t i = 0; i < n; ++i) {
    const double* p = pos + 3 * i;
    double kdotr = k * (dir[0] * p[0] + dir[1] * p[1] + dir[2] * p[2]);
    C64 phase = cexp(c(0.0, kdotr));
    C64 e0 = c(amplitude);
    for (int cc = 0; cc < 3; ++cc) {
      C64 pol = c(pol6[2 * cc], pol6[2 * cc + 1]);
      C64 v = (pol.im == 0.0) ? mul(muls(e0, pol.re), phase) : mul(mul(e0, pol), phase);
      E[3 * i + cc] = v;
    }
  }
}

// cda/mod.rs:187-235 C_ext / C_abs / C_sca.
// Pinned by: the single dipole's C_ext = k Im(alpha) (cda_diagnostic.rs:
// 119-167) with C_abs + C_sca = C_ext, and energy balance on a particle
// (tests/test_oracle.py).
void or_cross_sections(int64_t n, const double* alpha, int alpha_uniform, const double* p,
                       const double* einc, double k, double amplitude, double* out3) {
  const C64* A = reinterpret_cast<const C64*>(alpha);
  const C64* P = reinterpret_cast<const C64*>(p);
  const C64* E = reinterpret_cast<const C64*>(einc);
  double e0_sq = amplitude * amplitude;
  double cext = 0.0;
  for (int64_t i = 0; i < n; ++i)
    for (int cc = 0; cc < 3; ++cc) cext += mul(conj(E[3 * i + cc]), P[3 * i + cc]).im;
  cext *= k / e0_sq;
  double cabs = 0.0;
  double k3 = k * k * k;
  for (int64_t i = 0; i < n; ++i) {
    C64 inv[9];
    invert3(alpha_uniform ? A : A + 9 * i, inv);
    C64 s = c(0.0);
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        s = add(s, mul(mul(P[3 * i + a], conj(inv[3 * a + b])), conj(P[3 * i + b])));
    double psq = 0.0;
    for (int cc = 0; cc < 3; ++cc) psq += norm_sqr(P[3 * i + cc]);
    cabs += s.im - (2.0 / 3.0) * k3 * psq;
  }
  cabs *= k / e0_sq;
  out3[0] = cext;
  out3[1] = cabs;
  out3[2] = std::max(cext - cabs, 0.0);
}

// Adjoint gradients of C_ext = (k/E0^2) Im(E^H p), A p = E, A symmetric,
// A lam = conj(E) (no reference counterpart):
//   grad_eps[i] = (dC/dRe eps_i, dC/dIm eps_i) = (k/E0^2) (Im f, Re f),
//     f = -lam_i^T D_i p_i, D_i = d(alpha_i^-1)/d eps_i (caller's model);
//   grad_r[i][c] = (k/E0^2) Im( -i k dir_c conj(E_i).p_i + i k dir_c lam_i.E_i
//     + sum_{j != i} [lam_i^T dG_c(i,j) p_j + p_i^T dG_c(i,j) lam_j] ),
//   dG_c(i,j) = d G(r_i, r_j) / d r_i,c (FCD classification frozen as in
//   direct_tensor). Positions need free space: returns 1 (grad_r zero) with
//   a lattice or a substrate.
int or_adjoint_grad(const OrSpec* spec, const double* dir, const double* pol6, double amplitude,
                    const double* p, const double* lam, const double* dainv, int dainv_uniform,
                    double* grad_eps, double* grad_r, const int64_t* rows, int64_t nrows) {
  Lattice2D L;
  Problem P = to_problem(spec, &L);
  const int64_t n = P.n;
  const C64* Pv = reinterpret_cast<const C64*>(p);
  const C64* Lv = reinterpret_cast<const C64*>(lam);
  const C64* D = reinterpret_cast<const C64*>(dainv);
  const double K = P.k / (amplitude * amplitude);
  std::vector<C64> E(3 * n);
  or_incident(n, P.pos, dir, pol6, amplitude, P.k, reinterpret_cast<double*>(E.data()));
  for (int64_t i = 0; i < n; ++i) {
    const C64* Di = dainv_uniform ? D : D + 9 * i;
    C64 f = c(0.0);
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) f = add(f, mul(mul(Lv[3 * i + a], Di[3 * a + b]), Pv[3 * i + b]));
    f = neg(f);
    grad_eps[2 * i] = K * f.im;
    grad_eps[2 * i + 1] = K * f.re;
  }
  std::fill(grad_r, grad_r + 3 * n, 0.0);
  if (P.lattice || P.has_sub) return 1;
  // rows: the dipoles whose position gradient is wanted (all when null)
  const int64_t nr = rows ? nrows : n;
  for (int64_t q = 0; q < nr; ++q) {
    const int64_t i = rows ? rows[q] : q;
    C64 t1 = c(0.0), t2 = c(0.0);
    for (int a = 0; a < 3; ++a) {
      t1 = add(t1, mul(conj(E[3 * i + a]), Pv[3 * i + a]));
      t2 = add(t2, mul(Lv[3 * i + a], E[3 * i + a]));
    }
    C64 g[3];
    for (int cc = 0; cc < 3; ++cc)
      g[cc] = add(mul(c(0.0, -P.k * dir[cc]), t1), mul(c(0.0, P.k * dir[cc]), t2));
    for (int64_t j = 0; j < n; ++j) {
      if (j == i) continue;
      T333 dg;
      direct_tensor_grad(P, i, j, dg);
      for (int cc = 0; cc < 3; ++cc)
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) {
            g[cc] = add(g[cc], mul(mul(Lv[3 * i + a], dg[cc][a][b]), Pv[3 * j + b]));
            g[cc] = add(g[cc], mul(mul(Pv[3 * i + a], dg[cc][a][b]), Lv[3 * j + b]));
          }
    }
    for (int cc = 0; cc < 3; ++cc) grad_r[3 * i + cc] = K * g[cc].im;
  }
  return 0;
}

// fields.rs:18-48 field_at_point at m observation points: the incident
// field (types.rs:60-74) plus sum_i G(r_obs, r_i) p_i, dipoles with
// |r_obs - r_i|^2 < 1e-20 skipped; also |E|^2 (fields.rs:87).
// Pinned by: the scattered field at k r = 1e4 against k^2 e^{ikr}/(4 pi r)
// (I - rhat rhat) F, the incident plane wave alone for zero moments, and the
// skip of a coincident dipole (tests/test_oracle.py).
void or_field_at_points(int64_t n, const double* pos, const double* moments, double k,
                        const double* dir, const double* pol6, double amplitude, int64_t m,
                        const double* points, double* field, double* intensity) {
  const C64* P = reinterpret_cast<const C64*>(moments);
  C64* F = reinterpret_cast<C64*>(field);
  for (int64_t q = 0; q < m; ++q) {
    const double* o = points + 3 * q;
    C64 e[3];
    or_incident(1, o, dir, pol6, amplitude, k, reinterpret_cast<double*>(e));
    for (int64_t i = 0; i < n; ++i) {
      const double* r = pos + 3 * i;
      double dx = o[0] - r[0], dy = o[1] - r[1], dz = o[2] - r[2];
      double r_sq = dx * dx + dy * dy + dz * dz;
      if (r_sq < 1e-20) continue;
      T33 g;
      greens_point(o, r, k, g);
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) e[a] = add(e[a], mul(g[a][b], P[3 * i + b]));
    }
    for (int a = 0; a < 3; ++a) F[3 * q + a] = e[a];
    intensity[q] = norm_sqr(e[0]) + norm_sqr(e[1]) + norm_sqr(e[2]);
  }
}

// fields.rs:156-221 compute_far_field: theta = pi it / max(n_theta - 1, 1),
// phi = 2 pi ip / n_phi, F = sum_i p_i e^{-ik rhat.r_i}, intensity
// |(I - rhat rhat) F|^2.
// Pinned by: one dipole's |(I - rhat rhat) p|^2 (phase-free), two equal
// dipoles L apart on z giving that times |1 + e^{-ikL cos theta}|^2, and the
// theta/phi grid (tests/test_oracle.py).
void or_far_field(int64_t n, const double* pos, const double* moments, double k, int64_t n_theta,
                  int64_t n_phi, double* theta, double* phi, double* intensity) {
  const C64* P = reinterpret_cast<const C64*>(moments);
  int64_t t = 0;
  for (int64_t it = 0; it < n_theta; ++it) {
    double th = PI * (double)it / (double)std::max<int64_t>(n_theta - 1, 1);
    double st = std::sin(th), ct = std::cos(th);
    for (int64_t ip = 0; ip < n_phi; ++ip, ++t) {
      double ph = 2.0 * PI * (double)ip / (double)n_phi;
      double sp = std::sin(ph), cp = std::cos(ph);
      double rh[3] = {st * cp, st * sp, ct};
      C64 f[3] = {c(0.0), c(0.0), c(0.0)};
      for (int64_t i = 0; i < n; ++i) {
        const double* r = pos + 3 * i;
        double kdotr = k * (rh[0] * r[0] + rh[1] * r[1] + rh[2] * r[2]);
        C64 phase = cexp(c(0.0, -kdotr));
        for (int a = 0; a < 3; ++a) f[a] = add(f[a], mul(P[3 * i + a], phase));
      }
      C64 rdotf = add(add(muls(f[0], rh[0]), muls(f[1], rh[1])), muls(f[2], rh[2]));
      double in = 0.0;
      for (int a = 0; a < 3; ++a) in += norm_sqr(sub(f[a], muls(rdotf, rh[a])));
      theta[t] = th;
      phi[t] = ph;
      intensity[t] = in;
    }
  }
}

// Point-tensor gradient d G_ab / d (r_o)_c as [c][a][b] complex.
void or_greens_point_grad(const double* ro, const double* rs, double k, double* out) {
  T333 g;
  greens_point_grad(ro, rs, k, g);
  std::memcpy(out, g, sizeof(g));
}

}  // extern "C"
