// Host-side geometry preparation: near list, grid detection, Ewald records.
#include "lh_host.h"

#include <algorithm>
#include <array>
#include <map>
#include <cmath>
#include <unordered_map>

#include "lh_math.h"

namespace lh {

// ---------------------------------------------------------------------------
// Near list: spatial hash with bins of side sqrt(rskip2f); every pair within
// the exclusion radius is found by scanning the 27 neighbouring bins. The
// membership test is r2f_exact on FP32-rounded positions, the expression the
// far-field kernels skip on.
// ---------------------------------------------------------------------------
namespace {
struct DispKey {
  uint64_t a, b, c;
  int fcd;
  bool operator==(const DispKey& o) const {
    return a == o.a && b == o.b && c == o.c && fcd == o.fcd;
  }
};
struct DispHash {
  size_t operator()(const DispKey& k) const {
    return (size_t)(k.a * 0x9E3779B97F4A7C15ULL ^ (k.b + 0x7F4A7C15ULL) * 0xC2B2AE3D27D4EB4FULL ^
                    (k.c + 0x165667B1ULL) * 0x27D4EB2F165667C5ULL ^ (uint64_t)k.fcd);
  }
};
}  // namespace

void build_near_list(Geometry& g) {
  const int n = g.n;
  g.kw = 0;
  g.ell_col.clear();
  g.ell_tid.clear();
  g.disp.clear();
  g.ntab = 0;
  g.nnz = 0;
  if (!(g.rskip2f > 0.0f) || n == 0) return;
  std::vector<float> xf(n), yf(n), zf(n);
  for (int i = 0; i < n; ++i) {
    xf[i] = (float)g.x[i];
    yf[i] = (float)g.y[i];
    zf[i] = (float)g.z[i];
  }
  double mn[3] = {g.x[0], g.y[0], g.z[0]};
  double amax = 0.0;
  for (int i = 0; i < n; ++i) {
    mn[0] = std::min(mn[0], g.x[i]);
    mn[1] = std::min(mn[1], g.y[i]);
    mn[2] = std::min(mn[2], g.z[i]);
    amax = std::max({amax, std::fabs(g.x[i]), std::fabs(g.y[i]), std::fabs(g.z[i])});
  }
  // bin side: the exclusion radius plus the largest change FP32 rounding of
  // the positions can make to a displacement
  const double h = std::sqrt((double)g.rskip2f) * (1.0 + 1e-6) + 4.0 * amax * 1.2e-7 + 1e-12;
  auto key = [](long long a, long long b, long long c) {
    return (a * 73856093LL) ^ (b * 19349663LL) ^ (c * 83492791LL);
  };
  std::vector<long long> bx(n), by(n), bz(n);
  std::unordered_map<long long, std::vector<int>> bins;
  bins.reserve(n);
  for (int i = 0; i < n; ++i) {
    bx[i] = (long long)std::floor((g.x[i] - mn[0]) / h);
    by[i] = (long long)std::floor((g.y[i] - mn[1]) / h);
    bz[i] = (long long)std::floor((g.z[i] - mn[2]) / h);
    bins[key(bx[i], by[i], bz[i])].push_back(i);
  }
  std::vector<std::vector<int>> rows(n);
  for (int i = 0; i < n; ++i) {
    auto& cand = rows[i];
    for (int a = -1; a <= 1; ++a)
      for (int b = -1; b <= 1; ++b)
        for (int c = -1; c <= 1; ++c) {
          auto it = bins.find(key(bx[i] + a, by[i] + b, bz[i] + c));
          if (it == bins.end()) continue;
          for (int j : it->second) {
            if (j == i) continue;
            if (bx[j] != bx[i] + a || by[j] != by[i] + b || bz[j] != bz[i] + c) continue;
            if (r2f_exact(xf[i] - xf[j], yf[i] - yf[j], zf[i] - zf[j]) <= g.rskip2f)
              cand.push_back(j);
          }
        }
    std::sort(cand.begin(), cand.end());
    g.kw = std::max<int>(g.kw, (int)cand.size());
    g.nnz += (long long)cand.size();
  }
  g.ell_col.assign((size_t)g.kw * n, 0);
  g.ell_tid.assign((size_t)g.kw * n, -1);
  std::unordered_map<DispKey, int, DispHash> tab;
  for (int i = 0; i < n; ++i) {
    for (size_t kk = 0; kk < rows[i].size(); ++kk) {
      const int j = rows[i][kk];
      const double dx = g.x[i] - g.x[j], dy = g.y[i] - g.y[j], dz = g.z[i] - g.z[j];
      // greens.rs:102-106: dist = sqrt(dx^2 + dy^2 + dz^2); FCD unless dist > 2 cell,
      // on the classification positions when they are given
      double dist;
      if (g.cls.empty()) {
        dist = std::sqrt(r2_exact(dx, dy, dz));
      } else {
        const double* ci = &g.cls[3 * (size_t)i];
        const double* cj = &g.cls[3 * (size_t)j];
        dist = std::sqrt(r2_exact(ci[0] - cj[0], ci[1] - cj[1], ci[2] - cj[2]));
      }
      const bool fcd = g.use_fcd && !(dist > 2.0 * g.cell);
      DispKey dk{__builtin_bit_cast(uint64_t, dx), __builtin_bit_cast(uint64_t, dy),
                 __builtin_bit_cast(uint64_t, dz), fcd ? 1 : 0};
      auto it = tab.find(dk);
      int u;
      if (it == tab.end()) {
        u = g.ntab++;
        tab.emplace(dk, u);
        g.disp.push_back(dx);
        g.disp.push_back(dy);
        g.disp.push_back(dz);
        g.disp.push_back(fcd ? 1.0 : 0.0);
      } else {
        u = it->second;
      }
      g.ell_col[kk * (size_t)n + i] = j;
      g.ell_tid[kk * (size_t)n + i] = u;
    }
  }
}

// ---------------------------------------------------------------------------
// Lattice solves: grid detection and Ewald records.
// ---------------------------------------------------------------------------

void detect_grid(Geometry& g) {
  g.grid = 0;
  g.gpack.clear();
  g.gizj2.clear();
  if (g.n == 0) return;
  const std::vector<double>* ax[3] = {&g.x, &g.y, &g.z};
  std::vector<long long> idx[3];
  int gn[3];
  double g0[3], gh[3];
  for (int a = 0; a < 3; ++a) {
    const std::vector<double>& v = *ax[a];
    const double lo = *std::min_element(v.begin(), v.end());
    const double hi = *std::max_element(v.begin(), v.end());
    // each coordinate must sit within 1e-12 of the span scale of its grid
    // line, so that a grid offset times h is the pair's displacement to
    // about 1e-12 relative
    const double tol = 1e-12 * std::max({1.0, std::fabs(lo), std::fabs(hi)});
    std::vector<double> u(v);
    std::sort(u.begin(), u.end());
    double h = 0.0;
    for (size_t q = 1; q < u.size(); ++q) {
      const double d = u[q] - u[q - 1];
      if (d > tol && (h == 0.0 || d < h)) h = d;
    }
    idx[a].assign(g.n, 0);
    if (h == 0.0) {
      gn[a] = 1;
      g0[a] = lo;
      gh[a] = 1.0;
      continue;
    }
    const long long m = std::llround((hi - lo) / h);
    if (m <= 0 || m > (1 << 20)) return;
    h = (hi - lo) / (double)m;
    for (int i = 0; i < g.n; ++i) {
      const long long it = std::llround((v[i] - lo) / h);
      if (std::fabs(v[i] - (lo + (double)it * h)) > tol) return;
      idx[a][i] = it;
    }
    gn[a] = (int)m + 1;
    g0[a] = lo;
    gh[a] = h;
  }
  const long long nxd = 2LL * gn[0] - 1, nyd = 2LL * gn[1] - 1, nzd = 2LL * gn[2] - 1;
  if (nxd * nyd * nzd > (1LL << 30)) return;
  g.gpack.resize(g.n);
  g.gizj2.resize(g.n);
  for (int i = 0; i < g.n; ++i) {
    g.gpack[i] = (int)(idx[0][i] * nyd * nzd + idx[1][i] * nzd + idx[2][i]);
    g.gizj2[i] = (int)(2 * idx[2][i] - (gn[2] - 1));
  }
  for (int a = 0; a < 3; ++a) {
    g.gn[a] = gn[a];
    g.g0[a] = g0[a];
    g.gh[a] = gh[a];
  }
  g.grid = 1;
}

bool full_grid_indices(const Geometry& g, std::vector<int>& idx) {
  idx.clear();
  if (!g.grid || g.n < 2) return false;
  const long long cells = (long long)g.gn[0] * g.gn[1] * g.gn[2];
  if (cells != (long long)g.n) return false;
  std::vector<char> seen((size_t)cells, 0);
  idx.resize(3 * (size_t)g.n);
  const std::vector<double>* ax[3] = {&g.x, &g.y, &g.z};
  for (int i = 0; i < g.n; ++i) {
    long long flat = 0;
    for (int a = 0; a < 3; ++a) {
      const long long it =
          g.gn[a] == 1 ? 0 : std::llround(((*ax[a])[i] - g.g0[a]) / g.gh[a]);
      if (it < 0 || it >= g.gn[a]) return false;
      idx[3 * (size_t)i + a] = (int)it;
      flat = flat * g.gn[a] + it;
    }
    if (seen[(size_t)flat]) return false;
    seen[(size_t)flat] = 1;
  }
  return true;
}

bool fcd_offsets(const Geometry& g, const std::vector<int>& idx, std::vector<int>& offs) {
  offs.clear();
  std::map<std::array<int, 3>, int> flag;
  const int n = g.n;
  for (int kk = 0; kk < g.kw; ++kk)
    for (int i = 0; i < n; ++i) {
      const int u = g.ell_tid[(size_t)kk * n + i];
      if (u < 0) continue;
      const int j = g.ell_col[(size_t)kk * n + i];
      const std::array<int, 3> o = {idx[3 * (size_t)i] - idx[3 * (size_t)j],
                                    idx[3 * (size_t)i + 1] - idx[3 * (size_t)j + 1],
                                    idx[3 * (size_t)i + 2] - idx[3 * (size_t)j + 2]};
      const int f = g.disp[4 * (size_t)u + 3] != 0.0 ? 1 : 0;
      auto it = flag.find(o);
      if (it == flag.end()) flag.emplace(o, f);
      else if (it->second != f) return false;
    }
  for (const auto& e : flag)
    if (e.second)
      for (int c = 0; c < 3; ++c) offs.push_back(e.first[c]);
  return true;
}

void natural_spline(const double* xs, const double* ys, int n, double* y2) {
  std::vector<double> u((size_t)std::max(n - 1, 1), 0.0);
  for (int i = 0; i < n; ++i) y2[i] = 0.0;
  for (int i = 1; i < n - 1; ++i) {
    const double sig = (xs[i] - xs[i - 1]) / (xs[i + 1] - xs[i - 1]);
    const double p = sig * y2[i - 1] + 2.0;
    y2[i] = (sig - 1.0) / p;
    const double ui = (ys[i + 1] - ys[i]) / (xs[i + 1] - xs[i]) -
                      (ys[i] - ys[i - 1]) / (xs[i] - xs[i - 1]);
    u[i] = (6.0 * ui / (xs[i + 1] - xs[i - 1]) - sig * u[i - 1]) / p;
  }
  for (int k = n - 3; k >= 0; --k) y2[k + 1] = y2[k + 1] * y2[k + 2] + u[k + 1];
}

// lattice.rs:75-147 (cell_measure, eta, reciprocal_vectors), lattice.rs:168-233
// (point order) and the r-independent factors of ewald.rs:119-273.
double max_phase_revolutions(const Geometry& g, const Problem& p) {
  if (g.n == 0) return 0.0;
  double lo[3] = {g.x[0], g.y[0], g.z[0]}, hi[3] = {g.x[0], g.y[0], g.z[0]};
  for (int i = 0; i < g.n; ++i) {
    lo[0] = std::min(lo[0], g.x[i]);
    hi[0] = std::max(hi[0], g.x[i]);
    lo[1] = std::min(lo[1], g.y[i]);
    hi[1] = std::max(hi[1], g.y[i]);
    lo[2] = std::min(lo[2], g.z[i]);
    hi[2] = std::max(hi[2], g.z[i]);
  }
  double diag = std::sqrt((hi[0] - lo[0]) * (hi[0] - lo[0]) + (hi[1] - lo[1]) * (hi[1] - lo[1]) +
                          (hi[2] - lo[2]) * (hi[2] - lo[2]));
  if (p.has_sub)
    diag = 2.0 * diag + 2.0 * std::fabs(p.z_interface) +
           2.0 * std::max(std::fabs(lo[2]), std::fabs(hi[2]));
  return p.k * diag / (2.0 * M_PI);
}

void ewald_records(const Problem& p, std::vector<double>& out, int& nreal, int& nrecip,
                   double& eta) {
  out.clear();
  const double cx = p.a1[1] * p.a2[2] - p.a1[2] * p.a2[1];
  const double cy = p.a1[2] * p.a2[0] - p.a1[0] * p.a2[2];
  const double cz = p.a1[0] * p.a2[1] - p.a1[1] * p.a2[0];
  const double area = std::sqrt(cx * cx + cy * cy + cz * cz);
  eta = (p.eta > 0.0) ? p.eta : std::sqrt(M_PI / area);
  nreal = 0;
  for (int m1 = -p.n_real; m1 <= p.n_real; ++m1)
    for (int m2 = -p.n_real; m2 <= p.n_real; ++m2) {
      double R[3];
      for (int c = 0; c < 3; ++c) R[c] = (double)m1 * p.a1[c] + (double)m2 * p.a2[c];
      const bool zero = std::fabs(R[0]) < 1e-12 && std::fabs(R[1]) < 1e-12 && std::fabs(R[2]) < 1e-12;
      const double barg = p.kb[0] * R[0] + p.kb[1] * R[1] + p.kb[2] * R[2];
      out.insert(out.end(), {R[0], R[1], R[2], zero ? 1.0 : 0.0, std::cos(barg), std::sin(barg)});
      ++nreal;
    }
  // reciprocal vectors of the in-plane lattice (lattice.rs:113-147)
  const double tp = 2.0 * M_PI;
  const double a2xz[2] = {p.a2[1], -p.a2[0]};
  const double d1 = p.a1[0] * a2xz[0] + p.a1[1] * a2xz[1];
  const double b1[2] = {tp * a2xz[0] / d1, tp * a2xz[1] / d1};
  const double zxa1[2] = {-p.a1[1], p.a1[0]};
  const double d2 = p.a2[0] * zxa1[0] + p.a2[1] * zxa1[1];
  const double b2[2] = {tp * zxa1[0] / d2, tp * zxa1[1] / d2};
  const double pre = 1.0 / (2.0 * area);   // i / (2A) = (0, pre)
  const double k2 = p.k * p.k;
  nrecip = 0;
  for (int n1 = -p.n_recip; n1 <= p.n_recip; ++n1)
    for (int n2 = -p.n_recip; n2 <= p.n_recip; ++n2) {
      const double qx = p.kb[0] + ((double)n1 * b1[0] + (double)n2 * b2[0]);
      const double qy = p.kb[1] + ((double)n1 * b1[1] + (double)n2 * b2[1]);
      const double x = (qx * qx + qy * qy) - k2;
      // csqrt of a real argument with Im >= 0 (ewald.rs:334-338)
      const double qr = x >= 0.0 ? std::sqrt(x) : 0.0;
      const double qi = x >= 0.0 ? 0.0 : std::sqrt(-x);
      if (std::sqrt(qr * qr + qi * qi) < 1e-10) continue;
      // (0, pre) / (qr, qi)
      const double den = qr * qr + qi * qi;
      const double cr = pre * qi / den, ci = pre * qr / den;
      out.insert(out.end(), {qx, qy, qr, qi, cr, ci});
      ++nrecip;
    }
}

}  // namespace lh
