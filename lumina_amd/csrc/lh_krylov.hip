// Batched restarted GMRES(30) on the device (iterative.rs:33-206).
//
// All solves of a batch advance in lock-step through one restart cycle:
// the host enqueues the same kernel sequence for every solve and each kernel
// skips the solves whose per-solve state (GState) says they do not take
// part. The Krylov basis, Hessenberg matrix, Givens rotations and the
// least-squares right-hand side live in HBM for the whole solve; the host
// only reads back a counter per step.
//
// Orthogonalisation is classical Gram-Schmidt applied twice (CGS2): each
// pass is one batched dot kernel (all <v_i, w> at once) and one update
// kernel, instead of the reference's j+1 sequential MGS passes
// (iterative.rs:92-96). CGS2 gives a basis orthogonal to working precision,
// so the Hessenberg column agrees with MGS to rounding.
//
// Vector kernels work on items of LH_VB dipoles (3 LH_VB complex values);
// reductions write one partial per item and a per-solve kernel adds the
// partials in a fixed order, so every result is reproducible.
#include "lh_device.h"

namespace lh {

__device__ __forceinline__ bool part_of(const GState* st, int s, int which) {
  return takes_part(st, s, which);
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) {
  return make_double2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ double2 csub(double2 a, double2 b) {
  return make_double2(a.x - b.x, a.y - b.y);
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double cabs_(double2 a) { return hypot(a.x, a.y); }
// quotient as (a * conj(b)) / |b|^2 (the num-complex formula)
__device__ __forceinline__ double2 cdiv(double2 a, double2 b) {
  const double nsq = b.x * b.x + b.y * b.y;
  return make_double2((a.x * b.x + a.y * b.y) / nsq, (a.y * b.x - a.x * b.y) / nsq);
}

// ---------------------------------------------------------------------------
// Vector kernels.
// ---------------------------------------------------------------------------

// norm_part[item] = sum |v_slot|^2 over the item.
__global__ __launch_bounds__(LH_VB) void k_vnorm(const VecItem* __restrict__ items,
                                                const SolveDev* __restrict__ solves,
                                                const GState* __restrict__ st, int which,
                                                const double2* __restrict__ vec, int slot,
                                                double* __restrict__ norm_part) {
  __shared__ double lds[4];
  const VecItem it = items[blockIdx.x];
  if (!part_of(st, it.s, which)) return;
  const SolveDev S = solves[it.s];
  const long long n3 = 3LL * S.n;
  const double2* v = vec + S.vec_off + (long long)slot * n3;
  double acc = 0.0;
  const long long e0 = 3LL * it.t0;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const long long e = e0 + threadIdx.x + q * LH_VB;
    if (e < n3) {
      const double2 a = v[e];
      acc += a.x * a.x + a.y * a.y;
    }
  }
  const double t = block_sum(acc, lds);
  if (threadIdx.x == 0) norm_part[S.vi_off + it.b] = t;
}

// dst = src * nscale (per-solve scale from the state). dst_slot < 0 selects
// slot j+1 of the Krylov basis (after an Arnoldi step, j from the state).
__global__ __launch_bounds__(LH_VB) void k_scale(const VecItem* __restrict__ items,
                                                const SolveDev* __restrict__ solves,
                                                const GState* __restrict__ st, int which,
                                                double2* __restrict__ vec, int src_slot,
                                                int dst_slot) {
  const VecItem it = items[blockIdx.x];
  if (!part_of(st, it.s, which)) return;
  const SolveDev S = solves[it.s];
  const GState& g = st[it.s];
  const double sc = g.nscale;
  const int ds = dst_slot >= 0 ? dst_slot : g.j;
  const long long n3 = 3LL * S.n;
  const double2* a = vec + S.vec_off + (long long)src_slot * n3;
  double2* b = vec + S.vec_off + (long long)ds * n3;
  const long long e0 = 3LL * it.t0;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const long long e = e0 + threadIdx.x + q * LH_VB;
    if (e < n3) {
      const double2 x = a[e];
      b[e] = make_double2(x.x * sc, x.y * sc);
    }
  }
}

// dot_part[item][i] = sum conj(v_i) w over the item, i = i0..j.
__global__ __launch_bounds__(LH_VB) void k_dots(const VecItem* __restrict__ items,
                                               const SolveDev* __restrict__ solves,
                                               const GState* __restrict__ st,
                                               const double2* __restrict__ vec, int i0, int j,
                                               double2* __restrict__ dot_part) {
  __shared__ double lds[2][4][LH_MMAX + 1];
  const VecItem it = items[blockIdx.x];
  if (!part_of(st, it.s, 0)) return;
  const SolveDev S = solves[it.s];
  const long long n3 = 3LL * S.n;
  const double2* V = vec + S.vec_off;
  const double2* W = V + (long long)LH_SLOT_W * n3;
  const long long e0 = 3LL * it.t0;
  double2 w[3];
  bool ok[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const long long e = e0 + threadIdx.x + q * LH_VB;
    ok[q] = e < n3;
    w[q] = ok[q] ? W[e] : make_double2(0.0, 0.0);
  }
  const int wv = threadIdx.x >> 6;
  for (int i = i0; i <= j; ++i) {
    const double2* Vi = V + (long long)i * n3;
    double sr = 0.0, si = 0.0;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      if (ok[q]) {
        const double2 a = Vi[e0 + threadIdx.x + q * LH_VB];
        // conj(a) * w
        sr += a.x * w[q].x + a.y * w[q].y;
        si += a.x * w[q].y - a.y * w[q].x;
      }
    }
    sr = wave_sum(sr);
    si = wave_sum(si);
    if ((threadIdx.x & 63) == 0) {
      lds[0][wv][i] = sr;
      lds[1][wv][i] = si;
    }
  }
  __syncthreads();
  if (threadIdx.x >= i0 && threadIdx.x <= j) {
    const int i = threadIdx.x;
    const double sr = (lds[0][0][i] + lds[0][1][i]) + (lds[0][2][i] + lds[0][3][i]);
    const double si = (lds[1][0][i] + lds[1][1][i]) + (lds[1][2][i] + lds[1][3][i]);
    dot_part[(long long)(S.vi_off + it.b) * LH_NDOT + i] = make_double2(sr, si);
  }
}

// w -= sum_{i0<=i<=j} hp_i v_i ; norm_part[item] = |w|^2 over the item.
__global__ __launch_bounds__(LH_VB) void k_update(const VecItem* __restrict__ items,
                                                 const SolveDev* __restrict__ solves,
                                                 const GState* __restrict__ st,
                                                 double2* __restrict__ vec, int i0, int j,
                                                 double* __restrict__ norm_part) {
  __shared__ double lds[4];
  __shared__ double2 h[LH_MMAX + 1];
  const VecItem it = items[blockIdx.x];
  if (!part_of(st, it.s, 0)) return;
  const SolveDev S = solves[it.s];
  const GState& g = st[it.s];
  if (threadIdx.x >= i0 && threadIdx.x <= j)
    h[threadIdx.x] = make_double2(g.hp_re[threadIdx.x], g.hp_im[threadIdx.x]);
  __syncthreads();
  const long long n3 = 3LL * S.n;
  double2* V = vec + S.vec_off;
  double2* W = V + (long long)LH_SLOT_W * n3;
  const long long e0 = 3LL * it.t0;
  double acc = 0.0;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const long long e = e0 + threadIdx.x + q * LH_VB;
    if (e < n3) {
      double2 w = W[e];
      for (int i = i0; i <= j; ++i) w = csub(w, cmul(V[(long long)i * n3 + e], h[i]));
      W[e] = w;
      acc += w.x * w.x + w.y * w.y;
    }
  }
  const double t = block_sum(acc, lds);
  if (threadIdx.x == 0) norm_part[S.vi_off + it.b] = t;
}

// x += sum_{i<j} y_i v_i (iterative.rs:142-144) for solves in a cycle.
__global__ __launch_bounds__(LH_VB) void k_xupdate(const VecItem* __restrict__ items,
                                                  const SolveDev* __restrict__ solves,
                                                  const GState* __restrict__ st,
                                                  double2* __restrict__ vec) {
  __shared__ double2 y[LH_MMAX];
  const VecItem it = items[blockIdx.x];
  if (!part_of(st, it.s, 1)) return;
  const SolveDev S = solves[it.s];
  const GState& g = st[it.s];
  const int jj = g.j;
  if (threadIdx.x < jj) y[threadIdx.x] = make_double2(g.y_re[threadIdx.x], g.y_im[threadIdx.x]);
  __syncthreads();
  const long long n3 = 3LL * S.n;
  double2* V = vec + S.vec_off;
  double2* X = V + (long long)LH_SLOT_X * n3;
  const long long e0 = 3LL * it.t0;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const long long e = e0 + threadIdx.x + q * LH_VB;
    if (e < n3) {
      double2 x = X[e];
      for (int i = 0; i < jj; ++i) x = cadd(x, cmul(V[(long long)i * n3 + e], y[i]));
      X[e] = x;
    }
  }
}

// Cross-section partial sums over an item (cda/mod.rs:190-224):
//   out[item] = (sum_i Im(conj(E_ic) p_ic),
//                sum_i Im(sum_ab p_a conj(ainv_ab) conj(p_b)),
//                sum_i |p_i|^2)
// with p in the X slot and E in the E slot.
__global__ __launch_bounds__(LH_VB) void k_cross(const VecItem* __restrict__ items,
                                                const SolveDev* __restrict__ solves,
                                                const double2* __restrict__ vec,
                                                const double2* __restrict__ ainv,
                                                double* __restrict__ out) {
  __shared__ double lds[4];
  const VecItem it = items[blockIdx.x];
  const SolveDev S = solves[it.s];
  const int i = it.t0 + threadIdx.x;
  double ce = 0.0, ca = 0.0, ps = 0.0;
  if (i < S.n) {
    const long long n3 = 3LL * S.n;
    const double2* P = vec + S.vec_off + (long long)LH_SLOT_X * n3 + 3LL * i;
    const double2* E = vec + S.vec_off + (long long)LH_SLOT_E * n3 + 3LL * i;
    const double2* A = ainv + S.ainv_off + (S.ainv_uniform ? 0 : 9LL * i);
    double2 p[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      p[c] = P[c];
      const double2 e = E[c];
      ce += e.x * p[c].y - e.y * p[c].x;  // Im(conj(e) p)
      ps += p[c].x * p[c].x + p[c].y * p[c].y;
    }
    double2 s = make_double2(0.0, 0.0);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b)
        s = cadd(s, cmul(cmul(p[a], cconj(A[3 * a + b])), cconj(p[b])));
    ca = s.y;
  }
  const double t0 = block_sum(ce, lds);
  const double t1 = block_sum(ca, lds);
  const double t2 = block_sum(ps, lds);
  if (threadIdx.x == 0) {
    out[3LL * (S.vi_off + it.b) + 0] = t0;
    out[3LL * (S.vi_off + it.b) + 1] = t1;
    out[3LL * (S.vi_off + it.b) + 2] = t2;
  }
}

// dst_slot = src_slot (src_slot >= 0) or 0, for every solve.
__global__ __launch_bounds__(LH_VB) void k_copy_slot(const VecItem* __restrict__ items,
                                                    const SolveDev* __restrict__ solves,
                                                    double2* __restrict__ vec, int src_slot,
                                                    int dst_slot) {
  const VecItem it = items[blockIdx.x];
  const SolveDev S = solves[it.s];
  const long long n3 = 3LL * S.n;
  double2* b = vec + S.vec_off + (long long)dst_slot * n3;
  const double2* a = vec + S.vec_off + (long long)(src_slot >= 0 ? src_slot : 0) * n3;
  const long long e0 = 3LL * it.t0;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const long long e = e0 + threadIdx.x + q * LH_VB;
    if (e < n3) b[e] = (src_slot >= 0) ? a[e] : make_double2(0.0, 0.0);
  }
}

// ---------------------------------------------------------------------------
// Per-solve kernels: one block of 64 threads per solve.
// ---------------------------------------------------------------------------

__device__ __forceinline__ double sum_parts(const double* p, int n) {
  double a = 0.0;
  for (int b = threadIdx.x; b < n; b += 64) a += p[b];
  return wave_sum(a);
}

__global__ __launch_bounds__(64) void k_zero(int* __restrict__ counter) {
  if (threadIdx.x == 0) counter[0] = 0;
}

// hp_i = sum over items of the dot partials; hc accumulates over passes.
// Four waves per solve, wave w takes i = w, w + 4, ...; its lanes stride
// over the items and combine with a fixed butterfly (deterministic).
// Pass 0 also clears the step counter that k_fin_step of the same Arnoldi
// step counts into (block 0, before any block leaves).
//
// The per-solve kernels below run one block per solve: solve sl[b] for
// block b when a solve list is given (the solves of one solve group), else
// solve b.
__device__ __forceinline__ int solve_of_block(const int* sl) {
  return sl ? sl[blockIdx.x] : (int)blockIdx.x;
}

__global__ __launch_bounds__(256) void k_fin_dots(const SolveDev* __restrict__ solves,
                                                 GState* __restrict__ st, int i0, int j,
                                                 const double2* __restrict__ dot_part,
                                                 int pass, int* __restrict__ counter,
                                                 const int* __restrict__ sl) {
  const int s = solve_of_block(sl);
  if (pass == 0 && blockIdx.x == 0 && threadIdx.x == 0) *counter = 0;
  GState& g = st[s];
  if (!g.inner) return;
  const SolveDev S = solves[s];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int i = i0 + w; i <= j; i += 4) {
    double sr = 0.0, si = 0.0;
    for (int b = lane; b < S.nvi; b += 64) {
      const double2 d = dot_part[(long long)(S.vi_off + b) * LH_NDOT + i];
      sr += d.x;
      si += d.y;
    }
    sr = wave_sum(sr);
    si = wave_sum(si);
    if (lane == 0) {
      g.hp_re[i] = sr;
      g.hp_im[i] = si;
      if (pass == 0) {
        g.hc_re[i] = sr;
        g.hc_im[i] = si;
      } else {
        g.hc_re[i] += sr;
        g.hc_im[i] += si;
      }
    }
  }
}

// iterative.rs:98-135: h_{j+1,j} = |w|, basis vector scale, Givens
// rotations on the new column and on g, inner-loop exit test.
__global__ __launch_bounds__(64) void k_fin_step(const SolveDev* __restrict__ solves,
                                                GState* __restrict__ st, int j,
                                                const double* __restrict__ norm_part,
                                                int fp32, int* __restrict__ counter,
                                                const int* __restrict__ sl) {
  const int s = solve_of_block(sl);
  GState& g = st[s];
  if (!g.inner) {
    if (threadIdx.x == 0) g.nflag = 0;
    return;
  }
  const SolveDev S = solves[s];
  // the new column, the rotations so far: one parallel load into LDS, then
  // the serial Givens sweep of thread 0 reads them from there and carries
  // the rotated entry in a register
  __shared__ double2 col[LH_MMAX + 1], sns[LH_MMAX];
  __shared__ double css[LH_MMAX];
  const int tid = threadIdx.x;
  if (tid <= j) col[tid] = make_double2(g.hc_re[tid], g.hc_im[tid]);
  if (tid < j) {
    css[tid] = g.cs[tid];
    sns[tid] = make_double2(g.sn_re[tid], g.sn_im[tid]);
  }
  const double nsq = sum_parts(norm_part + S.vi_off, S.nvi);
  __syncthreads();
  if (tid != 0) return;
  const int M = LH_MMAX;
  const double hn = sqrt(nsq);
  g.nscale = (hn > 1e-30) ? 1.0 / hn : 0.0;  // lucky breakdown pushes a zero vector
  auto setH = [&](int r, int c, double2 v) {
    g.H_re[r * M + c] = v.x;
    g.H_im[r * M + c] = v.y;
  };
  // rotation i acts on rows i, i+1 of the column; row i+1 comes out of it
  // as the next rotation's row i
  double2 hij = col[0];
  for (int i = 0; i < j; ++i) {
    const double c = css[i];
    const double2 sn = sns[i];
    const double2 hi1 = col[i + 1];
    const double2 temp = cadd(make_double2(c * hij.x, c * hij.y), cmul(sn, hi1));
    const double2 nsn = make_double2(-sn.x, sn.y);  // -conj(sn)
    hij = cadd(cmul(nsn, hij), make_double2(c * hi1.x, c * hi1.y));
    setH(i, j, temp);
  }
  // givens_rotation (iterative.rs:180-191)
  const double2 a = hij, b = make_double2(hn, 0.0);
  double c;
  double2 sn;
  if (cabs_(b) < 1e-30) {
    c = 1.0;
    sn = make_double2(0.0, 0.0);
  } else if (cabs_(a) < 1e-30) {
    c = 0.0;
    sn = cdiv(b, make_double2(cabs_(b), 0.0));
  } else {
    const double r = sqrt((a.x * a.x + a.y * a.y) + (b.x * b.x + b.y * b.y));
    c = cabs_(a) / r;
    sn = cdiv(cmul(cdiv(a, make_double2(cabs_(a), 0.0)), cconj(b)), make_double2(r, 0.0));
  }
  g.cs[j] = c;
  g.sn_re[j] = sn.x;
  g.sn_im[j] = sn.y;
  setH(j, j, cadd(make_double2(c * a.x, c * a.y), cmul(sn, b)));
  setH(j + 1, j, make_double2(0.0, 0.0));
  const double2 gj = make_double2(g.g_re[j], g.g_im[j]);
  const double2 gj1 = make_double2(g.g_re[j + 1], g.g_im[j + 1]);
  const double2 t = cadd(make_double2(c * gj.x, c * gj.y), cmul(sn, gj1));
  const double2 u = cadd(cmul(make_double2(-sn.x, sn.y), gj), make_double2(c * gj1.x, c * gj1.y));
  g.g_re[j] = t.x;
  g.g_im[j] = t.y;
  g.g_re[j + 1] = u.x;
  g.g_im[j + 1] = u.y;
  g.total_iters += 1;
  g.matvecs += 1;
  if (fp32) g.matvecs32 += 1;
  g.j = j + 1;
  const bool brk = cabs_(u) < g.abs_tol;
  const bool more = !brk && (g.j < S.m) && (g.total_iters < S.max_iter);
  g.inner = more ? 1 : 0;
  g.nflag = more ? 1 : 0;
  if (more) atomicAdd(counter, 1);
}

// back_substitute (iterative.rs:194-206) for solves in a cycle.
__global__ __launch_bounds__(64) void k_backsub(GState* __restrict__ st,
                                               const int* __restrict__ sl) {
  const int s = solve_of_block(sl);
  GState& g = st[s];
  if (!g.started || threadIdx.x != 0) return;
  const int M = LH_MMAX;
  const int m = g.j;
  for (int i = 0; i < m; ++i) {
    g.y_re[i] = 0.0;
    g.y_im[i] = 0.0;
  }
  for (int i = m - 1; i >= 0; --i) {
    double2 sum = make_double2(g.g_re[i], g.g_im[i]);
    for (int l = i + 1; l < m; ++l)
      sum = csub(sum, cmul(make_double2(g.H_re[i * M + l], g.H_im[i * M + l]),
                           make_double2(g.y_re[l], g.y_im[l])));
    const double2 hii = make_double2(g.H_re[i * M + i], g.H_im[i * M + i]);
    if (cabs_(hii) > 1e-30) {
      const double2 yi = cdiv(sum, hii);
      g.y_re[i] = yi.x;
      g.y_im[i] = yi.y;
    }
  }
}

// Residual check at the top of a cycle (iterative.rs:53-67), which equals
// the check after the previous cycle (:146-158). init = 1 on the first call:
// ||rhs|| and abs_tol are set (rhs_norm < 1e-30 returns x = 0, :45-48).
// Solves that continue start a new cycle: beta = ||r||, g = beta e_1,
// nscale = 1 / beta for v_0 = r / beta (:72-83).
__global__ __launch_bounds__(64) void k_check(const SolveDev* __restrict__ solves,
                                             GState* __restrict__ st,
                                             const double* __restrict__ norm_part, int init,
                                             int count_matvec, int* __restrict__ counter,
                                             const int* __restrict__ sl) {
  const int s = solve_of_block(sl);
  GState& g = st[s];
  const SolveDev S = solves[s];
  if (!init && !g.started) return;
  const double nsq = sum_parts(norm_part + S.vi_off, S.nvi);
  if (threadIdx.x != 0) return;
  const double nrm = sqrt(nsq);
  if (init) {
    g.status = 0;
    g.inner = 0;
    g.j = 0;
    g.total_iters = 0;
    g.cycles = 0;
    g.nflag = 0;
    g.matvecs = 0;
    g.matvecs32 = 0;
    g.rhs_norm = nrm;
    g.abs_tol = S.tol * nrm;
    if (nrm < 1e-30) {
      g.status = 3;
      g.res_norm = 0.0;
      g.started = 0;
      return;
    }
  } else if (count_matvec) {
    g.matvecs += 1;
  }
  g.res_norm = nrm;
  g.started = 0;
  g.inner = 0;
  if (nrm < g.abs_tol) {
    g.status = 1;
    return;
  }
  if (g.total_iters >= S.max_iter) {
    g.status = 2;
    return;
  }
  g.cycles += 1;
  g.started = 1;
  g.inner = 1;
  g.j = 0;
  for (int i = 0; i <= LH_MMAX; ++i) {
    g.g_re[i] = 0.0;
    g.g_im[i] = 0.0;
  }
  g.g_re[0] = nrm;
  g.nscale = 1.0 / nrm;
  atomicAdd(counter, 1);
}

// ---------------------------------------------------------------------------
// Launchers.
// ---------------------------------------------------------------------------

void launch_vnorm(hipStream_t st, int nitems, const VecItem* items, const SolveDev* solves,
                  const GState* gs, int which, const double2* vec, int slot, double* np) {
  hipLaunchKernelGGL(k_vnorm, dim3(nitems), dim3(LH_VB), 0, st, items, solves, gs, which, vec,
                     slot, np);
}
void launch_scale(hipStream_t st, int nitems, const VecItem* items, const SolveDev* solves,
                  const GState* gs, int which, double2* vec, int src_slot, int dst_slot) {
  hipLaunchKernelGGL(k_scale, dim3(nitems), dim3(LH_VB), 0, st, items, solves, gs, which, vec,
                     src_slot, dst_slot);
}
void launch_dots(hipStream_t st, int nitems, const VecItem* items, const SolveDev* solves,
                 const GState* gs, const double2* vec, int i0, int j, double2* dp) {
  hipLaunchKernelGGL(k_dots, dim3(nitems), dim3(LH_VB), 0, st, items, solves, gs, vec, i0, j, dp);
}
void launch_update(hipStream_t st, int nitems, const VecItem* items, const SolveDev* solves,
                   const GState* gs, double2* vec, int i0, int j, double* np) {
  hipLaunchKernelGGL(k_update, dim3(nitems), dim3(LH_VB), 0, st, items, solves, gs, vec, i0, j,
                     np);
}
void launch_xupdate(hipStream_t st, int nitems, const VecItem* items, const SolveDev* solves,
                    const GState* gs, double2* vec) {
  hipLaunchKernelGGL(k_xupdate, dim3(nitems), dim3(LH_VB), 0, st, items, solves, gs, vec);
}
void launch_cross(hipStream_t st, int nitems, const VecItem* items, const SolveDev* solves,
                  const double2* vec, const double2* ainv, double* out) {
  hipLaunchKernelGGL(k_cross, dim3(nitems), dim3(LH_VB), 0, st, items, solves, vec, ainv, out);
}
void launch_copy_slot(hipStream_t st, int nitems, const VecItem* items, const SolveDev* solves,
                      double2* vec, int src_slot, int dst_slot) {
  hipLaunchKernelGGL(k_copy_slot, dim3(nitems), dim3(LH_VB), 0, st, items, solves, vec, src_slot,
                     dst_slot);
}
void launch_zero(hipStream_t st, int* counter) {
  hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, st, counter);
}
void launch_fin_dots(hipStream_t st, int ns, const SolveDev* solves, GState* gs, int i0, int j,
                     const double2* dp, int pass, int* counter, const int* sl) {
  hipLaunchKernelGGL(k_fin_dots, dim3(ns), dim3(256), 0, st, solves, gs, i0, j, dp, pass,
                     counter, sl);
}
void launch_fin_step(hipStream_t st, int ns, const SolveDev* solves, GState* gs, int j,
                     const double* np, int fp32, int* counter, const int* sl) {
  hipLaunchKernelGGL(k_fin_step, dim3(ns), dim3(64), 0, st, solves, gs, j, np, fp32, counter,
                     sl);
}
void launch_backsub(hipStream_t st, int ns, GState* gs, const int* sl) {
  hipLaunchKernelGGL(k_backsub, dim3(ns), dim3(64), 0, st, gs, sl);
}
void launch_check(hipStream_t st, int ns, const SolveDev* solves, GState* gs, const double* np,
                  int init, int count_matvec, int* counter, const int* sl) {
  hipLaunchKernelGGL(k_check, dim3(ns), dim3(64), 0, st, solves, gs, np, init, count_matvec,
                     counter, sl);
}

}  // namespace lh
