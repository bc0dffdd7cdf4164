// On-the-fly CDA matvec y = A x for a batch of solves (gfx950).
//
//   y_i = alpha_i^-1 x_i - sum_{j != i} G(r_i, r_j) x_j
//                        - sum_j (m o G(r_i, mirror r_j)) x_j        (substrate)
//
// (assembly.rs:209-318 matvec_on_the_fly). The pair sum is split in two:
//
//  * far field: every pair outside the exclusion radius (and j != i)
//    evaluates the point Green's tensor (greens.rs:34-71) on the fly in the
//    y += A x + B d (d . x) form -- no tensor is materialised. One target
//    dipole per lane. In the row form (k_far_row) source records stream
//    through the scalar cache (every lane of a wave reads the same record);
//    work items are tiles of (64, 128 or 256 targets) x (256 sources), each
//    writing the row sums of its targets to the partial slice of its source
//    tile (no atomics, reproducible results). The free-space forms with one
//    tensor per unordered pair (k_far_sym32 in FP32, sources staged in LDS;
//    k_far_sym64 in FP64) also write the column sums of their sources to the
//    slice of their target tile (k_far_sym32: to per-tile slots, with one
//    row-sum slot per run of source tiles). With a substrate
//    (assembly.rs:272-295) the image tensor at the mirrored source is applied
//    to (-D, -D, +D) o x_j, for every pair including j == i, skipped when
//    |d|^2 < 1e-20: in the row form by k_far_row<R, true>, and in the
//    one-tensor-per-pair form by k_far_sym32/64<NEAR, true>, where one image
//    tensor serves both directions of a pair.
//  * near field (k_near_tensors + k_reduce): pairs inside the exclusion
//    radius use FP64 tensors precomputed once per solve for each distinct
//    displacement of the geometry: the FCD 27-node cell average
//    (greens.rs:94-167) where dist <= 2 cell_size, the point tensor
//    otherwise. k_reduce adds alpha^-1 x, the far partials and the near
//    tensors (an ELL list of table ids) and writes y, or the residual
//    rhs - y and its norm.
//
// The exclusion test |d|^2 <= (2.1 cell)^2 is evaluated on FP32-rounded
// positions with r2f_exact in every kernel and in the host near-list
// builder, so the far field and the near list cover every pair exactly once.
#include <type_traits>

#include "lh_device.h"

namespace lh {

// ---------------------------------------------------------------------------
// Near-field tensors (once per solve), one thread per distinct displacement.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_near_tensors(const SolveDev* __restrict__ solves,
                                                     const int* __restrict__ item_s,
                                                     const int* __restrict__ item_u,
                                                     int nitems,
                                                     const double* __restrict__ disp,
                                                     double* __restrict__ tab) {
  const int b = blockIdx.x;
  if (b >= nitems) return;
  const SolveDev S = solves[item_s[b]];
  const int u = item_u[b] + threadIdx.x;
  if (u >= S.ntab || S.fft == 2) return;   // fft == 2: FCD inside the FFT kernel
  const double* d = disp + 4 * (S.dtab_off + u);
  const double dx = d[0], dy = d[1], dz = d[2];
  double t[12];
  if (d[3] != 0.0)
    fcd_average64(dx, dy, dz, S.cell, S.k, S.inv_k, S.c0, t);
  else
    g6_point64(dx, dy, dz, S.k, S.inv_k, S.c0, t);
  if (S.fft) {
    // the FFT matvec applies the point tensor to every pair: the near list
    // carries the FCD correction only (exactly zero for point pairs)
    double g[12];
    g6_point64(dx, dy, dz, S.k, S.inv_k, S.c0, g);
#pragma unroll
    for (int q = 0; q < 12; ++q) t[q] -= g[q];
  }
  double* o = tab + 12 * (S.tab_off + u);
#pragma unroll
  for (int q = 0; q < 12; ++q) o[q] = t[q];
}

// Every kernel below takes a participation code `which` (the kWhich codes
// of lh_types.h): the phase of the GMRES cycle, and in mixed Arnoldi steps
// the precision of the solve's products, that selects the solves it works on.

// ---------------------------------------------------------------------------
// Source records for one matvec: positions + input vector (+ image-scaled
// input). x is read from the Krylov slot `slot` of each solve.
// ---------------------------------------------------------------------------

__device__ __forceinline__ void set_pos(Src64& r, double x, double y, double z) {
  r.x = x; r.y = y; r.z = z;
  r.xf = (float)x; r.yf = (float)y; r.zf = (float)z; r.pf = 0.0f;
  r.pad = 0.0;
}
__device__ __forceinline__ void set_pos(Src32& r, double x, double y, double z) {
  r.x = (float)x; r.y = (float)y; r.z = (float)z; r.pf = 0.0f;
}

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
  set_pos(r, px[S.pos_off + j], py[S.pos_off + j], pz[S.pos_off + j]);
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    r.v[q] = static_cast<typename SRC::value_type>(v[q]);
    r.w[q] = static_cast<typename SRC::value_type>(w[q]);
  }
  src[S.src_off + j] = r;
}

// ---------------------------------------------------------------------------
// Pair arithmetic, FP64 and FP32.
// ---------------------------------------------------------------------------

template <typename R> struct SrcOf;
template <> struct SrcOf<double> { typedef Src64 type; };
template <> struct SrcOf<float> { typedef Src32 type; };

__device__ __forceinline__ float fx(const Src64& r) { return r.xf; }
__device__ __forceinline__ float fy(const Src64& r) { return r.yf; }
__device__ __forceinline__ float fz(const Src64& r) { return r.zf; }
__device__ __forceinline__ float fx(const Src32& r) { return r.x; }
__device__ __forceinline__ float fy(const Src32& r) { return r.y; }
__device__ __forceinline__ float fz(const Src32& r) { return r.z; }

// FP32: skipped pairs get 1/r = 0 (then A = B = 0; r2 is finite, so
// r = r2 / r stays finite). HW = true: sin/cos from the hardware
// v_sin_f32/v_cos_f32 with the phase in revolutions (k / 2 pi = krev);
// otherwise sincos_f32.
template <bool HW = false>
__device__ __forceinline__ void coeffs32(float r2, bool skip, float k, float krev, float inv_k,
                                         float c0, float& Ar, float& Ai, float& Br, float& Bi) {
  const float inv_r = skip ? 0.0f : __builtin_amdgcn_rsqf(r2);
  const float r = r2 * inv_r;
  float s, c;
  if constexpr (HW) {
    const float rev = r * krev;
    s = __builtin_amdgcn_sinf(rev);
    c = __builtin_amdgcn_cosf(rev);
  } else {
    sincos_f32(k * r, s, c);
  }
  const float u = inv_r * inv_k;       // 1 / (k r)
  const float pf = c0 * inv_r;         // k^2 / (4 pi r)
  const float pr = pf * c, pi = pf * s;  // g = pf e^{ikr}
  const float ar = fmaf(-u, u, 1.0f);  // a = (1 - u^2) + iu
  Ar = pr * ar - pi * u;
  Ai = pr * u + pi * ar;
  // b = (3u^2 - 1) - 3iu = 2 - 3a, so B = g b / r^2 = (g - 1.5 A)(2 / r^2):
  // two FMA and two products from A instead of forming b
  const float h = 2.0f * (inv_r * inv_r);
  Br = fmaf(-1.5f, Ar, pr) * h;
  Bi = fmaf(-1.5f, Ai, pi) * h;
}

// out += (A I + B d d^T) v, as fused multiply-add chains into the
// accumulators (3 FMA per output component).
template <typename R>
__device__ __forceinline__ void apply(R dx, R dy, R dz, R Ar, R Ai, R Br, R Bi, const R* v,
                                      R out[6]) {
  const R dr = fma(dz, v[4], fma(dy, v[2], dx * v[0]));
  const R di = fma(dz, v[5], fma(dy, v[3], dx * v[1]));
  const R wr = fma(-Bi, di, Br * dr);
  const R wi = fma(Bi, dr, Br * di);
  out[0] = fma(dx, wr, fma(-Ai, v[1], fma(Ar, v[0], out[0])));
  out[1] = fma(dx, wi, fma(Ai, v[0], fma(Ar, v[1], out[1])));
  out[2] = fma(dy, wr, fma(-Ai, v[3], fma(Ar, v[2], out[2])));
  out[3] = fma(dy, wi, fma(Ai, v[2], fma(Ar, v[3], out[3])));
  out[4] = fma(dz, wr, fma(-Ai, v[5], fma(Ar, v[4], out[4])));
  out[5] = fma(dz, wi, fma(Ai, v[4], fma(Ar, v[5], out[5])));
}

// out = (A I + B d d^T) v, starting each component with a multiply (a fresh
// sum, not an accumulation onto zero).
template <typename R>
__device__ __forceinline__ void apply_set(R dx, R dy, R dz, R Ar, R Ai, R Br, R Bi, const R* v,
                                          R out[6]) {
  const R dr = fma(dz, v[4], fma(dy, v[2], dx * v[0]));
  const R di = fma(dz, v[5], fma(dy, v[3], dx * v[1]));
  const R wr = fma(-Bi, di, Br * dr);
  const R wi = fma(Bi, dr, Br * di);
  out[0] = fma(dx, wr, fma(-Ai, v[1], Ar * v[0]));
  out[1] = fma(dx, wi, fma(Ai, v[0], Ar * v[1]));
  out[2] = fma(dy, wr, fma(-Ai, v[3], Ar * v[2]));
  out[3] = fma(dy, wi, fma(Ai, v[2], Ar * v[3]));
  out[4] = fma(dz, wr, fma(-Ai, v[5], Ar * v[4]));
  out[5] = fma(dz, wi, fma(Ai, v[4], Ar * v[5]));
}

// Error-free TwoSum of hi + v into (hi, lo).
__device__ __forceinline__ void two_sum(float& hi, float& lo, float v) {
  const float s = hi + v;
  const float bb = s - hi;
  const float err = (hi - (s - bb)) + (v - bb);
  hi = s;
  lo += err;
}

// Row sums: FP64 accumulates directly; FP32 sums runs of LH_RUN sources in
// FP32 and folds each run into a float-float (hi, lo) pair.
#define LH_RUN 32

template <typename R>
struct RowAcc {
  R hi[6], lo[6];
  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int q = 0; q < 6; ++q) hi[q] = lo[q] = R(0);
  }
  __device__ __forceinline__ void fold(const R acc[6]) {
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      if constexpr (std::is_same<R, float>::value) two_sum(hi[q], lo[q], acc[q]);
      else hi[q] += acc[q];
    }
  }
  __device__ __forceinline__ double get(int q) const { return (double)hi[q] + (double)lo[q]; }
};

// ---------------------------------------------------------------------------
// Far field: every ordered tile pair, row sums; SUB adds the Fresnel image.
// ---------------------------------------------------------------------------

// NEAR = false: items whose target and source boxes lie farther apart than
// the exclusion radius (checked on the host on the FP32 positions), so no
// pair of the item is excluded and no pair is i == j: the exclusion test
// drops out and the FP32 |d|^2 takes the FMA form.
template <typename R, bool SUB, bool HW, bool NEAR = true>
__global__ __launch_bounds__(256) void k_far_row(const FarItem* __restrict__ items,
                                                 const SolveDev* __restrict__ solves,
                                                 const GState* __restrict__ st, int which,
                                                 const typename SrcOf<R>::type* __restrict__ src,
                                                 double2* __restrict__ part, int p32) {
  typedef typename SrcOf<R>::type SRC;
  const FarItem it = items[blockIdx.x];
  if (!takes_part(st, it.s, which)) return;
  const SolveDev S = solves[it.s];
  const int i = it.t0 + threadIdx.x;
  const bool valid = i < S.n;
  const SRC* __restrict__ sp = src + S.src_off;
  const SRC& me = sp[valid ? i : 0];
  const float xif = fx(me), yif = fy(me), zif = fz(me);
  const R xi = me.x, yi = me.y, zi = me.z;
  // the prefactor k^2 / (4 pi) is common to every pair: the sums run with
  // c0 = 1 and the row total is scaled once
  const R k = (R)S.k, inv_k = (R)S.inv_k, c0 = R(1), z2 = (R)S.z2;
  const float krev = (float)(S.k * 0.15915494309189535);
  const float rskip2f = S.rskip2f;
  RowAcc<R> row;
  row.zero();
  // the next source record is loaded while the current one is used
  SRC nxt = sp[it.c0];
  for (int j0 = it.c0; j0 < it.c1; j0 += LH_RUN) {
    const int j1 = min(j0 + LH_RUN, it.c1);
    R acc[6] = {0, 0, 0, 0, 0, 0};
    for (int j = j0; j < j1; ++j) {
      const SRC r = nxt;
      nxt = sp[min(j + 1, it.c1 - 1)];
      const R dx = xi - r.x, dy = yi - r.y, dz = zi - r.z;
      R Ar, Ai, Br, Bi;
      if constexpr (!NEAR) {
        if constexpr (std::is_same<R, float>::value)
          coeffs32<HW>(fma(dz, dz, fma(dy, dy, dx * dx)), false, k, krev, inv_k, c0, Ar, Ai,
                       Br, Bi);
        else
          g_coeffs64(dx * dx + dy * dy + dz * dz, k, inv_k, c0, Ar, Ai, Br, Bi, false);
      } else if constexpr (std::is_same<R, float>::value) {
        // d2f <= rskip2f also holds for j == i (d2f = 0, rskip2f >= 0); the
        // exclusion test's |d|^2 is the value's |d|^2
        const float d2f = r2f_exact(xif - fx(r), yif - fy(r), zif - fz(r));
        coeffs32<HW>(d2f, d2f <= rskip2f, k, krev, inv_k, c0, Ar, Ai, Br, Bi);
      } else {
        const float d2f = r2f_exact(xif - fx(r), yif - fy(r), zif - fz(r));
        const bool skip = (d2f <= rskip2f) || (j == i);
        const R r2 = dx * dx + dy * dy + dz * dz;
        g_coeffs64(r2, k, inv_k, c0, Ar, Ai, Br, Bi, skip);
      }
      apply<R>(dx, dy, dz, Ar, Ai, Br, Bi, r.v, acc);
      if (SUB) {
        // image of source j: (x_j, y_j, 2 z0 - z_j)
        const R ez = zi - (z2 - r.z);
        const R e2 = dx * dx + dy * dy + ez * ez;
        const bool eskip = e2 < R(1e-20);
        if constexpr (std::is_same<R, float>::value)
          coeffs32<HW>(e2, eskip, k, krev, inv_k, c0, Ar, Ai, Br, Bi);
        else
          g_coeffs64(e2, k, inv_k, c0, Ar, Ai, Br, Bi, eskip);
        apply<R>(dx, dy, ez, Ar, Ai, Br, Bi, r.w, acc);
      }
    }
    row.fold(acc);
  }
  if (valid) {
    const long long e = S.part_off + (long long)(it.c0 / LH_TILE) * 3LL * S.n + 3LL * i;
#pragma unroll
    for (int c = 0; c < 3; ++c)
      put_part(part, e + c, row.get(2 * c) * S.c0, row.get(2 * c + 1) * S.c0, p32);
  }
}

// ---------------------------------------------------------------------------
// FP32 far field, one tensor per unordered pair, sources staged in LDS.
//
// G(r_i, r_j) = G(r_j, r_i) for the point tensor, so each unordered pair is
// evaluated once and applied both ways: y_i += G x_j and y_j += G x_i. A work
// item is a target tile T and a run of source tiles S <= T, 256 threads:
// wave w holds targets T*256 + 64 w + lane and takes the run's source tiles
// one after the other. On the diagonal tile (T == S)
// wave w takes sub-blocks b >= w only, and its own sub-block for steps 1..32
// (step 32 one way only), so each unordered pair is counted once. Forward
// sums of T over the whole run go to slot fs of T's dipoles, transposed sums
// of each S to slot T - S - 1 of S's dipoles (on the diagonal both go to slot
// fs); one forward slot per run instead of one per tile pair cuts the partial
// sums written and reduced. Missing sources of a partial tile carry
// x = 0 and a copy of a real position: their contribution is exactly 0 (a
// coincident pair is inside the exclusion radius and skipped).
// The four 64-source sub-blocks of S are taken one
// at a time by the whole block: sub-block b is staged in LDS twice over
// (entry e holds source e mod 64), so at step s lane l reads source
// (l + s) mod 64 at entry l + s with no wrap-around. Only the transposed
// accumulators move between lanes: each step's transposed contribution is
// added to the accumulator of lane l + 1 (v_add_f32 with DPP wave_rol:1), so
// after step s lane l holds the sum of source (l + s) mod 64. The next step's
// source is read from LDS while the current one is used, and a scheduling
// barrier keeps the steps apart, which holds the register count to what one
// step in flight needs.
// Per unordered pair: 27 VALU for the tensor coefficients and 2 x 28 for the
// two products (31 with the exclusion test), against 2 x 55 for the two
// ordered pairs in the row kernel. The prefactor k^2 / 4 pi is applied once
// per output. After each sub-block the four waves' transposed sums meet in
// LDS and wave b adds them in wave order, in FP64: results are deterministic.
// NEAR = false: tile pairs whose boxes lie beyond the exclusion radius (no
// excluded pair, no coincident pair, never the diagonal), |d|^2 by FMA.
// SUB = true (free-space solves on a substrate): each pair also carries the
// Fresnel image term. The image of source j sits at (x_j, y_j, 2 z0 - z_j), so
// the image displacement of (i, j) is (dx, dy, ez) with ez = z_i + z_j - 2 z0,
// and that of (j, i) is (-dx, -dy, ez): one image tensor (A, B of the common
// length) gives y_i += G(dx, dy, ez) w_j and y_j += G(-dx, -dy, ez) w_i, with
// w = (-D, -D, +D) o x the image-scaled moment of the source record. Image
// pairs are never excluded; the diagonal adds the self-image (j == i) one way.
// ---------------------------------------------------------------------------

// a + (the value of lane l + 1; lane 63 takes lane 0's): DPP wave_rol:1 on the
// first operand, written so that the move folds into v_add_f32_dpp (old value
// 0, the identity of the add, all rows and banks enabled)
__device__ __forceinline__ float rol1_add(float a, float t) {
  const float r = __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, a), 0x134, 0xF, 0xF, false));
  return r + t;
}

template <bool NEAR, bool SUB = false>
__global__ __launch_bounds__(256) void k_far_sym32(const FarItem* __restrict__ items,
                                                   const SolveDev* __restrict__ solves,
                                                   const GState* __restrict__ st, int which,
                                                   const Src32* __restrict__ src,
                                                   double2* __restrict__ part, int p32) {
  __shared__ float4 s_a[128];        // x, y, z, v0 of source (e mod 64) of the sub-block
  __shared__ float4 s_b[128];        // v1 .. v4
  __shared__ float s_c[128];         // v5
  __shared__ float4 s_d[SUB ? 128 : 1];   // SUB: w0 .. w3 (image-scaled x_j)
  __shared__ float2 s_e[SUB ? 128 : 1];   // SUB: w4, w5
  __shared__ float s_t[4][64][6];    // [wave][source lane][component]: transposed sums
  // [hi, lo][component][thread]: float-float row sums, each thread's own
  // entries, touched once per 32 steps (registers stay free for occupancy)
  __shared__ float s_hl[2][6][256];
  const FarItem it = items[blockIdx.x];
  if (!takes_part(st, it.s, which)) return;
  const SolveDev S = solves[it.s];
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int i = it.t0 + 64 * w + lane;
  const bool tvalid = i < S.n;
  // a wave with no valid target skips the pair work; its missing targets
  // carry x = 0 and the position of the wave's first target
  const bool live = it.t0 + 64 * w < S.n;
  const Src32* __restrict__ sp = src + S.src_off;
  const Src32& me = sp[tvalid ? i : (live ? it.t0 + 64 * w : it.t0)];
  const float xi = me.x, yi = me.y, zi = me.z;
  float vi[6], wi[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    vi[q] = tvalid ? me.v[q] : 0.0f;
    wi[q] = (SUB && tvalid) ? me.w[q] : 0.0f;
  }
  const float k = (float)S.k, inv_k = (float)S.inv_k;
  const float krev = (float)(S.k * 0.15915494309189535);
  const float rskip2f = S.rskip2f;
  const float z2 = SUB ? (float)S.z2 : 0.0f;
  const double c0 = S.c0;
  const long long n3 = 3LL * S.n;
  const int tt = it.t0 / LH_TILE;
  float run[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    run[q] = 0.0f;
    s_hl[0][q][tid] = s_hl[1][q][tid] = 0.0f;
  }

  // adds v[q] (and vl[q], a low part) to the row sums
  auto add_rows = [&](const float* v, const float* vl) {
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      float h = s_hl[0][q][tid], l = s_hl[1][q][tid];
      two_sum(h, l, v[q]);
      if (vl) l += vl[q];
      s_hl[0][q][tid] = h;
      s_hl[1][q][tid] = l;
    }
  };
  auto fold = [&]() {
    add_rows(run, nullptr);
#pragma unroll
    for (int q = 0; q < 6; ++q) run[q] = 0.0f;
  };
  // SUB: the image pair of this lane's target i and source j (position
  // offset dx, dy, source height zs, image-scaled moment ws of x_j): the
  // image of j lies at (x_j, y_j, 2 z0 - z_j), so the displacement is
  // (dx, dy, ez) with ez = z_i + z_j - 2 z0 in both directions. Forward
  // y_i += G(dx, dy, ez) w_j into run; with t, the transposed contribution
  // y_j += G(-dx, -dy, ez) w_i onto t (the same A and B).
  auto image = [&](float dx, float dy, float zs, const float* ws, float* t) {
    const float ez = zi - (z2 - zs);
    const float e2 = fmaf(ez, ez, fmaf(dy, dy, dx * dx));
    float Er, Ei, Fr, Fi;
    coeffs32<true>(e2, e2 < 1e-20f, k, krev, inv_k, 1.0f, Er, Ei, Fr, Fi);
    apply<float>(dx, dy, ez, Er, Ei, Fr, Fi, ws, run);
    if (t) apply<float>(-dx, -dy, ez, Er, Ei, Fr, Fi, wi, t);
  };

  for (int sc = it.c0; sc < it.c1; sc += LH_TILE) {
    const bool diag = NEAR && sc == it.t0;
    // transposed sums of source tile sc / 256 (off the diagonal)
    const long long tslot = S.part_off + (long long)(tt - sc / LH_TILE - 1) * n3;
    for (int b = 0; b < 4; ++b) {
      const int jb = sc + 64 * b;
      if (jb >= S.n) break;   // uniform over the block
      if (tid < 128) {
        // missing sources: x = 0 at the position of the sub-block's first source
        const int j = jb + (tid & 63);
        const bool sv = j < S.n;
        const Src32& r = sp[sv ? j : jb];
        s_a[tid] = make_float4(r.x, r.y, r.z, sv ? r.v[0] : 0.0f);
        s_b[tid] = sv ? make_float4(r.v[1], r.v[2], r.v[3], r.v[4]) : make_float4(0, 0, 0, 0);
        s_c[tid] = sv ? r.v[5] : 0.0f;
        if constexpr (SUB) {
          s_d[tid] = sv ? make_float4(r.w[0], r.w[1], r.w[2], r.w[3]) : make_float4(0, 0, 0, 0);
          s_e[tid] = sv ? make_float2(r.w[4], r.w[5]) : make_float2(0, 0);
        }
      }
      __syncthreads();
      float ta[6] = {0, 0, 0, 0, 0, 0};
      int home = 0;   // lane l ends holding the transposed sum of source (l + home) mod 64

      // steps s0 .. s1 - 1, both directions; the source of step s + 1 is read
      // while step s computes
      auto steps = [&](int s0, int s1) {
        float4 ca = s_a[lane + s0], cb = s_b[lane + s0];
        float cc = s_c[lane + s0];
        float4 cd;
        float2 ce;
        if constexpr (SUB) {
          cd = s_d[lane + s0];
          ce = s_e[lane + s0];
        }
#pragma unroll 4
        for (int s = s0; s < s1; ++s) {
          // entry lane + s + 1 <= 127
          const float4 na = s_a[lane + s + 1], nb = s_b[lane + s + 1];
          const float nc = s_c[lane + s + 1];
          float4 nd;
          float2 ne;
          if constexpr (SUB) {
            nd = s_d[lane + s + 1];
            ne = s_e[lane + s + 1];
          }
          const float vs[6] = {ca.w, cb.x, cb.y, cb.z, cb.w, cc};
          const float dx = xi - ca.x, dy = yi - ca.y, dz = zi - ca.z;
          float Ar, Ai, Br, Bi;
          if constexpr (NEAR) {
            const float d2f = r2f_exact(dx, dy, dz);
            coeffs32<true>(d2f, d2f <= rskip2f, k, krev, inv_k, 1.0f, Ar, Ai, Br, Bi);
          } else {
            coeffs32<true>(fmaf(dz, dz, fmaf(dy, dy, dx * dx)), false, k, krev, inv_k, 1.0f, Ar,
                           Ai, Br, Bi);
          }
          apply<float>(dx, dy, dz, Ar, Ai, Br, Bi, vs, run);
          float t6[6];
          apply_set<float>(dx, dy, dz, Ar, Ai, Br, Bi, vi, t6);
          if constexpr (SUB) {
            const float ws[6] = {cd.x, cd.y, cd.z, cd.w, ce.x, ce.y};
            image(dx, dy, ca.z, ws, t6);
            cd = nd;
            ce = ne;
          }
#pragma unroll
          for (int q = 0; q < 6; ++q) ta[q] = rol1_add(ta[q], t6[q]);
          ca = na;
          cb = nb;
          cc = nc;
          __builtin_amdgcn_sched_barrier(0);
        }
      };

      if (live && !(diag && b < w)) {
        if (diag && b == w) {
          // steps 1..31 both ways; step 32 pairs (l, l + 32) twice over, so
          // each lane adds its own direction only
          steps(1, 32);
          {
            const int e = lane + 32;
            const float4 A4 = s_a[e], B4 = s_b[e];
            const float vs[6] = {A4.w, B4.x, B4.y, B4.z, B4.w, s_c[e]};
            const float dx = xi - A4.x, dy = yi - A4.y, dz = zi - A4.z;
            const float d2f = r2f_exact(dx, dy, dz);
            float Ar, Ai, Br, Bi;
            coeffs32<true>(d2f, d2f <= rskip2f, k, krev, inv_k, 1.0f, Ar, Ai, Br, Bi);
            apply<float>(dx, dy, dz, Ar, Ai, Br, Bi, vs, run);
            if constexpr (SUB) {
              const float4 D4 = s_d[e];
              const float2 E2 = s_e[e];
              const float ws[6] = {D4.x, D4.y, D4.z, D4.w, E2.x, E2.y};
              image(dx, dy, A4.z, ws, nullptr);
            }
          }
          // SUB: the self-image, j == i (step 0, one way)
          if constexpr (SUB) image(0.0f, 0.0f, zi, wi, nullptr);
          fold();
          home = 31;
        } else {
          steps(0, 32);
          fold();
          steps(32, 64);
          fold();
          home = 63;
        }
      }
#pragma unroll
      for (int q = 0; q < 6; ++q) s_t[w][(lane + home) & 63][q] = ta[q];
      __syncthreads();
      // the staged sources are no longer read: the next sub-block may be staged
      // while wave b sums this one's transposed sums (read before the next
      // barrier, written after it)
      if (w == b) {
        double t[6];
#pragma unroll
        for (int q = 0; q < 6; ++q)
          t[q] = ((double)s_t[0][lane][q] + (double)s_t[1][lane][q]) +
                 ((double)s_t[2][lane][q] + (double)s_t[3][lane][q]);
        if (diag) {
          // transposed sums of the wave's own targets: into the row sums as
          // float-float
          float th[6], tl[6];
#pragma unroll
          for (int q = 0; q < 6; ++q) {
            th[q] = (float)t[q];
            tl[q] = (float)(t[q] - (double)th[q]);
          }
          add_rows(th, tl);
        } else if (jb + lane < S.n) {
          // transposed sums of source jb + lane
          const long long e = tslot + 3LL * (jb + lane);
#pragma unroll
          for (int c = 0; c < 3; ++c)
            put_part(part, e + c, t[2 * c] * c0, t[2 * c + 1] * c0, p32);
        }
      }
    }
  }
  if (tvalid) {
    // forward sums of the run (plus, on the diagonal, the transposed sums of
    // the same dipole) to slot fs
    const long long e = S.part_off + (long long)it.fs * n3 + 3LL * i;
    double y[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) y[q] = (double)s_hl[0][q][tid] + (double)s_hl[1][q][tid];
#pragma unroll
    for (int c = 0; c < 3; ++c) put_part(part, e + c, y[2 * c] * c0, y[2 * c + 1] * c0, p32);
  }
}

// ---------------------------------------------------------------------------
// Far field, one tensor per unordered pair (free-space solves), FP64, with
// every source record rotating around the wave: no LDS staging and no barrier
// inside the sub-block loop, which suits small batches whose few work items
// leave one wave per SIMD (the autotuner's FP64 candidate kSymRotBlock;
// k_far_sym64 below stages its sources in LDS and wins on large ones).
//
// G(r_i, r_j) = G(r_j, r_i) for the point tensor, so each unordered pair is
// evaluated once and applied both ways: y_i += G x_j and y_j += G x_i. A work
// item is a (target tile T, source tile S) pair with T >= S, 256 threads:
// wave w holds targets T*256 + 64 w + lane and walks the four 64-source
// sub-blocks of S. Each lane starts with its own source and the sources, with
// their transposed accumulators, rotate one lane per step around the wave
// (DPP wave_ror:1), so after 64 steps every target has met every source of the
// sub-block and every accumulator is back in its lane. On the diagonal tile
// (T == S) wave w takes sub-blocks b >= w only, and its own sub-block b == w
// for steps 1..32 (step 32 one way only), so each unordered pair is counted
// once. Transposed sums are combined over the four waves in LDS in a fixed
// order: results are deterministic. Forward sums of T go to partial slice S,
// transposed sums of S to slice T (on the diagonal both go to slice T).
// Missing sources of a partial tile carry x = 0 and a copy of a real position:
// their contribution is exactly 0 (a coincident pair is inside the exclusion
// radius and skipped).
// ---------------------------------------------------------------------------

// DPP wave_ror:1 (0x13C) on both halves: lane l takes the value of lane l - 1
// (lane 0 from 63)
__device__ __forceinline__ double wave_ror1(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffffLL), 0x13C, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0x13C, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (long long)(unsigned)lo);
}

// One pair (target lane, source rotated into the lane): |d|^2 in FP64, the
// exclusion decision on the FP32-rounded positions.
template <typename R> struct SymPos;
template <> struct SymPos<double> {
  double x, y, z;
  __device__ void set(const Src64& r) { x = r.x; y = r.y; z = r.z; }
  __device__ void rotate() { x = wave_ror1(x); y = wave_ror1(y); z = wave_ror1(z); }
};

template <typename R>
__device__ __forceinline__ void sym_coeffs(const SymPos<R>& t, const SymPos<R>& sp, float rskip2f,
                                           R k, float krev, R inv_k, R c0, R& dx, R& dy, R& dz,
                                           R& Ar, R& Ai, R& Br, R& Bi) {
  dx = t.x - sp.x;
  dy = t.y - sp.y;
  dz = t.z - sp.z;
  const float d2f = r2f_exact((float)t.x - (float)sp.x, (float)t.y - (float)sp.y,
                              (float)t.z - (float)sp.z);
  const double r2 = dx * dx + dy * dy + dz * dz;
  g_coeffs64(r2, k, inv_k, c0, Ar, Ai, Br, Bi, d2f <= rskip2f);
}

template <typename R>
__global__ __launch_bounds__(256) void k_far_sym(const FarItem* __restrict__ items,
                                                 const SolveDev* __restrict__ solves,
                                                 const GState* __restrict__ st, int which,
                                                 const typename SrcOf<R>::type* __restrict__ src,
                                                 double2* __restrict__ part) {
  typedef typename SrcOf<R>::type SRC;
  __shared__ R lds_t[4][4][64][6];   // [wave][sub-block][source lane][component]
  const FarItem it = items[blockIdx.x];
  if (!takes_part(st, it.s, which)) return;
  const SolveDev S = solves[it.s];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = it.t0 + 64 * w + lane;
  const bool tvalid = i < S.n;
  const SRC* __restrict__ sp = src + S.src_off;
  const SRC me = sp[tvalid ? i : 0];
  SymPos<R> tp;
  tp.set(me);
  R vi[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) vi[q] = tvalid ? me.v[q] : R(0);
  const R k = (R)S.k, inv_k = (R)S.inv_k, c0 = (R)S.c0;
  const float krev = (float)(S.k * 0.15915494309189535);
  const float rskip2f = S.rskip2f;
  const bool diag = it.t0 == it.c0;
  R hi[6], lo[6], run[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) hi[q] = lo[q] = run[q] = R(0);

  auto fold = [&]() {
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      if constexpr (std::is_same<R, float>::value) two_sum(hi[q], lo[q], run[q]);
      else hi[q] += run[q];
      run[q] = R(0);
    }
  };

  for (int b = 0; b < 4; ++b) {
    const int jb = it.c0 + 64 * b;
    R ta[6] = {0, 0, 0, 0, 0, 0};
    if (jb >= S.n || (diag && b < w)) {
#pragma unroll
      for (int q = 0; q < 6; ++q) lds_t[w][b][lane][q] = R(0);
      continue;
    }
    const int j = jb + lane;
    const bool sv = j < S.n;
    const SRC rs = sp[sv ? j : jb];
    SymPos<R> spos;
    spos.set(rs);
    R vs[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) vs[q] = sv ? rs.v[q] : R(0);

    // one unordered pair per lane: forward into run, transposed into ta
    auto step = [&](bool both) {
      R dx, dy, dz, Ar, Ai, Br, Bi;
      sym_coeffs<R>(tp, spos, rskip2f, k, krev, inv_k, c0, dx, dy, dz, Ar, Ai, Br, Bi);
      apply<R>(dx, dy, dz, Ar, Ai, Br, Bi, vs, run);
      if (both) apply<R>(dx, dy, dz, Ar, Ai, Br, Bi, vi, ta);
    };
    auto rotate = [&]() {
      spos.rotate();
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        vs[q] = wave_ror1(vs[q]);
        ta[q] = wave_ror1(ta[q]);
      }
    };
    int home;   // rotation of the accumulators at the end
    if (diag && b == w) {
      for (int kk = 1; kk < 32; ++kk) {
        rotate();
        step(true);
      }
      rotate();
      step(false);   // pairs (l, l - 32): each lane adds its own direction
      fold();
      home = 32;
    } else {
      for (int kk = 0; kk < 64; ++kk) {
        step(true);
        rotate();
        if (kk == 31) fold();
      }
      fold();
      home = 0;
    }
#pragma unroll
    for (int q = 0; q < 6; ++q) lds_t[w][b][(lane - home) & 63][q] = ta[q];
  }
  __syncthreads();
  // thread (w, lane): forward sum of target i; transposed sum of source
  // c0 + 64 w + lane over the four waves
  double tr[6];
#pragma unroll
  for (int q = 0; q < 6; ++q)
    tr[q] = ((double)lds_t[0][w][lane][q] + (double)lds_t[1][w][lane][q]) +
            ((double)lds_t[2][w][lane][q] + (double)lds_t[3][w][lane][q]);
  const long long n3 = 3LL * S.n;
  if (diag) {
    if (tvalid) {
      double2* o = part + S.part_off + (long long)(it.c0 / LH_TILE) * n3 + 3LL * i;
#pragma unroll
      for (int c = 0; c < 3; ++c)
        o[c] = make_double2(((double)hi[2 * c] + (double)lo[2 * c]) + tr[2 * c],
                            ((double)hi[2 * c + 1] + (double)lo[2 * c + 1]) + tr[2 * c + 1]);
    }
  } else {
    if (tvalid) {
      double2* o = part + S.part_off + (long long)(it.c0 / LH_TILE) * n3 + 3LL * i;
#pragma unroll
      for (int c = 0; c < 3; ++c)
        o[c] = make_double2((double)hi[2 * c] + (double)lo[2 * c],
                            (double)hi[2 * c + 1] + (double)lo[2 * c + 1]);
    }
    const int jj = it.c0 + 64 * w + lane;
    if (jj < S.n) {
      double2* o = part + S.part_off + (long long)(it.t0 / LH_TILE) * n3 + 3LL * jj;
#pragma unroll
      for (int c = 0; c < 3; ++c) o[c] = make_double2(tr[2 * c], tr[2 * c + 1]);
    }
  }
}

// ---------------------------------------------------------------------------
// FP64 far field, one tensor per unordered pair, sources staged in LDS.
//
// The scheme of k_far_sym32 in FP64: items T >= S, wave w holds targets
// T*256 + 64 w + lane, each 64-source sub-block of S staged in LDS twice over
// and read one step ahead at entry lane + s. DPP has no 64-bit add, so the
// transposed accumulators are moved first (two 32-bit wave_rol:1 moves per
// double, 12 per step) and the step's contribution is then accumulated onto
// them; after step s lane l holds the sum of source (l + s) mod 64. The
// exclusion test uses the FP32-rounded positions (the same r2f_exact test as
// every other kernel and the near list); NEAR = false for tile pairs whose
// boxes lie beyond the exclusion radius. Sums run in FP64; the prefactor
// k^2 / 4 pi is applied once per output. After each sub-block the waves'
// transposed sums meet in LDS and the owning wave adds them in wave order.
// SUB = true adds the Fresnel image term as in k_far_sym32<NEAR, true>.
// ---------------------------------------------------------------------------

__device__ __forceinline__ double wave_rol1(double v) {
  // DPP wave_rol:1 (0x134) on both halves: lane l takes lane l + 1's value
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffffLL), 0x134, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0x134, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (long long)(unsigned)lo);
}

template <bool NEAR, bool SUB = false>
__global__ __launch_bounds__(256) void k_far_sym64(const FarItem* __restrict__ items,
                                                   const SolveDev* __restrict__ solves,
                                                   const GState* __restrict__ st, int which,
                                                   const Src64* __restrict__ src,
                                                   double2* __restrict__ part) {
  // one stride for the five arrays: a step's reads share an address register
  __shared__ double2 s_a[128];       // x, y of source (e mod 64) of the sub-block
  __shared__ double2 s_b[128];       // z, v0
  __shared__ double2 s_c[128];       // v1, v2
  __shared__ double2 s_d[128];       // v3, v4
  __shared__ double2 s_e[128];       // v5, unused (SUB: w0)
  __shared__ double2 s_f[SUB ? 128 : 1];   // SUB: w1, w2 (image-scaled x_j)
  __shared__ double2 s_g[SUB ? 128 : 1];   // SUB: w3, w4
  __shared__ double s_h[SUB ? 128 : 1];    // SUB: w5
  __shared__ double s_t[4][64][6];   // [wave][source lane][component]: transposed sums
  const FarItem it = items[blockIdx.x];
  if (!takes_part(st, it.s, which)) return;
  const SolveDev S = solves[it.s];
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int i = it.t0 + 64 * w + lane;
  const bool tvalid = i < S.n;
  // a wave with no valid target skips the pair work; its missing targets
  // carry x = 0 and the position of the wave's first target
  const bool live = it.t0 + 64 * w < S.n;
  const Src64* __restrict__ sp = src + S.src_off;
  const Src64& me = sp[tvalid ? i : (live ? it.t0 + 64 * w : it.t0)];
  const double xi = me.x, yi = me.y, zi = me.z;
  const float xif = me.xf, yif = me.yf, zif = me.zf;
  double vi[6], wi[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    vi[q] = tvalid ? me.v[q] : 0.0;
    wi[q] = (SUB && tvalid) ? me.w[q] : 0.0;
  }
  const double k = S.k, inv_k = S.inv_k, z2 = S.z2;
  const float rskip2f = S.rskip2f;
  const bool diag = NEAR && it.t0 == it.c0;
  const double c0 = S.c0;
  const long long n3 = 3LL * S.n;
  double run[6] = {0, 0, 0, 0, 0, 0};

  for (int b = 0; b < 4; ++b) {
    const int jb = it.c0 + 64 * b;
    if (jb >= S.n) break;   // uniform over the block
    if (tid < 128) {
      // missing sources: x = 0 at the position of the sub-block's first source
      const int j = jb + (tid & 63);
      const bool sv = j < S.n;
      const Src64& r = sp[sv ? j : jb];
      s_a[tid] = make_double2(r.x, r.y);
      s_b[tid] = make_double2(r.z, sv ? r.v[0] : 0.0);
      s_c[tid] = sv ? make_double2(r.v[1], r.v[2]) : make_double2(0.0, 0.0);
      s_d[tid] = sv ? make_double2(r.v[3], r.v[4]) : make_double2(0.0, 0.0);
      s_e[tid] = make_double2(sv ? r.v[5] : 0.0, (SUB && sv) ? r.w[0] : 0.0);
      if constexpr (SUB) {
        s_f[tid] = sv ? make_double2(r.w[1], r.w[2]) : make_double2(0.0, 0.0);
        s_g[tid] = sv ? make_double2(r.w[3], r.w[4]) : make_double2(0.0, 0.0);
        s_h[tid] = sv ? r.w[5] : 0.0;
      }
    }
    __syncthreads();
    double ta[6] = {0, 0, 0, 0, 0, 0};
    int home = 0;   // lane l ends holding the transposed sum of source (l + home) mod 64

    // SUB: the image pair of this lane's target i and source j (position
    // offset dx, dy, source height zs, image-scaled moment ws of x_j): the
    // image of j lies at (x_j, y_j, 2 z0 - z_j), so the displacement is
    // (dx, dy, ez) with ez = z_i + z_j - 2 z0 in both directions. Forward
    // y_i += G(dx, dy, ez) w_j; with both, the transposed contribution
    // y_j += G(-dx, -dy, ez) w_i (the same A and B).
    auto image = [&](double dx, double dy, double zs, const double* ws, bool both) {
      const double ez = zi - (z2 - zs);
      const double e2 = dx * dx + dy * dy + ez * ez;
      double Er, Ei, Fr, Fi;
      g_coeffs64(e2, k, inv_k, 1.0, Er, Ei, Fr, Fi, e2 < 1e-20);
      apply<double>(dx, dy, ez, Er, Ei, Fr, Fi, ws, run);
      if (both) apply<double>(-dx, -dy, ez, Er, Ei, Fr, Fi, wi, ta);
    };
    // one unordered pair: target of this lane, the source held in a..e (and
    // with SUB its image-scaled moment in e.y, f, g, h); both = false adds
    // the forward direction only
    auto pair = [&](double2 a, double2 bb, double2 c, double2 d, double2 e, double2 f, double2 g,
                    double h, bool both) {
      const double vs[6] = {bb.y, c.x, c.y, d.x, d.y, e.x};
      const double dx = xi - a.x, dy = yi - a.y, dz = zi - bb.x;
      double Ar, Ai, Br, Bi;
      if constexpr (NEAR) {
        const float d2f = r2f_exact(xif - (float)a.x, yif - (float)a.y, zif - (float)bb.x);
        g_coeffs64(dx * dx + dy * dy + dz * dz, k, inv_k, 1.0, Ar, Ai, Br, Bi, d2f <= rskip2f);
      } else {
        g_coeffs64(dx * dx + dy * dy + dz * dz, k, inv_k, 1.0, Ar, Ai, Br, Bi, false);
      }
      apply<double>(dx, dy, dz, Ar, Ai, Br, Bi, vs, run);
      if (both) apply<double>(dx, dy, dz, Ar, Ai, Br, Bi, vi, ta);
      if constexpr (SUB) {
        const double ws[6] = {e.y, f.x, f.y, g.x, g.y, h};
        image(dx, dy, bb.x, ws, both);
      }
    };
    // steps s0 .. s1 - 1, both directions; the source of step s + 1 is read
    // while step s computes
    auto steps = [&](int s0, int s1) {
      double2 ca = s_a[lane + s0], cb = s_b[lane + s0], cc = s_c[lane + s0],
              cd = s_d[lane + s0], ce = s_e[lane + s0];
#pragma unroll 2
      for (int s = s0; s < s1; ++s) {
        // entry lane + s + 1 <= 127
        const int e1 = lane + s + 1;
        const double2 na = s_a[e1], nb = s_b[e1], nc = s_c[e1], nd = s_d[e1], ne = s_e[e1];
        // SUB: the rest of the image-scaled moment is read in its own step
        // (its loads overlap the direct pair), not one step ahead: fewer
        // registers held across the steps
        double2 cf{}, cg{};
        double ch = 0.0;
        if constexpr (SUB) {
          cf = s_f[lane + s];
          cg = s_g[lane + s];
          ch = s_h[lane + s];
        }
#pragma unroll
        for (int q = 0; q < 6; ++q) ta[q] = wave_rol1(ta[q]);
        pair(ca, cb, cc, cd, ce, cf, cg, ch, true);
        ca = na;
        cb = nb;
        cc = nc;
        cd = nd;
        ce = ne;
        __builtin_amdgcn_sched_barrier(0);
      }
    };

    if (live && !(diag && b < w)) {
      if (diag && b == w) {
        // steps 1..31 both ways; step 32 pairs (l, l + 32) twice over, so
        // each lane adds its own direction only
        steps(1, 32);
        const int e = lane + 32;
        if constexpr (SUB)
          pair(s_a[e], s_b[e], s_c[e], s_d[e], s_e[e], s_f[e], s_g[e], s_h[e], false);
        else
          pair(s_a[e], s_b[e], s_c[e], s_d[e], s_e[e], double2{}, double2{}, 0.0, false);
        // SUB: the self-image, j == i (step 0, one way)
        if constexpr (SUB) image(0.0, 0.0, zi, wi, false);
        home = 31;
      } else {
        steps(0, 64);
        home = 63;
      }
    }
#pragma unroll
    for (int q = 0; q < 6; ++q) s_t[w][(lane + home) & 63][q] = ta[q];
    __syncthreads();
    // the staged sources are no longer read: the next sub-block may be staged
    // while wave b sums this one's transposed sums (read before the next
    // barrier, written after it)
    if (w == b) {
      double t[6];
#pragma unroll
      for (int q = 0; q < 6; ++q)
        t[q] = (s_t[0][lane][q] + s_t[1][lane][q]) + (s_t[2][lane][q] + s_t[3][lane][q]);
      if (diag) {
        // transposed sums of the wave's own targets: into their row sums
#pragma unroll
        for (int q = 0; q < 6; ++q) run[q] += t[q];
      } else if (jb + lane < S.n) {
        // transposed sums of source jb + lane go to the target tile's slice
        double2* o = part + S.part_off + (long long)(it.t0 / LH_TILE) * n3 + 3LL * (jb + lane);
#pragma unroll
        for (int c = 0; c < 3; ++c) o[c] = make_double2(t[2 * c] * c0, t[2 * c + 1] * c0);
      }
    }
  }
  if (tvalid) {
    // forward sums (plus, on the diagonal, the transposed sums of the same
    // dipole) to the source tile's slice
    double2* o = part + S.part_off + (long long)(it.c0 / LH_TILE) * n3 + 3LL * i;
#pragma unroll
    for (int c = 0; c < 3; ++c) o[c] = make_double2(run[2 * c] * c0, run[2 * c + 1] * c0);
  }
}

// ---------------------------------------------------------------------------
// Reduction: y_i = alpha_i^-1 x_i - sum_tiles part - sum_near T_ij x_j.
// mode 0 writes y to slot `out_slot`; mode 1 writes r = rhs - y to the
// residual slot and a per-item partial of |r|^2 to norm_part.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(LH_VB) void k_reduce(
    const VecItem* __restrict__ items, const SolveDev* __restrict__ solves,
    const GState* __restrict__ st, int which, const int* __restrict__ ell_col,
    const int* __restrict__ ell_tid, const double* __restrict__ tab,
    const double2* __restrict__ ainv, double2* __restrict__ vec, int in_slot, int out_slot,
    const double2* __restrict__ part, int p32, const int* __restrict__ nslot, int mode,
    double* __restrict__ norm_part) {
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
    // the far partial slices, one per source tile (the FFT matvec fills
    // slice 0), or with nslot the slot count of this item's tile (the FP32
    // one-tensor-per-pair form's runs), each thread reading its dipole's
    // contiguous bytes per slice. Summing them in element order across the
    // block instead (fully contiguous block-wide loads, totals passed through
    // LDS) measured within 1% on the 1,024-solve batch and 1.7x slower on a
    // single 38,637-dipole solve, whose 151 blocks leave CUs idle.
    const int nslice = S.fft ? 1 : (nslot ? nslot[S.vi_off + it.b] : S.nt);
    for (int ch = 0; ch < nslice; ++ch) {
      const long long e = S.part_off + (long long)ch * n3 + 3LL * i;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const double2 q = get_part(part, e + c, p32);
        yr[c] -= q.x;
        yi[c] -= q.y;
      }
    }
    const int kw = S.fft == 2 ? 0 : S.kw;   // fft == 2: no near list
    for (int kk = 0; kk < kw; ++kk) {
      const long long e = S.ell_off + (long long)kk * S.n + i;
      const int u = ell_tid[e];
      if (u < 0) break;  // pads sit at the end of a row
      const int j = ell_col[e];
      const double* t = tab + 12 * (S.tab_off + u);
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
                         const int* item_u, const double* disp, double* tab) {
  if (nitems <= 0) return;
  hipLaunchKernelGGL(k_near_tensors, dim3(nitems), dim3(256), 0, st, solves, item_s, item_u,
                     nitems, disp, tab);
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
                  const SolveDev* solves, const GState* gs, int which, const Src64* src,
                  double2* part, int block) {
  if (nitems <= 0) return;
  if (sub)
    hipLaunchKernelGGL((k_far_row<double, true, false>), dim3(nitems), dim3(block), 0, st, items,
                       solves, gs, which, src, part, 0);
  else
    hipLaunchKernelGGL((k_far_row<double, false, false>), dim3(nitems), dim3(block), 0, st, items,
                       solves, gs, which, src, part, 0);
}

void launch_far32(hipStream_t st, bool sub, int nitems, const FarItem* items,
                  const SolveDev* solves, const GState* gs, int which, const Src32* src,
                  double2* part, bool hw_sincos, int block, int p32) {
  if (nitems <= 0) return;
  if (sub && hw_sincos)
    hipLaunchKernelGGL((k_far_row<float, true, true>), dim3(nitems), dim3(block), 0, st, items,
                       solves, gs, which, src, part, p32);
  else if (sub)
    hipLaunchKernelGGL((k_far_row<float, true, false>), dim3(nitems), dim3(block), 0, st, items,
                       solves, gs, which, src, part, p32);
  else if (hw_sincos)
    hipLaunchKernelGGL((k_far_row<float, false, true>), dim3(nitems), dim3(block), 0, st, items,
                       solves, gs, which, src, part, p32);
  else
    hipLaunchKernelGGL((k_far_row<float, false, false>), dim3(nitems), dim3(block), 0, st, items,
                       solves, gs, which, src, part, p32);
}

// Free-space items without excluded pairs (k_far_row<R, false, HW, false>).
void launch_far_clean64(hipStream_t st, int nitems, const FarItem* items, const SolveDev* solves,
                        const GState* gs, int which, const Src64* src, double2* part, int block) {
  if (nitems <= 0) return;
  hipLaunchKernelGGL((k_far_row<double, false, false, false>), dim3(nitems), dim3(block), 0, st,
                     items, solves, gs, which, src, part, 0);
}

void launch_far_clean32(hipStream_t st, int nitems, const FarItem* items, const SolveDev* solves,
                        const GState* gs, int which, const Src32* src, double2* part,
                        bool hw_sincos, int block, int p32) {
  if (nitems <= 0) return;
  if (hw_sincos)
    hipLaunchKernelGGL((k_far_row<float, false, true, false>), dim3(nitems), dim3(block), 0, st,
                       items, solves, gs, which, src, part, p32);
  else
    hipLaunchKernelGGL((k_far_row<float, false, false, false>), dim3(nitems), dim3(block), 0, st,
                       items, solves, gs, which, src, part, p32);
}

// near = false: items without excluded pairs (k_far_sym32<false>); sub:
// solves on a substrate (the image term, k_far_sym32<., true>).
void launch_far_sym32(hipStream_t st, bool near, bool sub, int nitems, const FarItem* items,
                      const SolveDev* solves, const GState* gs, int which, const Src32* src,
                      double2* part, int p32) {
  if (nitems <= 0) return;
  if (sub) {
    if (near)
      hipLaunchKernelGGL((k_far_sym32<true, true>), dim3(nitems), dim3(256), 0, st, items, solves,
                         gs, which, src, part, p32);
    else
      hipLaunchKernelGGL((k_far_sym32<false, true>), dim3(nitems), dim3(256), 0, st, items,
                         solves, gs, which, src, part, p32);
  } else if (near) {
    hipLaunchKernelGGL((k_far_sym32<true>), dim3(nitems), dim3(256), 0, st, items, solves, gs,
                       which, src, part, p32);
  } else {
    hipLaunchKernelGGL((k_far_sym32<false>), dim3(nitems), dim3(256), 0, st, items, solves, gs,
                       which, src, part, p32);
  }
}

// The rotating form (k_far_sym<double>), every item T >= S with the
// exclusion test.
void launch_far_symrot64(hipStream_t st, int nitems, const FarItem* items, const SolveDev* solves,
                         const GState* gs, int which, const Src64* src, double2* part) {
  if (nitems <= 0) return;
  hipLaunchKernelGGL((k_far_sym<double>), dim3(nitems), dim3(256), 0, st, items, solves, gs,
                     which, src, part);
}

// near = false: items without excluded pairs (k_far_sym64<false>); sub:
// solves on a substrate (the image term, k_far_sym64<., true>).
void launch_far_sym64(hipStream_t st, bool near, bool sub, int nitems, const FarItem* items,
                      const SolveDev* solves, const GState* gs, int which, const Src64* src,
                      double2* part) {
  if (nitems <= 0) return;
  if (sub) {
    if (near)
      hipLaunchKernelGGL((k_far_sym64<true, true>), dim3(nitems), dim3(256), 0, st, items, solves,
                         gs, which, src, part);
    else
      hipLaunchKernelGGL((k_far_sym64<false, true>), dim3(nitems), dim3(256), 0, st, items,
                         solves, gs, which, src, part);
  } else if (near) {
    hipLaunchKernelGGL((k_far_sym64<true>), dim3(nitems), dim3(256), 0, st, items, solves, gs,
                       which, src, part);
  } else {
    hipLaunchKernelGGL((k_far_sym64<false>), dim3(nitems), dim3(256), 0, st, items, solves, gs,
                       which, src, part);
  }
}

void launch_reduce(hipStream_t st, int nitems, const VecItem* items, const SolveDev* solves,
                   const GState* gs, int which, const int* ell_col, const int* ell_tid,
                   const double* tab, const double2* ainv, double2* vec, int in_slot,
                   int out_slot, const double2* part, int p32, const int* nslot, int mode,
                   double* norm_part) {
  hipLaunchKernelGGL(k_reduce, dim3(nitems), dim3(LH_VB), 0, st, items, solves, gs, which,
                     ell_col, ell_tid, tab, ainv, vec, in_slot, out_slot, part, p32, nslot, mode,
                     norm_part);
}

}  // namespace lh
