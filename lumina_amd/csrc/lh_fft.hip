// FFT block-Toeplitz matvec for dipoles filling a regular grid
// (fft_matvec.rs:99-258): G x = IFFT( FFT(K) . FFT(x) ) / len over a
// circulant embedding of the point tensor.
//
//  * The embedding period along each axis is the smallest 7-smooth
//    P >= 2n - 1 (the reference pads to 2n; any P >= 2n - 1 keeps the
//    displacements -(n-1) .. n-1 apart, and a 7-smooth length keeps hipFFT
//    on its radix kernels).
//  * K_ab at padded index (dl mod P0, dm mod P1, dn mod P2) holds the point
//    tensor G(dl h0, dm h1, dn h2) (greens.rs:34-71) for |dl| < n0, ...,
//    zero at the origin and in the padding. Six components (G is symmetric),
//    stored component-major [6][len], FP64 and an FP32 copy.
//  * With FCD, when the near pairs' FCD flags are a function of the grid
//    offset alone (checked on the host), k_fft_fcd replaces K at the FCD
//    offsets by the cell average and the near list is not used (fft = 2);
//    otherwise the near list carries T_FCD - G_point (fft = 1).
//  * One matvec: k_fft_scatter (x of every dipole to its grid point, zero
//    elsewhere, 3 component buffers) -> forward FFT (batch 3) -> k_fft_mul
//    (y_a = sum_b K_ab x_b per frequency, in place) -> inverse FFT (batch 3)
//    -> k_fft_gather (y / len at the dipoles, into the far partial slice
//    that k_reduce subtracts).
#include <climits>
#include <type_traits>

#include "lh_device.h"

namespace lh {

// Point tensor of every circulant offset, one thread per padded point.
__global__ __launch_bounds__(256) void k_fft_kernel(long long len, int P0, int P1, int P2, int n0,
                                                    int n1, int n2, double h0, double h1,
                                                    double h2, double k, double inv_k, double c0,
                                                    double2* __restrict__ K) {
  const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
  if (q >= len) return;
  const int c = (int)(q % P2);
  const long long r = q / P2;
  const int b = (int)(r % P1);
  const int a = (int)(r / P1);
  // the offset a index stands for: a, or a - P for the negative half
  const int dl = a < n0 ? a : (a > P0 - n0 ? a - P0 : INT_MIN);
  const int dm = b < n1 ? b : (b > P1 - n1 ? b - P1 : INT_MIN);
  const int dn = c < n2 ? c : (c > P2 - n2 ? c - P2 : INT_MIN);
  double t[12];
#pragma unroll
  for (int u = 0; u < 12; ++u) t[u] = 0.0;
  if (dl != INT_MIN && dm != INT_MIN && dn != INT_MIN && (dl | dm | dn) != 0)
    g6_point64(dl * h0, dm * h1, dn * h2, k, inv_k, c0, t);
#pragma unroll
  for (int u = 0; u < 6; ++u) K[u * len + q] = make_double2(t[2 * u], t[2 * u + 1]);
}

// The FCD cell average (greens.rs:94-167) at the listed grid offsets, over
// the point tensor k_fft_kernel wrote there.
__global__ __launch_bounds__(64) void k_fft_fcd(int m, const int* __restrict__ offs,
                                                long long len, int P0, int P1, int P2, double h0,
                                                double h1, double h2, double cell, double k,
                                                double inv_k, double c0,
                                                double2* __restrict__ K) {
  const int e = blockIdx.x * 64 + threadIdx.x;
  if (e >= m) return;
  const int dl = offs[3 * e], dm = offs[3 * e + 1], dn = offs[3 * e + 2];
  double t[12];
  fcd_average64(dl * h0, dm * h1, dn * h2, cell, k, inv_k, c0, t);
  const long long a = dl < 0 ? dl + P0 : dl, b = dm < 0 ? dm + P1 : dm, c = dn < 0 ? dn + P2 : dn;
  const long long q = (a * P1 + b) * P2 + c;
#pragma unroll
  for (int u = 0; u < 6; ++u) K[u * len + q] = make_double2(t[2 * u], t[2 * u + 1]);
}

__global__ __launch_bounds__(256) void k_fft_to32(long long m, const double2* __restrict__ a,
                                                  float2* __restrict__ b) {
  const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
  if (q < m) b[q] = make_float2((float)a[q].x, (float)a[q].y);
}

template <class T2>
__device__ __forceinline__ T2 mk2(double x, double y) {
  if constexpr (sizeof(T2) == 8) return make_float2((float)x, (float)y);
  else return make_double2(x, y);
}

// x of every dipole to its grid point, zero elsewhere: X[c][q].
template <class T2>
__global__ __launch_bounds__(256) void k_fft_scatter(long long len,
                                                     const int* __restrict__ pt2dip,
                                                     const double2* __restrict__ x,
                                                     T2* __restrict__ X) {
  const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
  if (q >= len) return;
  const int j = pt2dip[q];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double2 v = j >= 0 ? x[3LL * j + c] : make_double2(0.0, 0.0);
    X[c * len + q] = mk2<T2>(v.x, v.y);
  }
}

// y_a = sum_b K_ab x_b at every frequency, in place.
template <class T2>
__global__ __launch_bounds__(256) void k_fft_mul(long long len, const T2* __restrict__ K,
                                                 T2* __restrict__ X) {
  const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
  if (q >= len) return;
  using R = typename std::conditional<sizeof(T2) == 8, float, double>::type;
  R xr[3], xi[3], kr[6], ki[6];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const T2 v = X[c * len + q];
    xr[c] = v.x;
    xi[c] = v.y;
  }
#pragma unroll
  for (int u = 0; u < 6; ++u) {
    const T2 v = K[u * len + q];
    kr[u] = v.x;
    ki[u] = v.y;
  }
  // (xx yy zz xy xz yz)
  const int idx[3][3] = {{0, 3, 4}, {3, 1, 5}, {4, 5, 2}};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    R yr = 0, yi = 0;
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      const int u = idx[a][b];
      yr = fma(kr[u], xr[b], fma(-ki[u], xi[b], yr));
      yi = fma(kr[u], xi[b], fma(ki[u], xr[b], yi));
    }
    T2 o;
    o.x = yr;
    o.y = yi;
    X[a * len + q] = o;
  }
}

// (G x)_i = X[c][point of i] / len into the far partial slice.
template <class T2>
// part is the start of the partial-slice buffer and part_off the solve's
// slice 0 in complex elements: with p32 the slices are float2 from the
// buffer's start, so the offset is applied as an element index, as every
// reader (k_reduce) does.
__global__ __launch_bounds__(256) void k_fft_gather(int n, long long len,
                                                    const int* __restrict__ dip2pt,
                                                    const T2* __restrict__ X, double scale,
                                                    double2* __restrict__ part,
                                                    long long part_off, int p32) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const long long q = dip2pt[i];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const T2 v = X[c * len + q];
    put_part(part, part_off + 3LL * i + c, (double)v.x * scale, (double)v.y * scale, p32);
  }
}

// ---------------------------------------------------------------------------
// Launchers (host).
// ---------------------------------------------------------------------------

static dim3 grid_of(long long m) { return dim3((unsigned)((m + 255) / 256)); }

void launch_fft_kernel(hipStream_t st, long long len, const int P[3], const int n[3],
                       const double h[3], double k, double2* K) {
  hipLaunchKernelGGL(k_fft_kernel, grid_of(len), dim3(256), 0, st, len, P[0], P[1], P[2], n[0],
                     n[1], n[2], h[0], h[1], h[2], k, 1.0 / k, k * k / (4.0 * M_PI), K);
}

void launch_fft_fcd(hipStream_t st, int m, const int* offs, long long len, const int P[3],
                    const double h[3], double cell, double k, double2* K) {
  if (m <= 0) return;
  hipLaunchKernelGGL(k_fft_fcd, dim3((m + 63) / 64), dim3(64), 0, st, m, offs, len, P[0], P[1],
                     P[2], h[0], h[1], h[2], cell, k, 1.0 / k, k * k / (4.0 * M_PI), K);
}

void launch_fft_to32(hipStream_t st, long long m, const double2* a, float2* b) {
  hipLaunchKernelGGL(k_fft_to32, grid_of(m), dim3(256), 0, st, m, a, b);
}

void launch_fft_scatter64(hipStream_t st, long long len, const int* pt2dip, const double2* x,
                          double2* X) {
  hipLaunchKernelGGL(k_fft_scatter<double2>, grid_of(len), dim3(256), 0, st, len, pt2dip, x, X);
}

void launch_fft_scatter32(hipStream_t st, long long len, const int* pt2dip, const double2* x,
                          float2* X) {
  hipLaunchKernelGGL(k_fft_scatter<float2>, grid_of(len), dim3(256), 0, st, len, pt2dip, x, X);
}

void launch_fft_mul64(hipStream_t st, long long len, const double2* K, double2* X) {
  hipLaunchKernelGGL(k_fft_mul<double2>, grid_of(len), dim3(256), 0, st, len, K, X);
}

void launch_fft_mul32(hipStream_t st, long long len, const float2* K, float2* X) {
  hipLaunchKernelGGL(k_fft_mul<float2>, grid_of(len), dim3(256), 0, st, len, K, X);
}

void launch_fft_gather64(hipStream_t st, int n, long long len, const int* dip2pt,
                         const double2* X, double scale, double2* part, long long part_off) {
  hipLaunchKernelGGL(k_fft_gather<double2>, grid_of(n), dim3(256), 0, st, n, len, dip2pt, X,
                     scale, part, part_off, 0);
}

void launch_fft_gather32(hipStream_t st, int n, long long len, const int* dip2pt,
                         const float2* X, double scale, double2* part, long long part_off,
                         int p32) {
  hipLaunchKernelGGL(k_fft_gather<float2>, grid_of(n), dim3(256), 0, st, n, len, dip2pt, X,
                     scale, part, part_off, p32);
}

}  // namespace lh
