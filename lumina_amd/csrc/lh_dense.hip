// Stored-matrix session kernels, FP64 complex, A row-major n x n:
// y = A x (MatvecSession::matvec, gpu.rs:428-488), one wave per row, lanes
// striding the row so every load instruction reads 64 consecutive complex
// values; and the device assembly of the free-space interaction matrix
// (k_assemble).
#include "lh_device.h"

namespace lh {

__global__ __launch_bounds__(256) void k_dense_matvec(int n, const double2* __restrict__ A,
                                                     const double2* __restrict__ x,
                                                     double2* __restrict__ y) {
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= n) return;
  const double2* a = A + (long long)row * n;
  double sr = 0.0, si = 0.0;
  for (int j = lane; j < n; j += 64) {
    const double2 m = a[j], v = x[j];
    sr += m.x * v.x - m.y * v.y;
    si += m.x * v.y + m.y * v.x;
  }
  sr = wave_sum(sr);
  si = wave_sum(si);
  if (lane == 0) y[row] = make_double2(sr, si);
}

void launch_dense_matvec(hipStream_t st, int n, const double2* A, const double2* x, double2* y) {
  hipLaunchKernelGGL(k_dense_matvec, dim3((n + 3) / 4), dim3(256), 0, st, n, A, x, y);
}

// Dense interaction matrix A = blockdiag(alpha^-1) - G for the stored-matrix
// session (assemble_and_create_session, backend.rs:109-117; the wgpu form is
// gpu.rs:508-607 with matrix_assemble.wgsl). G is the point tensor
// (greens.rs:34-71) of the displacement r_i - r_j; a pair closer than
// 1e-10 nm (|d|^2 < 1e-20) gets a zero block, as in matrix_assemble.wgsl.
// One thread builds the 3x3 block (i, j) from one tensor evaluation in FP64
// and copies the alpha^-1 block on the diagonal. Block row i is
// blockIdx.y and the threads run over j, so each of the block's three matrix
// rows is 48 contiguous bytes per thread, contiguous across the wave.
__global__ __launch_bounds__(256) void k_assemble(int n, const double* __restrict__ pos, double k,
                                                 double inv_k, double c0,
                                                 const double2* __restrict__ diag,
                                                 double2* __restrict__ A) {
  const int i = blockIdx.y;
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= n) return;
  const long long dim = 3LL * n;
  double2 blk[9];
  if (i == j) {
#pragma unroll
    for (int q = 0; q < 9; ++q) blk[q] = diag[9LL * i + q];
  } else {
    const double d[3] = {pos[3LL * i] - pos[3LL * j], pos[3LL * i + 1] - pos[3LL * j + 1],
                         pos[3LL * i + 2] - pos[3LL * j + 2]};
    const double r2 = r2_exact(d[0], d[1], d[2]);
    double Ar = 0.0, Ai = 0.0, Br = 0.0, Bi = 0.0;
    if (r2 >= 1e-20) g_coeffs64(r2, k, inv_k, c0, Ar, Ai, Br, Bi);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const double dd = d[a] * d[c];
        double gr = Br * dd, gi = Bi * dd;
        if (a == c) {
          gr += Ar;
          gi += Ai;
        }
        blk[3 * a + c] = make_double2(-gr, -gi);
      }
  }
  double2* out = A + 3LL * i * dim + 3LL * j;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int c = 0; c < 3; ++c) out[a * dim + c] = blk[3 * a + c];
}

void launch_assemble(hipStream_t st, int n, const double* pos, double k, const double2* diag,
                     double2* A) {
  hipLaunchKernelGGL(k_assemble, dim3((n + 255) / 256, n), dim3(256), 0, st, n, pos, k, 1.0 / k,
                     k * k / (4.0 * M_PI), diag, A);
}

}  // namespace lh
