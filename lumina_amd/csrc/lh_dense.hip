// Stored-matrix y = A x (MatvecSession::matvec, gpu.rs:428-488), FP64
// complex, A row-major n x n. One wave per row: lanes stride the row so
// every load instruction reads 64 consecutive complex values.
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

}  // namespace lh
