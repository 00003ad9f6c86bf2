// cc_post.hip -- post-processing of the channel state on the device: time
// averages of the plane statistics, the fluctuation field and the
// Q-criterion, as the reference's stand-alone programs compute them
// (v1.9/codes/avg_v1.7.f90, v1.9/codes/fluc.f90,
// v1.8_validation/codes/qInvariant.f90).  Compiled with -ffp-contract=off
// and written in those programs' operation order, so that on the same
// inputs the results are the programs' own to the last bit.
//
// Fields are in the reference's Fortran order: f[((c*n3 + k)*n2 + j)*n1 + i].
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/cc_abi.h"
#include "cc_device.h"
#include "cc_kernels.h"

namespace cc {

// var - mean(k) for nf fields (fluc.f90:55-62); mean is stat1(n3, nf).
__global__ void k_fluc(long long N, int n12, int n3, int nf, const double* __restrict__ var,
                       const double* __restrict__ mean, double* __restrict__ out) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N * nf) return;
  const int c = (int)(t / N);
  const int k = (int)((t % N) / n12);
  out[t] = var[t] - mean[(long long)c * n3 + k];
}

// Q = (P^2 - S:S + R:R)/2 at the interior nodes (qInvariant.f90:26-81);
// U holds the three velocity fluctuations; boundary nodes get 0.
__global__ void __launch_bounds__(256) k_qcrit(int n1, int n2, int n3,
                                               const double* __restrict__ x,
                                               const double* __restrict__ y,
                                               const double* __restrict__ z,
                                               const double* __restrict__ U,
                                               double* __restrict__ Q) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  const int k = blockIdx.z;
  if (i >= n1) return;
  const long long N = (long long)n1 * n2 * n3;
  const long long n = ((long long)k * n2 + j) * n1 + i;
  if (i == 0 || i == n1 - 1 || j == 0 || j == n2 - 1 || k == 0 || k == n3 - 1) {
    Q[n] = 0.0;
    return;
  }
  const double haf = 0.5;
  const double dx1 = x[i + 1] - x[i], dy1 = y[j + 1] - y[j], dz1 = z[k + 1] - z[k];
  const double dx2 = x[i] - x[i - 1], dy2 = y[j] - y[j - 1], dz2 = z[k] - z[k - 1];
  const long long sx = 1, sy = n1, sz = (long long)n1 * n2;
  double Jt[3][3];
  for (int c = 0; c < 3; c++) {
    const double* u = U + c * N + n;
    const double u0 = u[0];
    Jt[c][0] = ((dx2 / dx1) * (u[sx] - u0) + (dx1 / dx2) * (u0 - u[-sx])) / (dx1 + dx2);
    Jt[c][1] = ((dy2 / dy1) * (u[sy] - u0) + (dy1 / dy2) * (u0 - u[-sy])) / (dy1 + dy2);
    Jt[c][2] = ((dz2 / dz1) * (u[sz] - u0) + (dz1 / dz2) * (u0 - u[-sz])) / (dz1 + dz2);
  }
  double S[3][3], R[3][3];
  S[0][0] = Jt[0][0];
  S[0][1] = haf * (Jt[0][1] + Jt[1][0]);
  S[0][2] = haf * (Jt[0][2] + Jt[2][0]);
  S[1][0] = haf * (Jt[1][0] + Jt[0][1]);
  S[1][1] = Jt[1][1];
  S[1][2] = haf * (Jt[2][1] + Jt[1][2]);
  S[2][0] = haf * (Jt[2][0] + Jt[0][2]);
  S[2][1] = haf * (Jt[2][1] + Jt[1][2]);
  S[2][2] = Jt[2][2];
  R[0][0] = 0.0;
  R[0][1] = haf * (Jt[0][1] - Jt[1][0]);
  R[0][2] = haf * (Jt[0][2] - Jt[2][0]);
  R[1][0] = haf * (Jt[1][0] - Jt[0][1]);
  R[1][1] = 0.0;
  R[1][2] = haf * (Jt[1][2] - Jt[2][1]);
  R[2][0] = haf * (Jt[2][0] - Jt[0][2]);
  R[2][1] = haf * (Jt[2][1] - Jt[1][2]);
  R[2][2] = 0.0;
  // sum(S**2), sum(-R**2) over the 3x3 arrays in column-major order
  double a1 = 0.0, a2 = 0.0;
  for (int col = 0; col < 3; col++)
    for (int row = 0; row < 3; row++) {
      a1 = a1 + S[row][col] * S[row][col];
      a2 = a2 + (-(R[row][col] * R[row][col]));
    }
  const double P = -(S[0][0] + S[1][1] + S[2][2]);
  Q[n] = haf * (P * P - a1 - a2);
}

// Running sums of the plane statistics (23 x n3) and a sample count, added
// after every statistics step while averaging is on; skipped on a halted run
// exactly as the statistics kernels are.
__global__ void k_avg_accum(int m, const double* __restrict__ stats, double* __restrict__ sums,
                            const int* halt, int step) {
  const int h = *halt;
  if (h != 0 && h <= step) return;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < m) sums[t] += stats[t];
  if (t == 0) sums[m] += 1.0;
}

void launch_avg_accum(int n3, const double* stats, double* sums, const int* halt, int step,
                      hipStream_t s) {
  const int m = kNStat * n3;
  hipLaunchKernelGGL(k_avg_accum, dim3((m + 255) / 256), dim3(256), 0, s, m, stats, sums, halt,
                     step);
}

void launch_fluc(int n1, int n2, int n3, int nf, const double* var, const double* mean,
                 double* out, hipStream_t s) {
  const long long N = (long long)n1 * n2 * n3;
  const long long tot = N * nf;
  hipLaunchKernelGGL(k_fluc, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, N, n1 * n2,
                     n3, nf, var, mean, out);
}

void launch_qcrit(int n1, int n2, int n3, const double* x, const double* y, const double* z,
                  const double* U, double* Q, hipStream_t s) {
  hipLaunchKernelGGL(k_qcrit, dim3((n1 + 255) / 256, n2, n3), dim3(256), 0, s, n1, n2, n3, x, y,
                     z, U, Q);
}

}  // namespace cc
