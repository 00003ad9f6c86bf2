// cc_device.h -- device-side layout and physics helpers shared by the HIP
// kernels of the channel time step.
//
// Device layout (HBM): the reference's periodic endpoints are folded away.
// A field holds the P1 x P2 unique nodes of each of the n3 planes
// (P1 = n1-2, P2 = n2-2; reference i = ii+2, j = jj+2, k = kk+1), i fastest,
// planes contiguous, components outermost:
//     f[((c*n3 + kk)*P2 + jj)*P1 + ii]
// Periodic neighbours wrap in ii and jj.  Planes kk = 0 and kk = n3-1 are the
// walls; the conserved state there is the initial (file) state and is never
// updated, as in the reference (main_v1.9.f90:284-294 vs the update ranges
// 2..n3-1 at :565-586).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cc {

struct Geom {
  int P1, P2, n3;       // unique x, unique y, planes (incl. walls)
  long long plane;      // P1*P2
  long long field;      // P1*P2*n3
};

// Constants of one channel instance (by value into every kernel).
struct Phys {
  double Re, invRe, ggM, gM2, dt;     // gM2 = gama*Mac**2
  double dx, dy, dXi, invdx, invdy, invdXi;
  double Pex, Pey, Pez;
  double dz1, dzN;
  double f1d2c0, f1d2c1, f1d2c2, f2d1c0, f2d1c1, f2d1c2;
  double b1d2c0, b1d2c1, b1d2c2, b2d1c0, b2d1c1, b2d1c2;
  double gkRP;                        // gama*invRe*invPr
  double Ri;                          // buoyancy: body force -Ri rho along z (0 = reference)
  double Ls;                          // Navier slip length of both walls (0 = no slip)
};

constexpr double kGama = 1.4;
constexpr double kInvPr = 1.0 / 0.71;   // comdata.f90:12
constexpr double kTwo3rd = 2.0 / 3.0;
constexpr double kOne12th = 1.0 / 12.0;

__device__ __forceinline__ int wrap_m(int i, int P) { return i < 0 ? i + P : i; }
__device__ __forceinline__ int wrap_p(int i, int P) { return i >= P ? i - P : i; }
__device__ __forceinline__ int wrapi(int i, int P) {
  return i < 0 ? i + P : (i >= P ? i - P : i);
}

__device__ __forceinline__ long long idx(const Geom& g, int ii, int jj, int kk) {
  return ((long long)kk * g.P2 + jj) * g.P1 + ii;
}

}  // namespace cc
