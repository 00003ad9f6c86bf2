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
  double invdz1, invdzN;              // their reciprocals (fused kernels)
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
// Arrays that no kernel writes while the reading kernel runs (the metric
// jac/invJac, the forcing tau/rho-bar, a normalisation vector): read through
// the constant address space, a uniform index becomes a scalar load.  Its
// wait then counts scalar loads only, not the vector loads in flight, which
// complete in order: a vector load of one such value would wait for every
// state load issued before it.
__device__ __forceinline__ double cread(const double* p, long long i) {
  return ((const __attribute__((address_space(4))) double*)p)[i];
}
__device__ __forceinline__ float creadf(const float* p, long long i) {
  return ((const __attribute__((address_space(4))) float*)p)[i];
}

__device__ __forceinline__ int wrapi(int i, int P) {
  return i < 0 ? i + P : (i >= P ? i - P : i);
}

__device__ __forceinline__ long long idx(const Geom& g, int ii, int jj, int kk) {
  return ((long long)kk * g.P2 + jj) * g.P1 + ii;
}

// Wall condition of one wall from the reference's parameter studies (the
// device copy of cc_wall; all zero = isothermal no-slip wall).
struct WallSpec {
  int vel, stripes, temp, reserved;
  double amp, kx, omega, phase_odd;
  double disc_omega, grad, Tw, dTdz;
};
enum { kWallNoSlip = 0, kWallWave = 1, kWallDiscs = 2, kWallStripes = 3 };
enum { kWallT1 = 0, kWallTFixed = 1, kWallTGrad = 2 };

// Duplicate nodes of a wall plane: the reference's boundary loop sets every
// node 1..n1 x 1..n2, periodic duplicates included, each at its own
// coordinates.  Slots: ref i = 1 (jj), i = n1 (P2 + jj), j = 1 (2 P2 + ii),
// j = n2 (2 P2 + P1 + ii), then the corners (1,1), (n1,1), (1,n2), (n1,n2).
// Fields per wall: u, v, T, rho, p.
constexpr int kRingFields = 5;
__host__ __device__ __forceinline__ int ring_size(int P1, int P2) { return 2 * P2 + 2 * P1 + 4; }

// Polar angle in [0, 2 pi) of the rotating-disc studies
// (v1.9_RD/v1.9_rd1/codes/comdata.f90:180-193); on the axes, where that
// function leaves its result unset, the angle of the axis.
__device__ __forceinline__ double polar_angle(double A1, double A2) {
  const double pi = 4.0 * atan(1.0);
  const double tmp = atan(fabs(A2 / A1));
  if (A1 > 0 && A2 > 0) return tmp;
  if (A1 < 0 && A2 > 0) return pi - tmp;
  if (A1 < 0 && A2 < 0) return pi + tmp;
  if (A1 > 0 && A2 < 0) return 2.0 * pi - tmp;
  if (A1 == 0 && A2 > 0) return 0.5 * pi;
  if (A1 == 0 && A2 < 0) return 1.5 * pi;
  if (A1 < 0 && A2 == 0) return pi;
  return 0.0;
}

// u, v, T of the wall node at reference indices (i, j) (1-based, duplicates
// included) under wall condition w, from the primitives of the two planes
// above it (a: first interior plane, b: second).  c0..c2 are the wall's
// one-sided first-derivative coefficients.  The bottom-wall blocks of
// main_v1.9_s1.f90:609, main_v1.9_s3.f90:609-610, main_v1.9_rd1.f90:611-640,
// main_v1.9_g1.f90:609-618, main_v1.9_gst3.f90, main_v1.9_gst5.f90,
// main_v1.9_st1.f90:612, main_v1.9_st3.f90:612.
__device__ __forceinline__ void wall_bc(const WallSpec& w, int i, int j, double time,
                                        const double* __restrict__ x1f,
                                        const double* __restrict__ x2f, int n1, int n2,
                                        double c0, double c1, double c2, double ua, double ub,
                                        double va, double vb, double Ta, double Tb, double& u,
                                        double& v, double& T) {
#pragma clang fp contract(off)
  // evaluated in the studies' order, without fused multiply-adds
  const double x = x1f[i - 1], y = x2f[j - 1];
  u = 0.0;
  v = 0.0;
  if (w.vel == kWallWave) {
    const double ph = (j % 2 == 1) ? w.phase_odd : 0.0;
    v = w.amp * sin(w.kx * x + w.omega * time + ph);
  } else if (w.vel == kWallDiscs) {
    const double R = x2f[n2 - 1] / 10.0;
    const double cy[3] = {2.0, 5.0, 8.0};
    for (int d = 0; d < 3; d++) {
      const double ax = x - 2.0 * R, ay = y - cy[d] * R;
      const double r = sqrt(ax * ax + ay * ay);
      if (r < R) {
        const double th = polar_angle(ax, ay);
        u = -(r * w.disc_omega * sin(th));
        v = r * w.disc_omega * cos(th);
        break;
      }
    }
  } else if (w.vel == kWallStripes) {
    bool on;
    if (w.stripes == 0) {
      on = ((6 * j / n2) % 2) == 0;
    } else if (w.stripes == 1) {
      const double r = (x1f[n1 - 1] + x2f[n2 - 1]) / 8.0;
      on = ((int)floor((x - y + x2f[n2 - 1]) / r) % 2) == 0;
    } else {
      const double r = x2f[n2 - 1] / 8.0;
      on = (((int)floor((y - r * sin(x)) / r) + 1) % 3) == 0;
    }
    if (on) {
      u = (w.grad - c1 * ua - c2 * ub) / c0;
      v = (w.grad - c1 * va - c2 * vb) / c0;
    }
  }
  if (w.temp == kWallTFixed)
    T = w.Tw;
  else if (w.temp == kWallTGrad)
    T = (w.dTdz - c1 * Ta - c2 * Tb) / c0;
  else
    T = 1.0;
}

}  // namespace cc
