// cc_strict.hip -- FP64 stage kernels in the reference's evaluation order.
//
// Compiled with -ffp-contract=off: every product and sum is rounded where
// the Fortran source rounds it, divisions are IEEE divisions, and the
// residual is summed as src - (((((F1+F2)+G1)+G2)+H1)+H2)
// (v1.10/codes/main_v1.9.f90:572).  The only differences from the CPU
// restatement come from pow() (device libm vs host libm, <= 1 ulp) and from
// the order of the plane-statistics sums.
//
// Per RK stage: k_rhs_strict (convective + viscous fluxes + source + RK
// update), k_prim_strict (primitive recovery + mu), k_wallval_strict (wall
// values) and k_wallp_strict (wall pressure extrapolation).  The primitive field W holds
// (u, v, w, E, T, p, mu) on every plane, as the reference's var/mu arrays do.
#include "cc_device.h"
#include "cc_kernels.h"

namespace cc {

enum { WU = 0, WV = 1, WW = 2, WE = 3, WT = 4, WP = 5, WMU = 6 };

// QUICK face pair (main_v1.9.f90:339-361 for x; :377-399 y; :419-441 z).
__device__ __forceinline__ double quick(double ue, double uw, double qmm, double qm, double qc,
                                        double qp, double qpp) {
  double f = 0.0;
  if (ue >= 0.0 && uw >= 0.0)
    f = ue * (6.0 * qc + 3.0 * qp - 1.0 * qm) - uw * (6.0 * qm + 3.0 * qc - 1.0 * qmm);
  if (ue < 0.0 && uw < 0.0)
    f = ue * (6.0 * qp + 3.0 * qc - 1.0 * qpp) - uw * (6.0 * qc + 3.0 * qm - 1.0 * qp);
  if (ue < 0.0 && uw >= 0.0)
    f = ue * (6.0 * qp + 3.0 * qc - 1.0 * qpp) - uw * (6.0 * qm + 3.0 * qc - 1.0 * qmm);
  if (ue >= 0.0 && uw < 0.0)
    f = ue * (6.0 * qc + 3.0 * qp - 1.0 * qm) - uw * (6.0 * qc + 3.0 * qm - 1.0 * qp);
  return f;
}

// Derivatives and point values of the primitives at one node
// (main_v1.9.f90:461-494 in the oracle numbering 461-494).
struct Pt {
  double u, v, w, T, p, mu;
  double dudx, dudy, dudz, dvdx, dvdy, dvdz, dwdx, dwdy, dwdz, dTdx, dTdy, dTdz;
};

__device__ __forceinline__ void point(const Geom& g, const Phys& ph, const double* __restrict__ W,
                                      const double* __restrict__ invJac, int ii, int jj, int kk,
                                      Pt& P) {
  const long long F = g.field;
  auto at = [&](int c, int i, int j, int k) { return W[c * F + idx(g, i, j, k)]; };
  const int im = wrapi(ii - 1, g.P1), ip = wrapi(ii + 1, g.P1);
  const int jm = wrapi(jj - 1, g.P2), jp = wrapi(jj + 1, g.P2);
  P.u = at(WU, ii, jj, kk);
  P.v = at(WV, ii, jj, kk);
  P.w = at(WW, ii, jj, kk);
  P.T = at(WT, ii, jj, kk);
  P.p = at(WP, ii, jj, kk);
  P.mu = at(WMU, ii, jj, kk);
  P.dudx = (at(WU, ip, jj, kk) - at(WU, im, jj, kk)) * 0.5 * ph.invdx;
  P.dudy = (at(WU, ii, jp, kk) - at(WU, ii, jm, kk)) * 0.5 * ph.invdy;
  P.dvdx = (at(WV, ip, jj, kk) - at(WV, im, jj, kk)) * 0.5 * ph.invdx;
  P.dvdy = (at(WV, ii, jp, kk) - at(WV, ii, jm, kk)) * 0.5 * ph.invdy;
  P.dwdx = (at(WW, ip, jj, kk) - at(WW, im, jj, kk)) * 0.5 * ph.invdx;
  P.dwdy = (at(WW, ii, jp, kk) - at(WW, ii, jm, kk)) * 0.5 * ph.invdy;
  P.dTdx = (at(WT, ip, jj, kk) - at(WT, im, jj, kk)) * 0.5 * ph.invdx;
  P.dTdy = (at(WT, ii, jp, kk) - at(WT, ii, jm, kk)) * 0.5 * ph.invdy;
  if (kk == 0) {
    P.dudz = (at(WU, ii, jj, 1) - at(WU, ii, jj, 0)) / ph.dz1;
    P.dvdz = (at(WV, ii, jj, 1) - at(WV, ii, jj, 0)) / ph.dz1;
    P.dwdz = (at(WW, ii, jj, 1) - at(WW, ii, jj, 0)) / ph.dz1;
    P.dTdz = (at(WT, ii, jj, 1) - at(WT, ii, jj, 0)) / ph.dz1;
  } else if (kk == g.n3 - 1) {
    const int N = g.n3 - 1;
    P.dudz = (at(WU, ii, jj, N) - at(WU, ii, jj, N - 1)) / ph.dzN;
    P.dvdz = (at(WV, ii, jj, N) - at(WV, ii, jj, N - 1)) / ph.dzN;
    P.dwdz = (at(WW, ii, jj, N) - at(WW, ii, jj, N - 1)) / ph.dzN;
    P.dTdz = (at(WT, ii, jj, N) - at(WT, ii, jj, N - 1)) / ph.dzN;
  } else {
    const double iJ = invJac[kk];
    P.dudz = (at(WU, ii, jj, kk + 1) - at(WU, ii, jj, kk - 1)) * 0.5 * ph.invdXi * iJ;
    P.dvdz = (at(WV, ii, jj, kk + 1) - at(WV, ii, jj, kk - 1)) * 0.5 * ph.invdXi * iJ;
    P.dwdz = (at(WW, ii, jj, kk + 1) - at(WW, ii, jj, kk - 1)) * 0.5 * ph.invdXi * iJ;
    P.dTdz = (at(WT, ii, jj, kk + 1) - at(WT, ii, jj, kk - 1)) * 0.5 * ph.invdXi * iJ;
  }
}

// Point fluxes fD, gD, hD components 2..5 (main_v1.9.f90:483-518); the
// first component is zero.
__device__ __forceinline__ void fD(const Phys& ph, const Pt& P, double f[4]) {
  const double invRe = ph.invRe, ggM = ph.ggM, muT = P.mu, ktT = P.mu;
  double divg = P.dudx + P.dvdy + P.dwdz;
  double dissF = ggM * invRe *
                 (kTwo3rd * muT * P.u * divg - 2.0 * muT * P.u * P.dudx -
                  muT * P.v * (P.dvdx + P.dudy) - muT * P.w * (P.dwdx + P.dudz));
  f[0] = P.p + kTwo3rd * muT * invRe * divg - 2.0 * muT * invRe * P.dudx;
  f[1] = -muT * invRe * (P.dvdx + P.dudy);
  f[2] = -muT * invRe * (P.dwdx + P.dudz);
  f[3] = -kGama * ktT * invRe * kInvPr * P.dTdx + ggM * P.u * P.p + dissF;
}
__device__ __forceinline__ void gD(const Phys& ph, const Pt& P, double f[4]) {
  const double invRe = ph.invRe, ggM = ph.ggM, muT = P.mu, ktT = P.mu;
  double divg = P.dudx + P.dvdy + P.dwdz;
  double dissG = ggM * invRe *
                 (-muT * P.u * (P.dvdx + P.dudy) + kTwo3rd * muT * P.v * divg -
                  2.0 * muT * P.v * P.dvdy - muT * P.w * (P.dwdy + P.dvdz));
  f[0] = -muT * invRe * (P.dvdx + P.dudy);
  f[1] = P.p + kTwo3rd * muT * invRe * divg - 2.0 * muT * invRe * P.dvdy;
  f[2] = -muT * invRe * (P.dwdy + P.dvdz);
  f[3] = -kGama * ktT * invRe * kInvPr * P.dTdy + ggM * P.v * P.p + dissG;
}
__device__ __forceinline__ void hD(const Phys& ph, const Pt& P, double f[4]) {
  const double invRe = ph.invRe, ggM = ph.ggM, muT = P.mu, ktT = P.mu;
  double divg = P.dudx + P.dvdy + P.dwdz;
  double dissH = ggM * invRe *
                 (-muT * P.u * (P.dwdx + P.dudz) - muT * P.v * (P.dwdy + P.dvdz) +
                  kTwo3rd * muT * P.w * divg - 2.0 * muT * P.w * P.dwdz);
  f[0] = -muT * invRe * (P.dwdx + P.dudz);
  f[1] = -muT * invRe * (P.dwdy + P.dvdz);
  f[2] = P.p + kTwo3rd * muT * invRe * divg - 2.0 * muT * invRe * P.dwdz;
  f[3] = -kGama * ktT * invRe * kInvPr * P.dTdz + ggM * P.w * P.p + dissH;
}

__device__ __forceinline__ bool halted(const int* halt, int step) {
  int h = *halt;
  return h != 0 && h < step;
}

// Fluxes, source, residual and RK2 update for the updated nodes
// (main_v1.9.f90:329-586).  stage 1: Uout = Uin + (dt/2) R;
// stage 2: Uout = Ubase + dt R.
__global__ void __launch_bounds__(256) k_rhs_strict(Geom g, Phys ph, const double* __restrict__ Uin,
                                                    const double* Ubase, double* Uout,
                                                    const double* __restrict__ W,
                                                    const double* __restrict__ jac,
                                                    const double* __restrict__ invJac,
                                                    const double* __restrict__ force, int stage,
                                                    const int* halt, int step) {
  const int ii = blockIdx.x * blockDim.x + threadIdx.x;
  const int jj = blockIdx.y * blockDim.y + threadIdx.y;
  const int kk = blockIdx.z + 1;
  if (ii >= g.P1 || jj >= g.P2) return;
  if (halted(halt, step)) return;
  const long long F = g.field;
  const int P1 = g.P1, P2 = g.P2, n3 = g.n3;
  auto Uc = [&](int c, int i, int j, int k) { return Uin[c * F + idx(g, i, j, k)]; };
  auto Wc = [&](int c, int i, int j, int k) { return W[c * F + idx(g, i, j, k)]; };
  const int im1 = wrapi(ii - 1, P1), im2 = wrapi(ii - 2, P1);
  const int ip1 = wrapi(ii + 1, P1), ip2 = wrapi(ii + 2, P1);
  const int jm1 = wrapi(jj - 1, P2), jm2 = wrapi(jj - 2, P2);
  const int jp1 = wrapi(jj + 1, P2), jp2 = wrapi(jj + 2, P2);
  const double mu_c = Wc(WMU, ii, jj, kk);
  const int k1 = kk + 1;  // reference k

  double F1[5], G1[5], H1[5];
  // x (main_v1.9.f90:333-369)
  double Pe = ph.Re * fabs(Uc(1, ii, jj, kk)) * ph.dx / mu_c;
  if (Pe > ph.Pex) {
    double t1 = (Wc(WU, ii, jj, kk) + Wc(WU, ip1, jj, kk)) * 0.5;
    double t2 = (Wc(WU, ii, jj, kk) + Wc(WU, im1, jj, kk)) * 0.5;
    for (int c = 0; c < 5; c++)
      F1[c] = quick(t1, t2, Uc(c, im2, jj, kk), Uc(c, im1, jj, kk), Uc(c, ii, jj, kk),
                    Uc(c, ip1, jj, kk), Uc(c, ip2, jj, kk)) * ph.invdx / 8.0;
  } else {
    for (int c = 0; c < 5; c++)
      F1[c] = (-Wc(WU, ip2, jj, kk) * Uc(c, ip2, jj, kk) +
               8.0 * Wc(WU, ip1, jj, kk) * Uc(c, ip1, jj, kk) -
               8.0 * Wc(WU, im1, jj, kk) * Uc(c, im1, jj, kk) +
               Wc(WU, im2, jj, kk) * Uc(c, im2, jj, kk)) * ph.invdx * kOne12th;
  }
  // y (main_v1.9.f90:371-407)
  Pe = ph.Re * fabs(Uc(2, ii, jj, kk)) * ph.dy / mu_c;
  if (Pe > ph.Pey) {
    double t1 = (Wc(WV, ii, jj, kk) + Wc(WV, ii, jp1, kk)) * 0.5;
    double t2 = (Wc(WV, ii, jj, kk) + Wc(WV, ii, jm1, kk)) * 0.5;
    for (int c = 0; c < 5; c++)
      G1[c] = quick(t1, t2, Uc(c, ii, jm2, kk), Uc(c, ii, jm1, kk), Uc(c, ii, jj, kk),
                    Uc(c, ii, jp1, kk), Uc(c, ii, jp2, kk)) * ph.invdy / 8.0;
  } else {
    for (int c = 0; c < 5; c++)
      G1[c] = (-Wc(WV, ii, jp2, kk) * Uc(c, ii, jp2, kk) +
               8.0 * Wc(WV, ii, jp1, kk) * Uc(c, ii, jp1, kk) -
               8.0 * Wc(WV, ii, jm1, kk) * Uc(c, ii, jm1, kk) +
               Wc(WV, ii, jm2, kk) * Uc(c, ii, jm2, kk)) * ph.invdy * kOne12th;
  }
  // z (main_v1.9.f90:409-450)
  Pe = ph.Re * fabs(Uc(3, ii, jj, kk)) * ph.dXi * jac[kk] / mu_c;
  const double iJ = invJac[kk];
  if (k1 < 3 || k1 > n3 - 2) {
    for (int c = 0; c < 5; c++)
      H1[c] = (Wc(WW, ii, jj, kk + 1) * Uc(c, ii, jj, kk + 1) -
               Wc(WW, ii, jj, kk - 1) * Uc(c, ii, jj, kk - 1)) * ph.invdXi * iJ * 0.5;
  } else if (Pe > ph.Pez) {
    double t1 = (Wc(WW, ii, jj, kk) + Wc(WW, ii, jj, kk + 1)) * 0.5;
    double t2 = (Wc(WW, ii, jj, kk) + Wc(WW, ii, jj, kk - 1)) * 0.5;
    for (int c = 0; c < 5; c++)
      H1[c] = quick(t1, t2, Uc(c, ii, jj, kk - 2), Uc(c, ii, jj, kk - 1), Uc(c, ii, jj, kk),
                    Uc(c, ii, jj, kk + 1), Uc(c, ii, jj, kk + 2)) * ph.invdXi * iJ / 8.0;
  } else {
    for (int c = 0; c < 5; c++)
      H1[c] = (-Wc(WW, ii, jj, kk + 2) * Uc(c, ii, jj, kk + 2) +
               8.0 * Wc(WW, ii, jj, kk + 1) * Uc(c, ii, jj, kk + 1) -
               8.0 * Wc(WW, ii, jj, kk - 1) * Uc(c, ii, jj, kk - 1) +
               Wc(WW, ii, jj, kk - 2) * Uc(c, ii, jj, kk - 2)) * ph.invdXi * iJ * kOne12th;
  }

  // viscous divergence (main_v1.9.f90:551-563)
  double F2[5], G2[5], H2[5];
  {
    Pt a, b;
    double fa[4], fb[4];
    point(g, ph, W, invJac, ip1, jj, kk, a);
    point(g, ph, W, invJac, im1, jj, kk, b);
    fD(ph, a, fa);
    fD(ph, b, fb);
    F2[0] = (0.0 - 0.0) * 0.5 * ph.invdx;
    for (int c = 0; c < 4; c++) F2[c + 1] = (fa[c] - fb[c]) * 0.5 * ph.invdx;
    point(g, ph, W, invJac, ii, jp1, kk, a);
    point(g, ph, W, invJac, ii, jm1, kk, b);
    gD(ph, a, fa);
    gD(ph, b, fb);
    G2[0] = (0.0 - 0.0) * 0.5 * ph.invdy;
    for (int c = 0; c < 4; c++) G2[c + 1] = (fa[c] - fb[c]) * 0.5 * ph.invdy;
    point(g, ph, W, invJac, ii, jj, kk + 1, a);
    point(g, ph, W, invJac, ii, jj, kk - 1, b);
    hD(ph, a, fa);
    hD(ph, b, fb);
    H2[0] = (0.0 - 0.0) * 0.5 * ph.invdXi * iJ;
    for (int c = 0; c < 4; c++) H2[c + 1] = (fa[c] - fb[c]) * 0.5 * ph.invdXi * iJ;
  }

  // source, residual, RK2 (main_v1.9.f90:565-586)
  const double tdiff = force[0] - force[1];
  const double rb = force[2 + kk];
  double src[5];
  src[0] = 0.0;
  src[1] = 0.5 * tdiff * rb;
  src[2] = 0.0;
  src[3] = 0.0;
  src[4] = ph.ggM * Wc(WU, ii, jj, kk) * 0.5 * tdiff * rb;
  const long long n = idx(g, ii, jj, kk);
  for (int c = 0; c < 5; c++) {
    double R = src[c] - (F1[c] + F2[c] + G1[c] + G2[c] + H1[c] + H2[c]);
    if (stage == 1)
      Uout[c * F + n] = Uin[c * F + n] + 0.5 * ph.dt * R;
    else
      Uout[c * F + n] = Ubase[c * F + n] + ph.dt * R;
  }
}

// Primitive recovery and mu on the updated nodes (main_v1.9.f90:599-609,
// 643); blow-up check of main_v1.9.f90:728-740 on stage 2.
__global__ void __launch_bounds__(256) k_prim_strict(Geom g, Phys ph, const double* __restrict__ U,
                                                     double* __restrict__ W, int check, int* halt,
                                                     unsigned long long* blow, int step, int n1,
                                                     int n2) {
  const int ii = blockIdx.x * blockDim.x + threadIdx.x;
  const int jj = blockIdx.y * blockDim.y + threadIdx.y;
  const int kk = blockIdx.z + 1;
  if (ii >= g.P1 || jj >= g.P2) return;
  if (halted(halt, step)) return;
  const long long F = g.field, n = idx(g, ii, jj, kk);
  const double rho = U[n];
  const double u = U[F + n] / rho, v = U[2 * F + n] / rho, w = U[3 * F + n] / rho;
  const double E = U[4 * F + n] / rho;
  const double T = E - ph.ggM * 0.5 * (u * u + v * v + w * w);
  const double p = rho * T / ph.gM2;
  W[WU * F + n] = u;
  W[WV * F + n] = v;
  W[WW * F + n] = w;
  W[WE * F + n] = E;
  W[WT * F + n] = T;
  W[WP * F + n] = p;
  W[WMU * F + n] = pow(T, 0.7);
  if (check && (rho < 0.0 || T < 0.0)) {
    unsigned long long lin =
        ((unsigned long long)kk * n2 + (unsigned long long)(jj + 1)) * n1 + (ii + 1);
    atomicMin(blow, lin);
    atomicCAS(halt, 0, step);
  }
}

// Wall values (main_v1.9.f90:622-641) on both walls; rw receives the wall
// density set from the previous wall pressure.  A kernel of its own: the
// pressure extrapolation below reads the wall values of the neighbours.
__global__ void __launch_bounds__(256) k_wallval_strict(Geom g, Phys ph, double* __restrict__ W,
                                                        double* __restrict__ rw, int check,
                                                        int* halt, unsigned long long* blow,
                                                        int step, int n1, int n2) {
  const int ii = blockIdx.x * blockDim.x + threadIdx.x;
  const int jj = blockIdx.y * blockDim.y + threadIdx.y;
  if (ii >= g.P1 || jj >= g.P2) return;
  if (halted(halt, step)) return;
  const long long F = g.field, PL = g.plane;
  const int N = g.n3 - 1;
  for (int s = 0; s < 2; s++) {
    const long long n = idx(g, ii, jj, s ? N : 0);
    W[WU * F + n] = 0.0;
    W[WV * F + n] = 0.0;
    W[WW * F + n] = 0.0;
    W[WT * F + n] = 1.0;
    W[WE * F + n] = 1.0 + ph.ggM * 0.5 * (0.0 * 0.0 + 0.0 * 0.0 + 0.0 * 0.0);
    W[WMU * F + n] = pow(1.0, 0.7);  // mu = T**0.7 (main_v1.9.f90:643)
    const double r = ph.gM2 * W[WP * F + n] / 1.0;
    rw[s * PL + idx(g, ii, jj, 0)] = r;
    if (check && r < 0.0) {
      unsigned long long lin =
          ((unsigned long long)(s ? N : 0) * n2 + (unsigned long long)(jj + 1)) * n1 + (ii + 1);
      atomicMin(blow, lin);
      atomicCAS(halt, 0, step);
    }
  }
}

// Wall-pressure extrapolation from the z-momentum equation at an impervious
// wall (main_v1.9.f90:645-716).
__global__ void __launch_bounds__(256) k_wallp_strict(Geom g, Phys ph, const double* __restrict__ U,
                                                      double* __restrict__ W,
                                                      const double* __restrict__ rw,
                                                      const int* halt, int step) {
  const int ii = blockIdx.x * blockDim.x + threadIdx.x;
  const int jj = blockIdx.y * blockDim.y + threadIdx.y;
  if (ii >= g.P1 || jj >= g.P2) return;
  if (halted(halt, step)) return;
  const long long F = g.field, PL = g.plane;
  const int N = g.n3 - 1;
  auto Wc = [&](int c, int i, int j, int k) { return W[c * F + idx(g, i, j, k)]; };
  const int im = wrapi(ii - 1, g.P1), ip = wrapi(ii + 1, g.P1);
  const int jm = wrapi(jj - 1, g.P2), jp = wrapi(jj + 1, g.P2);
  const long long nb = idx(g, ii, jj, 0), nt = idx(g, ii, jj, N);
  const double r1b = rw[idx(g, ii, jj, 0)], r1t = rw[PL + idx(g, ii, jj, 0)];
  // bottom
  {
    double dudxdz = ((ph.f1d2c0 * Wc(WMU, ip, jj, 0) * Wc(WU, ip, jj, 0) +
                      ph.f1d2c1 * Wc(WMU, ip, jj, 1) * Wc(WU, ip, jj, 1) +
                      ph.f1d2c2 * Wc(WMU, ip, jj, 2) * Wc(WU, ip, jj, 2)) -
                     (ph.f1d2c0 * Wc(WMU, im, jj, 0) * Wc(WU, im, jj, 0) +
                      ph.f1d2c1 * Wc(WMU, im, jj, 1) * Wc(WU, im, jj, 1) +
                      ph.f1d2c2 * Wc(WMU, im, jj, 2) * Wc(WU, im, jj, 2))) *
                    0.5 * ph.invdx;
    double dvdydz = ((ph.f1d2c0 * Wc(WMU, ii, jp, 0) * Wc(WV, ii, jp, 0) +
                      ph.f1d2c1 * Wc(WMU, ii, jp, 1) * Wc(WV, ii, jp, 1) +
                      ph.f1d2c2 * Wc(WMU, ii, jp, 2) * Wc(WV, ii, jp, 2)) -
                     (ph.f1d2c0 * Wc(WMU, ii, jm, 0) * Wc(WV, ii, jm, 0) +
                      ph.f1d2c1 * Wc(WMU, ii, jm, 1) * Wc(WV, ii, jm, 1) +
                      ph.f1d2c2 * Wc(WMU, ii, jm, 2) * Wc(WV, ii, jm, 2))) *
                    0.5 * ph.invdy;
    double d2wdz2 = ph.f2d1c0 * Wc(WMU, ii, jj, 0) * Wc(WW, ii, jj, 0) +
                    ph.f2d1c1 * Wc(WMU, ii, jj, 1) * Wc(WW, ii, jj, 1) +
                    ph.f2d1c2 * Wc(WMU, ii, jj, 2) * Wc(WW, ii, jj, 2);
    double w1 = Wc(WW, ii, jj, 0), w2 = Wc(WW, ii, jj, 1), w3 = Wc(WW, ii, jj, 2);
    double drhow2dz = ph.f1d2c0 * r1b * (w1 * w1) + ph.f1d2c1 * U[idx(g, ii, jj, 1)] * (w2 * w2) +
                      ph.f1d2c2 * U[idx(g, ii, jj, 2)] * (w3 * w3);
    double lhs = ph.invRe * (-dudxdz - dvdydz + kTwo3rd * (dudxdz + dvdydz - 2.0 * d2wdz2)) + drhow2dz;
    W[WP * F + nb] = -(lhs + ph.f1d2c1 * Wc(WP, ii, jj, 1) + ph.f1d2c2 * Wc(WP, ii, jj, 2)) / ph.f1d2c0;
  }
  // top
  {
    double dudxdz = ((ph.b1d2c0 * Wc(WMU, ip, jj, N) * Wc(WU, ip, jj, N) +
                      ph.b1d2c1 * Wc(WMU, ip, jj, N - 1) * Wc(WU, ip, jj, N - 1) +
                      ph.b1d2c2 * Wc(WMU, ip, jj, N - 2) * Wc(WU, ip, jj, N - 2)) -
                     (ph.b1d2c0 * Wc(WMU, im, jj, N) * Wc(WU, im, jj, N) +
                      ph.b1d2c1 * Wc(WMU, im, jj, N - 1) * Wc(WU, im, jj, N - 1) +
                      ph.b1d2c2 * Wc(WMU, im, jj, N - 2) * Wc(WU, im, jj, N - 2))) *
                    0.5 * ph.invdx;
    double dvdydz = ((ph.b1d2c0 * Wc(WMU, ii, jp, N) * Wc(WV, ii, jp, N) +
                      ph.b1d2c1 * Wc(WMU, ii, jp, N - 1) * Wc(WV, ii, jp, N - 1) +
                      ph.b1d2c2 * Wc(WMU, ii, jp, N - 2) * Wc(WV, ii, jp, N - 2)) -
                     (ph.b1d2c0 * Wc(WMU, ii, jm, N) * Wc(WV, ii, jm, N) +
                      ph.b1d2c1 * Wc(WMU, ii, jm, N - 1) * Wc(WV, ii, jm, N - 1) +
                      ph.b1d2c2 * Wc(WMU, ii, jm, N - 2) * Wc(WV, ii, jm, N - 2))) *
                    0.5 * ph.invdy;
    double d2wdz2 = ph.b2d1c0 * Wc(WMU, ii, jj, N) * Wc(WW, ii, jj, N) +
                    ph.b2d1c1 * Wc(WMU, ii, jj, N - 1) * Wc(WW, ii, jj, N - 1) +
                    ph.b2d1c2 * Wc(WMU, ii, jj, N - 2) * Wc(WW, ii, jj, N - 2);
    double w1 = Wc(WW, ii, jj, N), w2 = Wc(WW, ii, jj, N - 1), w3 = Wc(WW, ii, jj, N - 2);
    double drhow2dz = ph.b1d2c0 * r1t * (w1 * w1) +
                      ph.b1d2c1 * U[idx(g, ii, jj, N - 1)] * (w2 * w2) +
                      ph.b1d2c2 * U[idx(g, ii, jj, N - 2)] * (w3 * w3);
    double lhs = ph.invRe * (-dudxdz - dvdydz + kTwo3rd * (dudxdz + dvdydz - 2.0 * d2wdz2)) + drhow2dz;
    W[WP * F + nt] = -(lhs + ph.b1d2c1 * Wc(WP, ii, jj, N - 1) + ph.b1d2c2 * Wc(WP, ii, jj, N - 2)) / ph.b1d2c0;
  }
}

// ---- host launchers ----------------------------------------------------
void launch_stage_strict(const Geom& g, const Phys& ph, const double* Uin, const double* Ubase,
                         double* Uout, double* W, double* rw, const double* jac,
                         const double* invJac, const double* force, int stage, int* halt,
                         unsigned long long* blow, int step, int n1, int n2, hipStream_t s,
                         hipEvent_t ev0, hipEvent_t ev1) {
  dim3 blk(64, 4);
  dim3 grd((g.P1 + 63) / 64, (g.P2 + 3) / 4, g.n3 - 2);
  if (ev0) (void)hipEventRecord(ev0, s);
  hipLaunchKernelGGL(k_rhs_strict, grd, blk, 0, s, g, ph, Uin, Ubase, Uout, W, jac, invJac, force,
                     stage, halt, step);
  if (ev1) (void)hipEventRecord(ev1, s);
  hipLaunchKernelGGL(k_prim_strict, grd, blk, 0, s, g, ph, Uout, W, stage == 2, halt, blow, step,
                     n1, n2);
  dim3 g2((g.P1 + 63) / 64, (g.P2 + 3) / 4, 1);
  hipLaunchKernelGGL(k_wallval_strict, g2, blk, 0, s, g, ph, W, rw, stage == 2, halt, blow, step,
                     n1, n2);
  hipLaunchKernelGGL(k_wallp_strict, g2, blk, 0, s, g, ph, Uout, W, rw, halt, step);
}

}  // namespace cc
