// cc_common.hip -- state upload/download and the FP64 reductions that feed
// the bulk forcing (wall shear and plane-mean density), shared by every
// precision mode.  Compiled with -ffp-contract=off: the per-node primitive
// recovery here must give the same bits as the stage kernels' own.
#include <mutex>

#include "cc_device.h"
#include "cc_kernels.h"

namespace cc {

// Fortran var(i,j,k,c), 0-based indices.
__device__ __forceinline__ long long fidx(int n1, int n2, int n3, int i, int j, int k, int c) {
  return (((long long)c * n3 + k) * n2 + j) * n1 + i;
}

// Primitives of a conserved state (main_v1.9.f90:599-609).
struct Prim {
  double rho, u, v, w, E, T, p;
};
__device__ __forceinline__ Prim prim_of(const Phys& ph, double r, double ru, double rv, double rw,
                                        double rE) {
  Prim P;
  P.rho = r;
  P.u = ru / r;
  P.v = rv / r;
  P.w = rw / r;
  P.E = rE / r;
  P.T = P.E - ph.ggM * 0.5 * (P.u * P.u + P.v * P.v + P.w * P.w);
  P.p = P.rho * P.T / ph.gM2;
  return P;
}

// The 23 plane-statistics quantities of main_v1.9.f90:993-1009.
__device__ __forceinline__ void stat_terms(const Prim& P, double q[kNStat]) {
  const double v[7] = {P.rho, P.u, P.v, P.w, P.E, P.T, P.p};
  for (int c = 0; c < 7; c++) {
    q[c] = v[c];
    q[7 + c] = v[c] * v[c];
  }
  q[14] = P.rho * P.w;
  q[15] = P.rho * P.T;
  q[16] = P.u * P.v;
  q[17] = P.u * P.w;
  q[18] = P.v * P.w;
  q[19] = P.v * P.T;
  q[20] = P.T * P.w;
  q[21] = P.p * P.w;
  q[22] = P.p * P.T;
}

// Wall state after the boundary update (main_v1.9.f90:624-639): u=v=w=0
// (or the slip or study velocity), T=1 (or the study temperature),
// E=T+ggM/2|u|^2, rho = gama Ma^2 p / T.
__device__ __forceinline__ Prim wall_prim(const Phys& ph, double rho, double p, double u = 0.0,
                                          double v = 0.0, double T = 1.0) {
  Prim P;
  P.rho = rho;
  P.u = u;
  P.v = v;
  P.w = 0.0;
  P.T = T;
  P.E = T + ph.ggM * 0.5 * (u * u + v * v + 0.0 * 0.0);
  P.p = p;
  return P;
}

// State of a wall-plane duplicate node from the ring (study walls).
__device__ __forceinline__ Prim ring_prim(const Phys& ph, const double* rg, int NRG, int slot) {
  return wall_prim(ph, rg[3 * NRG + slot], rg[4 * NRG + slot], rg[slot], rg[NRG + slot],
                   rg[2 * NRG + slot]);
}

// Ring slot of the duplicate at Fortran 0-based (i, j) of a wall plane, or -1.
__device__ __forceinline__ int ring_slot(int n1, int n2, int P1, int P2, int i, int j) {
  const bool ei = (i == 0 || i == n1 - 1), ej = (j == 0 || j == n2 - 1);
  if (ei && ej) return 2 * P2 + 2 * P1 + (i == 0 ? 0 : 1) + (j == 0 ? 0 : 2);
  if (ei) return (i == 0 ? 0 : P2) + (j - 1);
  if (ej) return 2 * P2 + (j == 0 ? 0 : P1) + (i - 1);
  return -1;
}

// ---- upload: conserved state from var (main_v1.9.f90:284-294) -----------
__global__ void k_upload(Geom g, Phys, int n1, int n2, const double* __restrict__ V,
                         double* __restrict__ UA, double* __restrict__ UB, double* __restrict__ W,
                         WallP pw, double* __restrict__ rw) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.field) return;
  const int ii = (int)(t % g.P1);
  const int jj = (int)((t / g.P1) % g.P2);
  const int kk = (int)(t / g.plane);
  const int n3 = g.n3;
  double v[7];
  for (int c = 0; c < 7; c++) v[c] = V[fidx(n1, n2, n3, ii + 1, jj + 1, kk, c)];
  const long long F = g.field;
  UA[t] = v[0];
  UB[t] = v[0];
  for (int c = 1; c < 5; c++) {
    UA[c * F + t] = v[0] * v[c];
    UB[c * F + t] = v[0] * v[c];
  }
  if (W) {  // reference-order path keeps the file's primitives and mu = T**0.7
    W[0 * F + t] = v[1];
    W[1 * F + t] = v[2];
    W[2 * F + t] = v[3];
    W[3 * F + t] = v[4];
    W[4 * F + t] = v[5];
    W[5 * F + t] = v[6];
    W[6 * F + t] = pow(v[5], 0.7);
  }
  const long long n2d = (long long)jj * g.P1 + ii;
  if (kk == 0) {
    pw.p[n2d] = v[6];
    rw[n2d] = v[0];
  } else if (kk == n3 - 1) {
    pw.p[pw.top_off + n2d] = v[6];
    rw[g.plane + n2d] = v[0];
  }
}

// Constant part of the plane sums: nodes the reference never refreshes.
// Interior planes: the four corner columns (i,j) in {1,n1}x{1,n2}
// (main_v1.9.f90:589-596 never writes them).  Wall planes: the whole halo
// ring i in {1,n1} or j in {1,n2}, whose pressure is never recomputed
// (:645 loops over updated i,j only) and whose other values are set by the
// wall update (:624-639).
__global__ void k_consts(Geom g, Phys ph, int n1, int n2, const double* __restrict__ V,
                         double* __restrict__ C, int wflags) {
  const int kk = blockIdx.x * blockDim.x + threadIdx.x;
  const int n3 = g.n3;
  if (kk >= n3) return;
  double acc[kNStat];
  for (int q = 0; q < kNStat; q++) acc[q] = 0.0;
  double tq[kNStat];
  if (kk > 0 && kk < n3 - 1) {
    const int ci[2] = {0, n1 - 1}, cj[2] = {0, n2 - 1};
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++) {
        Prim P;
        P.rho = V[fidx(n1, n2, n3, ci[a], cj[b], kk, 0)];
        P.u = V[fidx(n1, n2, n3, ci[a], cj[b], kk, 1)];
        P.v = V[fidx(n1, n2, n3, ci[a], cj[b], kk, 2)];
        P.w = V[fidx(n1, n2, n3, ci[a], cj[b], kk, 3)];
        P.E = V[fidx(n1, n2, n3, ci[a], cj[b], kk, 4)];
        P.T = V[fidx(n1, n2, n3, ci[a], cj[b], kk, 5)];
        P.p = V[fidx(n1, n2, n3, ci[a], cj[b], kk, 6)];
        stat_terms(P, tq);
        for (int q = 0; q < kNStat; q++) acc[q] += tq[q];
      }
  } else if (ph.Ls == 0.0 && !((wflags >> (kk != 0)) & 1)) {
    // (slip walls: the ring holds periodic images; study walls: the ring
    // is updated every stage and summed by k_stats_part)
    for (int j = 0; j < n2; j++)
      for (int i = 0; i < n1; i++) {
        if (!(i == 0 || i == n1 - 1 || j == 0 || j == n2 - 1)) continue;
        const double p = V[fidx(n1, n2, n3, i, j, kk, 6)];
        Prim P = wall_prim(ph, ph.gM2 * p / 1.0, p);
        stat_terms(P, tq);
        for (int q = 0; q < kNStat; q++) acc[q] += tq[q];
      }
  }
  for (int q = 0; q < kNStat; q++) C[q * n3 + kk] = acc[q];
}

// Wall arrays of the fused path from the uploaded file: before the first
// wall update the wall planes hold the file's values (main_v1.9.f90:272-294:
// var as read, mu = T**0.7), duplicates included.  Threads [0, 2 plane) fill
// the own wall nodes, the rest the duplicate ring (u, v, T, rho, p).
__global__ void k_wall_init(Geom g, Phys ph, int n1, int n2, const double* __restrict__ V,
                            WallP pw, double* __restrict__ dw) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int n3 = g.n3, NRG = ring_size(g.P1, g.P2);
  if (t < 2 * g.plane) {
    const int top = (int)(t / g.plane);
    const long long n = t % g.plane;
    const int ii = (int)(n % g.P1), jj = (int)(n / g.P1), k = top ? n3 - 1 : 0;
    auto at = [&](int i, int j, int c) { return V[fidx(n1, n2, n3, i, j, k, c)]; };
    pw.uw[t] = at(ii + 1, jj + 1, 1);
    pw.vw[t] = at(ii + 1, jj + 1, 2);
    const double T = at(ii + 1, jj + 1, 5);
    pw.tw[t] = T;
    pw.mw[t] = pow(T, 0.7);
    // divergence correction of the edge nodes: file duplicate vs the
    // periodic image the stage kernel reads (see wall_node)
    double dc = 0.0;
    if (ii == 0) dc += (at(n1 - 2, jj + 1, 1) - at(0, jj + 1, 1)) * 0.5 * ph.invdx;
    if (ii == g.P1 - 1) dc += (at(n1 - 1, jj + 1, 1) - at(1, jj + 1, 1)) * 0.5 * ph.invdx;
    if (jj == 0) dc += (at(ii + 1, n2 - 2, 2) - at(ii + 1, 0, 2)) * 0.5 * ph.invdy;
    if (jj == g.P2 - 1) dc += (at(ii + 1, n2 - 1, 2) - at(ii + 1, 1, 2)) * 0.5 * ph.invdy;
    dw[t] = (2.0 / 3.0) * pow(T, 0.7) * ph.invRe * dc;  // as in wall_node
    return;
  }
  const long long q = t - 2 * g.plane;
  if (q >= 2LL * NRG) return;
  const int top = (int)(q / NRG), slot = (int)(q % NRG), k = top ? n3 - 1 : 0;
  int i, j;  // Fortran 0-based position of the slot
  if (slot < g.P2) { i = 0; j = slot + 1; }
  else if (slot < 2 * g.P2) { i = n1 - 1; j = slot - g.P2 + 1; }
  else if (slot < 2 * g.P2 + g.P1) { j = 0; i = slot - 2 * g.P2 + 1; }
  else if (slot < 2 * g.P2 + 2 * g.P1) { j = n2 - 1; i = slot - 2 * g.P2 - g.P1 + 1; }
  else {
    const int c = slot - 2 * g.P2 - 2 * g.P1;
    i = (c & 1) ? n1 - 1 : 0;
    j = (c & 2) ? n2 - 1 : 0;
  }
  double* rg = pw.ring + (long long)top * kRingFields * NRG;
  rg[slot] = V[fidx(n1, n2, n3, i, j, k, 1)];
  rg[NRG + slot] = V[fidx(n1, n2, n3, i, j, k, 2)];
  rg[2 * NRG + slot] = V[fidx(n1, n2, n3, i, j, k, 5)];
  rg[3 * NRG + slot] = V[fidx(n1, n2, n3, i, j, k, 0)];
  rg[4 * NRG + slot] = V[fidx(n1, n2, n3, i, j, k, 6)];
}

// ---- download: the reference's var(n1,n2,n3,7) with every halo value ----
// var(i,j,k,1..7) (Fortran 0-based i,j,k) as the reference holds it, from
// the device state: periodic images of the unique nodes, the stale corner
// columns of the interior planes (the file's values in V), the wall values
// and, on a wall, its duplicate ring.
__device__ __forceinline__ Prim node_var(const Geom& g, const Phys& ph, int n1, int n2,
                                         const double* __restrict__ U, const WallP& pw,
                                         const double* __restrict__ rw,
                                         const double* __restrict__ V, int i, int j, int k) {
  const int n3 = g.n3;
  const bool ei = (i == 0 || i == n1 - 1), ej = (j == 0 || j == n2 - 1);
  const bool wall = (k == 0 || k == n3 - 1);
  Prim P;
  if (!wall && ei && ej) {  // stale corner column
    P.rho = V[fidx(n1, n2, n3, i, j, k, 0)];
    P.u = V[fidx(n1, n2, n3, i, j, k, 1)];
    P.v = V[fidx(n1, n2, n3, i, j, k, 2)];
    P.w = V[fidx(n1, n2, n3, i, j, k, 3)];
    P.E = V[fidx(n1, n2, n3, i, j, k, 4)];
    P.T = V[fidx(n1, n2, n3, i, j, k, 5)];
    P.p = V[fidx(n1, n2, n3, i, j, k, 6)];
    return P;
  }
  const int top = k == n3 - 1;
  if (wall && (ei || ej) && ((pw.wflags >> top) & 1)) {
    const int NRG = ring_size(g.P1, g.P2);
    P = ring_prim(ph, pw.ring + (long long)top * kRingFields * NRG, NRG,
                  ring_slot(n1, n2, g.P1, g.P2, i, j));
  } else if (wall && (ei || ej) && ph.Ls == 0.0) {
    const double p = V[fidx(n1, n2, n3, i, j, k, 6)];
    P = wall_prim(ph, ph.gM2 * p / 1.0, p);
  } else {
    const int ii = (i == 0) ? g.P1 - 1 : (i == n1 - 1 ? 0 : i - 1);
    const int jj = (j == 0) ? g.P2 - 1 : (j == n2 - 1 ? 0 : j - 1);
    const long long n2d = (long long)jj * g.P1 + ii;
    if (wall) {
      const long long wo = (top ? g.plane : 0) + n2d;
      P = wall_prim(ph, rw[wo], pw.p[(top ? pw.top_off : 0) + n2d], pw.uw ? pw.uw[wo] : 0.0,
                    pw.vw ? pw.vw[wo] : 0.0, pw.tw ? pw.tw[wo] : 1.0);
    } else {
      const long long F = g.field, n = idx(g, ii, jj, k);
      P = prim_of(ph, U[n], U[F + n], U[2 * F + n], U[3 * F + n], U[4 * F + n]);
    }
  }
  return P;
}

__global__ void k_download(Geom g, Phys ph, int n1, int n2, const double* __restrict__ U,
                           WallP pw, const double* __restrict__ rw, const double* __restrict__ V,
                           double* __restrict__ Vout) {
  const long long total = (long long)n1 * n2 * g.n3;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int i = (int)(t % n1);
  const int j = (int)((t / n1) % n2);
  const int k = (int)(t / ((long long)n1 * n2));
  const Prim P = node_var(g, ph, n1, n2, U, pw, rw, V, i, j, k);
  const double v[7] = {P.rho, P.u, P.v, P.w, P.E, P.T, P.p};
  for (int c = 0; c < 7; c++) Vout[fidx(n1, n2, g.n3, i, j, k, c)] = v[c];
}

// ---- deterministic block reduction ---------------------------------------
template <int NV>
__device__ __forceinline__ void block_sum(double (&a)[NV], double* lds /* [nwaves][NV] */) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int q = 0; q < NV; q++) {
    double x = a[q];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    a[q] = x;
  }
  if (lane == 0)
    for (int q = 0; q < NV; q++) lds[wave * NV + q] = a[q];
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = blockDim.x >> 6;
    for (int q = 0; q < NV; q++) {
      double s = lds[q];
      for (int w = 1; w < nw; w++) s += lds[w * NV + q];
      a[q] = s;
    }
  }
}

// Wall shear (main_v1.9.f90:929-988) in two deterministic passes: blocks
// (b, wall) sum a contiguous range of the wall's nodes, then one thread per
// wall adds the kTauBlocks partials in order.
constexpr int kTauBlocks = 32;
__device__ __forceinline__ void tauwall_part(const Geom& g, const Phys& ph,
                                             const double* __restrict__ U, const WallP& pw,
                                             double* __restrict__ part, int blk, bool top,
                                             double* lds) {
  const long long F = g.field;
  const int N = g.n3 - 1;
  const long long n0 = g.plane * blk / kTauBlocks, n1e = g.plane * (blk + 1) / kTauBlocks;
  double acc[1] = {0.0};
  for (long long n = n0 + threadIdx.x; n < n1e; n += blockDim.x) {
    const int ii = (int)(n % g.P1), jj = (int)(n / g.P1);
    // wall values after the boundary update: w = 0, mu = T_w**0.7
    const double mu = pw.mw ? pw.mw[(top ? g.plane : 0) + n] : pow(1.0, 0.7);
    const double dwdx = mu * (0.0 - 0.0) * 0.5 * ph.invdx;
    const long long a = idx(g, ii, jj, top ? N - 1 : 1), b = idx(g, ii, jj, top ? N - 2 : 2);
    const double u2 = U[F + a] / U[a], u3 = U[F + b] / U[b];
    const double u1 = pw.uw ? pw.uw[(top ? g.plane : 0) + n] : 0.0;
    double dudz;
    if (!top)
      dudz = mu * (ph.f1d2c0 * u1 + ph.f1d2c1 * u2 + ph.f1d2c2 * u3);
    else
      dudz = mu * (ph.b1d2c0 * u1 + ph.b1d2c1 * u2 + ph.b1d2c2 * u3);
    acc[0] += (dwdx + dudz) * ph.invRe;
  }
  block_sum<1>(acc, lds);
  if (threadIdx.x == 0) part[(top ? kTauBlocks : 0) + blk] = acc[0];
}

// Plane statistics (main_v1.9.f90:993-1025) in two deterministic passes:
// blocks (s, kk) sum one of `split` node ranges of plane kk, then one thread
// per quantity and plane adds the partials in order.  Each unique interior node is
// weighted by the number of times the reference's all-node sum sees it (1 +
// an x duplicate + a y duplicate); the stale corners and wall rings come in
// through C, a study wall's duplicates through its ring.
constexpr int kStatSplitMax = 8;
__device__ __forceinline__ void planestats_part(const Geom& g, const Phys& ph,
                                                const double* __restrict__ U, const WallP& pw,
                                                const double* __restrict__ rw,
                                                double* __restrict__ part, int kk, int sp,
                                                int split, double* lds) {
  const int n3 = g.n3;
  const bool wall = (kk == 0 || kk == n3 - 1);
  const long long F = g.field;
  double acc[kNStat];
  for (int q = 0; q < kNStat; q++) acc[q] = 0.0;
  double tq[kNStat];
  // a plane holds fewer than 2^31 nodes (checked at create): 32-bit index
  // arithmetic
  const int n0 = (int)(g.plane * sp / split), n1e = (int)(g.plane * (sp + 1) / split);
  const int bd = (int)blockDim.x;
  if (wall) {
    for (int n = n0 + (int)threadIdx.x; n < n1e; n += bd) {
      const int jj = n / g.P1, ii = n - jj * g.P1;
      double wgt = 1.0;
      const bool top = kk == n3 - 1;
      const long long wo = (top ? g.plane : 0) + n;
      const Prim P = wall_prim(ph, rw[wo], pw.p[(top ? pw.top_off : 0) + n],
                               pw.uw ? pw.uw[wo] : 0.0, pw.vw ? pw.vw[wo] : 0.0,
                               pw.tw ? pw.tw[wo] : 1.0);
      if (ph.Ls != 0.0 && !((pw.wflags >> top) & 1))  // periodic-image ring: every duplicate counts
        wgt = ((ii == 0 || ii == g.P1 - 1) ? 2.0 : 1.0) * ((jj == 0 || jj == g.P2 - 1) ? 2.0 : 1.0);
      stat_terms(P, tq);
      for (int q = 0; q < kNStat; q++) acc[q] += wgt * tq[q];
    }
  } else {
    const long long base = (long long)kk * g.plane;
    for (int n = n0 + (int)threadIdx.x; n < n1e; n += bd) {
      const int jj = n / g.P1, ii = n - jj * g.P1;
      const long long m = base + n;
      const Prim P = prim_of(ph, U[m], U[F + m], U[2 * F + m], U[3 * F + m], U[4 * F + m]);
      const double wgt =
          1.0 + ((ii == 0 || ii == g.P1 - 1) ? 1.0 : 0.0) + ((jj == 0 || jj == g.P2 - 1) ? 1.0 : 0.0);
      stat_terms(P, tq);
      for (int q = 0; q < kNStat; q++) acc[q] += wgt * tq[q];
    }
  }
  if (wall && sp == 0 && ((pw.wflags >> (kk != 0)) & 1)) {
    // study wall: the duplicates as the last wall update left them
    const int NRG = ring_size(g.P1, g.P2);
    const double* rg = pw.ring + (long long)(kk != 0) * kRingFields * NRG;
    for (int q0 = threadIdx.x; q0 < NRG; q0 += blockDim.x) {
      stat_terms(ring_prim(ph, rg, NRG, q0), tq);
      for (int q = 0; q < kNStat; q++) acc[q] += tq[q];
    }
  }
  block_sum<kNStat>(acc, lds);
  if (threadIdx.x == 0)
    for (int q = 0; q < kNStat; q++) part[((long long)q * n3 + kk) * split + sp] = acc[q];
}

// First pass of the statistics step in one launch: blocks [0, split n3)
// take the plane statistics (block b: plane b / split, range b % split),
// the 2 kTauBlocks after them the wall shear (bottom wall first).
__global__ void __launch_bounds__(256) k_stats_part(Geom g, Phys ph, const double* __restrict__ U,
                                                    WallP pw, const double* __restrict__ rw,
                                                    double* __restrict__ part, int split,
                                                    const int* halt, int step) {
  __shared__ double lds[4 * kNStat];
  const int h = *halt;
  if (h != 0 && h <= step) return;
  const int b = blockIdx.x, nps = split * g.n3;
  if (b < nps) {
    planestats_part(g, ph, U, pw, rw, part, b / split, b % split, split, lds);
  } else {
    const int t = b - nps;
    tauwall_part(g, ph, U, pw, part + (long long)kNStat * g.n3 * split, t % kTauBlocks,
                 t >= kTauBlocks, lds);
  }
}

// Second pass: thread e < kNStat n3 adds the partials of quantity q = e / n3
// on plane kk = e % n3 in order (all of them loaded first: a loop of
// dependent loads made this pass latency-bound); threads kNStat n3 and
// kNStat n3 + 1 the two walls' shear partials.
__device__ __forceinline__ void stats_final_elem(int e, int n1, int n2, int n3,
                                                 const double* __restrict__ part, int split,
                                                 const double* __restrict__ C,
                                                 double* __restrict__ force,
                                                 double* __restrict__ stats) {
  if (e < kNStat * n3) {
    const double* pp = part + (long long)e * split;  // [q][kk][split] with e = q n3 + kk
    double v[kStatSplitMax];
#pragma unroll
    for (int sp = 0; sp < kStatSplitMax; sp++) v[sp] = sp < split ? pp[sp] : 0.0;
    double a = 0.0;
#pragma unroll
    for (int sp = 0; sp < kStatSplitMax; sp++)
      if (sp < split) a += v[sp];
    const double t = (a + C[e]) / (double)(n1 * n2);
    stats[e] = t;
    if (e < n3) force[2 + e] = t;  // q = 0: the mean density of plane e
  } else if (e < kNStat * n3 + 2) {
    const int w = e - kNStat * n3;
    const double* tp = part + (long long)kNStat * n3 * split;
    double s = 0.0;
    for (int b = 0; b < kTauBlocks; b++) s += tp[w * kTauBlocks + b];
    const double t = s / (double)((n1 - 2) * (n2 - 2));
    force[w] = t;
    stats[kNStat * n3 + w] = t;
  }
}

__global__ void k_stats_final(int n1, int n2, int n3, const double* __restrict__ part, int split,
                              const double* __restrict__ C, double* __restrict__ force,
                              double* __restrict__ stats, const int* halt, int step) {
  const int h = *halt;
  if (h != 0 && h <= step) return;
  stats_final_elem(blockIdx.x * blockDim.x + threadIdx.x, n1, n2, n3, part, split, C, force,
                   stats);
}

// The same two passes for the members of a batch, in one launch each: a
// block belongs to the last member whose first block it has reached (the
// members' blocks are contiguous), and runs that member's block of the
// single-instance pass, so every member's sums are the single launch's.
__device__ __forceinline__ int stats_member(const StatDesc* __restrict__ d, int nm, bool fin) {
  const int b = (int)blockIdx.x;
  int m = 0;
  for (int i = 1; i < nm; i++)
    if ((fin ? d[i].fblk0 : d[i].blk0) <= b) m = i;
  return __builtin_amdgcn_readfirstlane(m);
}

__global__ void __launch_bounds__(256) k_stats_part_batched(const StatDesc* __restrict__ d,
                                                            int nm, int step) {
  __shared__ double lds[4 * kNStat];
  const StatDesc& D = d[stats_member(d, nm, false)];
  const int h = *D.halt;
  if (h != 0 && h <= step) return;
  const int b = (int)blockIdx.x - D.blk0, split = D.split, nps = split * D.g.n3;
  if (b < nps) {
    planestats_part(D.g, D.ph, D.U, D.pw, D.rw, D.part, b / split, b % split, split, lds);
  } else {
    const int t = b - nps;
    tauwall_part(D.g, D.ph, D.U, D.pw, D.part + (long long)kNStat * D.g.n3 * split,
                 t % kTauBlocks, t >= kTauBlocks, lds);
  }
}

__global__ void __launch_bounds__(256) k_stats_final_batched(const StatDesc* __restrict__ d,
                                                             int nm, int step) {
  const StatDesc& D = d[stats_member(d, nm, true)];
  const int h = *D.halt;
  if (h != 0 && h <= step) return;
  stats_final_elem(((int)blockIdx.x - D.fblk0) * 256 + (int)threadIdx.x, D.n1, D.n2, D.g.n3,
                   D.part, D.split, D.C, D.force, D.stats);
}

// ---- bulk diagnostics (main_v1.9.f90:799-906) --------------------------------
// The diagnostics read var(i,j,k,c) of every node.  Two sources give the same
// values: the reconstructed array (VarArray, what cc_download returns) or the
// device state directly (VarState: node_var per node, no 7-field copy).
struct VarArray {
  const double* V;
  long long total;
  __device__ void uvw(long long t, double& u, double& v, double& w) const {
    u = V[total + t];
    v = V[2 * total + t];
    w = V[3 * total + t];
  }
  __device__ void ru(long long t, double& r, double& u) const {
    r = V[t];
    u = V[total + t];
  }
};
struct VarState {
  Geom g;
  Phys ph;
  int n1, n2;
  const double* U;
  WallP pw;
  const double* rw;
  const double* V;
  __device__ Prim at(long long t) const {
    const int i = (int)(t % n1);
    const int j = (int)((t / n1) % n2);
    const int k = (int)(t / ((long long)n1 * n2));
    return node_var(g, ph, n1, n2, U, pw, rw, V, i, j, k);
  }
  __device__ void uvw(long long t, double& u, double& v, double& w) const {
    const Prim P = at(t);
    u = P.u;
    v = P.v;
    w = P.w;
  }
  __device__ void ru(long long t, double& r, double& u) const {
    const Prim P = at(t);
    r = P.rho;
    u = P.u;
  }
};

// Pass 1: per-block partial sums of u^2+v^2+w^2 over every node (halo values
// included, as the reference's all-node loop).
template <typename A>
__global__ void __launch_bounds__(256) k_ke_partial(long long total, A var,
                                                    double* __restrict__ part) {
  __shared__ double lds[4];
  double acc[1] = {0.0};
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    double u, v, w;
    var.uvw(t, u, v, w);
    acc[0] += u * u + v * v + w * w;
  }
  block_sum<1>(acc, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
}

// Pass 2 (one block): KE from the partials in a fixed order, then the three
// bulk plane integrals for each of the five x-planes with the trapezoidal
// quarter-sum of main_v1.9.f90:827-834.
template <typename A>
__global__ void __launch_bounds__(256) k_diag_final(int n1, int n2, int n3, double dy, A var,
                                                    const double* __restrict__ part, int npart,
                                                    const double* __restrict__ x3f, double x2e,
                                                    double* __restrict__ out) {
  __shared__ double lds[4];
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int b = 0; b < npart; b++) s += part[b];
    out[0] = s / (double)(n1 * n2 * n3);
  }
  const int plane[5] = {5, 10, 70, 100, 125};
  const double Ayz = x3f[n3 - 1] * x2e;
  for (int q = 0; q < 3; q++)
    for (int m = 0; m < 5; m++) {
      const int p = plane[m];  // 1-based x index
      double acc[1] = {0.0};
      if (p > 2 && p < n1 - 1) {
        const int i = p - 1;
        const long long nj = n2 - 1, nk = n3 - 1;
        for (long long t = threadIdx.x; t < nj * nk; t += blockDim.x) {
          const int j = (int)(t % nj), k = (int)(t / nj);
          const double dAyz = dy * (x3f[k + 1] - x3f[k]);
          auto val = [&](int jj, int kk) {
            const long long n = ((long long)kk * n2 + jj) * n1 + i;
            double r, u;
            var.ru(n, r, u);
            if (q == 0) return r;
            if (q == 1) return u;
            return r * u;
          };
          acc[0] += 0.25 * (val(j, k) + val(j + 1, k) + val(j, k + 1) + val(j + 1, k + 1)) * dAyz;
        }
      }
      block_sum<1>(acc, lds);
      if (threadIdx.x == 0) out[1 + 5 * q + m] = acc[0] / Ayz;
      __syncthreads();
    }
}

void launch_diag(int n1, int n2, int n3, double dy, const double* Vdev, double* part,
                 const double* x3f, double x2e, double* out, hipStream_t s) {
  const long long total = (long long)n1 * n2 * n3;
  const VarArray A{Vdev, total};
  hipLaunchKernelGGL(k_ke_partial<VarArray>, dim3(kDiagParts), dim3(256), 0, s, total, A, part);
  hipLaunchKernelGGL(k_diag_final<VarArray>, dim3(1), dim3(256), 0, s, n1, n2, n3, dy, A, part,
                     kDiagParts, x3f, x2e, out);
}

// ---- the allocation-free diagnostics' scratch --------------------------------
// One static device array per GPU the library runs on, held until the process
// ends: kDiagParts partial sums, the 16 results and the file's x3 (n3 <=
// kDiagScratchMaxN3).  Every call on a GPU shares it; each call writes every
// element it reads, so no clearing is needed, and calls that overlap on one
// GPU race.  In host code the array's name is a host-side stand-in, not a
// device address: only the pointer hipGetSymbolAddress returns for a device,
// with that device current, is passed to kernels and copies.
static __device__ __align__(256) double scratch[kDiagScratchSize];

namespace {
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

int diag_scratch_init(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return -1;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[dev]) return 0;
  int cur = 0;
  if (hipGetDevice(&cur) != hipSuccess) return -1;
  if (cur != dev && hipSetDevice(dev) != hipSuccess) return -1;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (cur != dev) (void)hipSetDevice(cur);
  if (e != hipSuccess || !p) return -1;
  g_scratch_ptr[dev] = (double*)p;
  return 0;
}

double* diag_scratch(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch_ptr[dev];
}

void launch_diag_state(const Geom& g, const Phys& ph, int n1, int n2, const double* U, WallP pw,
                       const double* rw, const double* Vfile, double* part, const double* x3f,
                       double x2e, double* out, hipStream_t s) {
  const long long total = (long long)n1 * n2 * g.n3;
  const VarState A{g, ph, n1, n2, U, pw, rw, Vfile};
  hipLaunchKernelGGL(k_ke_partial<VarState>, dim3(kDiagParts), dim3(256), 0, s, total, A, part);
  hipLaunchKernelGGL(k_diag_final<VarState>, dim3(1), dim3(256), 0, s, n1, n2, g.n3, ph.dy, A,
                     part, kDiagParts, x3f, x2e, out);
}

void launch_upload(const Geom& g, const Phys& ph, int n1, int n2, const double* Vfull, double* UA,
                   double* UB, double* W, WallP pw, double* rw, double* dw, double* consts,
                   hipStream_t s) {
  const int B = 256;
  hipLaunchKernelGGL(k_upload, dim3((unsigned)((g.field + B - 1) / B)), dim3(B), 0, s, g, ph, n1,
                     n2, Vfull, UA, UB, W, pw, rw);
  launch_wall_init(g, ph, n1, n2, Vfull, pw, dw, consts, s);
}

void launch_wall_init(const Geom& g, const Phys& ph, int n1, int n2, const double* Vfull,
                      WallP pw, double* dw, double* consts, hipStream_t s) {
  if (pw.uw) {
    const long long nw = 2 * g.plane + 2LL * ring_size(g.P1, g.P2);
    hipLaunchKernelGGL(k_wall_init, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, s, g, ph,
                       n1, n2, Vfull, pw, dw);
  }
  hipLaunchKernelGGL(k_consts, dim3((g.n3 + 63) / 64), dim3(64), 0, s, g, ph, n1, n2, Vfull,
                     consts, pw.wflags);
}

void launch_download(const Geom& g, const Phys& ph, int n1, int n2, const double* U, WallP pw,
                     const double* rw, const double* Vfull, double* Vout, hipStream_t s) {
  const long long total = (long long)n1 * n2 * g.n3;
  const int B = 256;
  hipLaunchKernelGGL(k_download, dim3((unsigned)((total + B - 1) / B)), dim3(B), 0, s, g, ph, n1,
                     n2, U, pw, rw, Vfull, Vout);
}

int stats_scratch_size(int n3) { return kNStat * n3 * kStatSplitMax + 2 * kTauBlocks; }

// Node ranges per plane of the statistics pass: as many as let the whole
// first pass (split n3 + 2 kTauBlocks blocks) run in one round on the
// current device, at most kStatSplitMax (a second, partly filled round of
// blocks cost more than the longer ranges).
static int stats_split(int n3) {
  static int resident[kMaxDevices] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) dev = 0;
  if (resident[dev] == 0) {
    int ncu = 0, bpc = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        ncu <= 0)
      ncu = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, k_stats_part, 256, 0) != hipSuccess ||
        bpc <= 0)
      bpc = 1;
    resident[dev] = ncu * bpc;
  }
  const int sp = (resident[dev] - 2 * kTauBlocks) / n3;
  return sp < 1 ? 1 : (sp > kStatSplitMax ? kStatSplitMax : sp);
}

StatDesc stats_desc(const Geom& g, const Phys& ph, int n1, int n2, const double* U, WallP pw,
                    const double* rw, const double* consts, double* force, double* stats,
                    double* part, const int* halt) {
  StatDesc d{};
  d.g = g;
  d.ph = ph;
  d.n1 = n1;
  d.n2 = n2;
  d.split = stats_split(g.n3);
  d.U = U;
  d.pw = pw;
  d.rw = rw;
  d.C = consts;
  d.force = force;
  d.stats = stats;
  d.part = part;
  d.halt = halt;
  return d;
}
int stats_part_blocks(const StatDesc& d) { return d.split * d.g.n3 + 2 * kTauBlocks; }
int stats_final_blocks(const StatDesc& d) { return (kNStat * d.g.n3 + 2 + 255) / 256; }

void launch_stats_batched(const StatDesc* d_desc, int nm, int part_blocks, int final_blocks,
                          int step, hipStream_t s) {
  hipLaunchKernelGGL(k_stats_part_batched, dim3((unsigned)part_blocks), dim3(256), 0, s, d_desc,
                     nm, step);
  hipLaunchKernelGGL(k_stats_final_batched, dim3((unsigned)final_blocks), dim3(256), 0, s,
                     d_desc, nm, step);
}

void launch_stats(const Geom& g, const Phys& ph, int n1, int n2, const double* U, WallP pw,
                  const double* rw, const double* consts, double* force, double* stats,
                  double* part, const int* halt, int step, hipStream_t s) {
  // stats layout: [23][n3] plane statistics, then tau[2]; part: the plane
  // partials [23][n3][split], then the wall-shear partials [2][kTauBlocks]
  const int split = stats_split(g.n3);
  hipLaunchKernelGGL(k_stats_part, dim3((unsigned)(split * g.n3 + 2 * kTauBlocks)), dim3(256), 0,
                     s, g, ph, U, pw, rw, part, split, halt, step);
  hipLaunchKernelGGL(k_stats_final, dim3((kNStat * g.n3 + 2 + 255) / 256), dim3(256), 0, s, n1,
                     n2, g.n3, part, split, consts, force, stats, halt, step);
}

}  // namespace cc
