// cc_fused.hip -- one fused kernel per RK stage: primitive recovery, hybrid
// central/QUICK convective flux, viscous/pressure point fluxes and their
// divergence, forcing source and RK update, streaming along the wall-normal
// direction (2.5-D blocking).
//
// Math is the reference's (v1.10/codes/main_v1.9.f90:329-609); the
// evaluation order is free (FMA contraction allowed), so FP64 results agree
// with the reference-order path to rounding, not bit for bit.  With
// T = float the flux and stencil arithmetic is FP32 while the conserved
// state U^n, the RK accumulation and every reduction stay FP64; the stage-1
// predictor U*, whose only use is stage 2's FP32 stencil input, is stored
// as FP32 (a kernel per stage).
//
// Work decomposition.  A block owns a TX x TY tile of (ii, jj) columns and a
// chunk of output planes [kb, ke).  It walks planes t = kb-2 .. ke+1; at each
// t it
//   1. loads U(t) on the tile plus a 2-node x/y halo, recovers the
//      primitives once per loaded node and stages (u,v,w,T), (p,mu) and U in
//      LDS (two plane slots, ping-pong);
//   2. computes the x and y convective fluxes of plane t;
//   3. computes the point fluxes fD, gD (tile + 1-node halo) and hD (tile)
//      of plane t-1, then the x and y divergences of plane t-1, and adds
//      hD(t-1) into the z divergences of planes t-2 and t;
//   4. completes plane t-2 with its z convective flux and z divergence
//      (register windows along the column), adds the source and writes the
//      RK update.
// HBM traffic per node and stage: U in (40 B), U out (40 B), plus the base
// state in stage 2 (40 B) and the x/y halo re-reads, which hit L2; mixed
// moves U* as 20 B (stage 1 out, stage 2 in).
//
// Latency.  Phases 2 and 3 share no barrier (fD/gD have their own LDS).
// The per-plane metric and forcing values are scalar loads (cread), so no
// vector-load wait drains the state loads in flight.  FP32 prefetches the
// next plane into registers after phase 3 and keeps its z window in FP32,
// which holds it to 168 VGPRs and three waves per SIMD; in stage 2 it loads
// the base state U^n of phase 4 after the phase-1 barrier.  FP64 warms the
// next plane in L2 with 4-byte loads whose values are discarded, one per
// cache line, during phases 2-3.  Stage 1 (one-instance kernel) then loads it into registers
// just before phase 4, which needs no base state from memory; stage 2, out
// of registers, reads it from L2 in phase 1, and warms U^n the same way
// and loads it after the phase-3 barrier, so that phase 3b hides the L2
// hit.  The batched kernel reads the next plane from L2 in both stages.
#include <type_traits>

#include "cc_device.h"
#include "cc_kernels.h"

namespace cc {

// The convective flux difference at a node is linear in the five stencil
// values (qmm, qm, qc, qp, qpp) along the direction, with weights that
// depend only on the velocities, so the weights are formed once per node and
// every field costs five FMAs.
//
// QUICK (main_v1.9.f90:335-358), scaled by s: ue fe - uw fw, where the east
// face takes its upwind-biased value from the side ue blows from,
//   fe = 6 qc + 3 qp - qm (ue >= 0), 6 qp + 3 qc - qpp (ue < 0),
// and the west face from the side uw blows from,
//   fw = 6 qm + 3 qc - qmm (uw >= 0), 6 qc + 3 qm - qp (uw < 0).
template <typename T>
__device__ __forceinline__ void quick_w(T ue, T uw, T s, T (&c)[5]) {
  const bool ep = ue >= T(0), wp = uw >= T(0);
  const T e = ue * s, w = uw * s;
  c[0] = wp ? w : T(0);
  c[1] = (ep ? -e : T(0)) - (wp ? T(6) * w : T(3) * w);
  c[2] = (ep ? T(6) * e : T(3) * e) - (wp ? T(3) * w : T(6) * w);
  c[3] = (ep ? T(3) * e : T(6) * e) + (wp ? T(0) : w);
  c[4] = ep ? T(0) : -e;
}
// fourth-order central (main_v1.9.f90:360-366), scaled by k:
//   -vpp qpp + 8 vp qp - 8 vm qm + vmm qmm, with the velocities vmm..vpp at
//   the stencil points
template <typename T>
__device__ __forceinline__ void central_w(T vmm, T vm, T vp, T vpp, T k, T (&c)[5]) {
  c[0] = vmm * k;
  c[1] = T(-8) * vm * k;
  c[2] = T(0);
  c[3] = T(8) * vp * k;
  c[4] = -vpp * k;
}
template <typename T>
__device__ __forceinline__ T apply_w(const T (&c)[5], T qmm, T qm, T qc, T qp, T qpp, T acc) {
  acc = fma(c[0], qmm, acc);
  acc = fma(c[1], qm, acc);
  acc = fma(c[2], qc, acc);
  acc = fma(c[3], qp, acc);
  return fma(c[4], qpp, acc);
}

template <typename T>
struct Deriv {
  T dudx, dudy, dudz, dvdx, dvdy, dvdz, dwdx, dwdy, dwdz, dTdx, dTdy, dTdz;
};

// 1/x for the positive normal values of the fused path (density, 1 + m):
// the hardware reciprocal refined by two Newton steps, within 1 ulp of the
// quotient for double (an IEEE division costs ~10 instructions); the
// hardware reciprocal alone (1 ulp) for float.
template <typename T>
__device__ __forceinline__ T rcp(T x);
template <>
__device__ __forceinline__ double rcp<double>(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}
template <>
__device__ __forceinline__ float rcp<float>(float x) {
  return __builtin_amdgcn_rcpf(x);
}

// T**0.7 for T > 0 (mu = kt = T**0.7, main_v1.9.f90:643,720).
// Double, T in [2^-100, 2^100]: the hardware FP32 log2/exp2 give a seed y0
// within a few FP32 ulp; one Halley step on f(y) = y^10 - T^7 (0.7 = 7/10),
//   y1 = y0 (9 y0^10 + 11 T^7) / (11 y0^10 + 9 T^7),
// converges cubically (relative error ~8 e0^3, below 1e-19), leaving the
// rounding of the powers and the quotient: within 5 ulp of the correctly
// rounded T**0.7 (2 ulp for 96 % of arguments), which the 1e-12 FP64
// tolerance absorbs.  It costs ~20 instructions where a log/exp series in
// FP64 costs ~80.  Outside that range the series below (2 ulp) serves all
// positive finite T; 0 and +inf return themselves (0**0.7 = 0,
// inf**0.7 = inf), negative T and NaN give NaN, as the reference's
// T**0.7d0, so that a negative temperature propagates NaN.
// Series: x = m 2^e with m in [sqrt(1/2), sqrt(2)); ln m = 2 atanh(s),
// s = (m-1)/(m+1), |s| < 0.172, by its odd series to s^21; then
// 2^(0.7 log2 x) = 2^n exp(f ln 2) with n = rint(.), |f| <= 1/2 and the
// exponential by its Taylor series to degree 14.
// Float: the hardware log2/exp2 (v_log_f32, v_exp_f32), which give the
// special values themselves.  The library's log2f/exp2f wrap the same
// instructions in rescaling for denormal inputs and results (12 more
// instructions), which a temperature never reaches.
__device__ __noinline__ double pow07d_series(double x) {
  int e;
  double m = frexp(x, &e);
  if (m < 0.70710678118654752440) {
    m *= 2.0;
    e -= 1;
  }
  const double s = (m - 1.0) * rcp<double>(m + 1.0);
  const double s2 = s * s;
  double p = 1.0 / 21.0;
  p = fma(p, s2, 1.0 / 19.0);
  p = fma(p, s2, 1.0 / 17.0);
  p = fma(p, s2, 1.0 / 15.0);
  p = fma(p, s2, 1.0 / 13.0);
  p = fma(p, s2, 1.0 / 11.0);
  p = fma(p, s2, 1.0 / 9.0);
  p = fma(p, s2, 1.0 / 7.0);
  p = fma(p, s2, 1.0 / 5.0);
  p = fma(p, s2, 1.0 / 3.0);
  const double lnm = fma(2.0 * s * s2, p, 2.0 * s);
  constexpr double kLn2 = 0.69314718055994530942;
  const double y = fma(0.7 / kLn2, lnm, 0.7 * (double)e);
  const double n = rint(y);
  const double g = (y - n) * kLn2;
  double q = 1.0 / 87178291200.0;  // 1/14!
  q = fma(q, g, 1.0 / 6227020800.0);
  q = fma(q, g, 1.0 / 479001600.0);
  q = fma(q, g, 1.0 / 39916800.0);
  q = fma(q, g, 1.0 / 3628800.0);
  q = fma(q, g, 1.0 / 362880.0);
  q = fma(q, g, 1.0 / 40320.0);
  q = fma(q, g, 1.0 / 5040.0);
  q = fma(q, g, 1.0 / 720.0);
  q = fma(q, g, 1.0 / 120.0);
  q = fma(q, g, 1.0 / 24.0);
  q = fma(q, g, 1.0 / 6.0);
  q = fma(q, g, 0.5);
  q = fma(q, g, 1.0);
  q = fma(q, g, 1.0);
  const double r = ldexp(q, (int)n);
  if (x > 0.0 && x < __builtin_huge_val()) return r;
  return (x == 0.0 || x > 0.0) ? x : __builtin_nan("");
}
__device__ __forceinline__ double pow07d(double x) {
  if (!(x >= 0x1p-100 && x <= 0x1p100)) return pow07d_series(x);  // NaN takes this branch
  const double y = (double)__builtin_amdgcn_exp2f(0.7f * __builtin_amdgcn_logf((float)x));
  const double x2 = x * x, c = (x2 * x2) * (x2 * x);  // x^7
  const double y2 = y * y, y4 = y2 * y2, y10 = (y4 * y4) * y2;
  const double num = fma(9.0, y10, 11.0 * c), den = fma(11.0, y10, 9.0 * c);
  return y * num * rcp<double>(den);
}

template <typename T>
__device__ __forceinline__ T pow07(T x);
template <>
__device__ __forceinline__ double pow07<double>(double x) {
  return pow07d(x);
}
template <>
__device__ __forceinline__ float pow07<float>(float x) {
  return __builtin_amdgcn_exp2f(0.7f * __builtin_amdgcn_logf(x));
}

// Element of a plane array: a wave-uniform base (the field and plane) plus
// a 32-bit byte offset (the node within the plane).  The access then takes
// its base from SGPRs and its offset from one VGPR (the saddr form), instead
// of a 64-bit address per access built by 64-bit vector adds.
// With kOpaque the base is held opaque in SGPRs as a global-address-space
// pointer, so that the compiler neither rebuilds the five fields' addresses
// from the first one with vector adds nor falls back to flat addressing.
// The FP64 kernels, short of SGPRs for ten such bases (they spill them to
// VGPR lanes), index plainly by the element offset o.
template <typename V>
using gptr = __attribute__((address_space(1))) V*;
template <bool kOpaque, typename V>
__device__ __forceinline__ gptr<const V> at_b(const V* base, int o, unsigned boff) {
  if constexpr (kOpaque) {
    auto b = (gptr<const char>)base;
    asm volatile("" : "+s"(b));
    return (gptr<const V>)(b + boff);
  } else {
    return (gptr<const V>)(base + o);
  }
}
template <bool kOpaque, typename V>
__device__ __forceinline__ gptr<V> at_b(V* base, int o, unsigned boff) {
  if constexpr (kOpaque) {
    auto b = (gptr<char>)base;
    asm volatile("" : "+s"(b));
    return (gptr<V>)(b + boff);
  } else {
    return (gptr<V>)(base + o);
  }
}

// I is Inst (a kernel argument) or Inst in the constant address space (a
// batched launch's descriptor): either way its fields are scalar loads.
// S: the stage (1, 2) as a constant, or 0 for the argument stage_arg (FP64).
template <typename T, int TX, int TY, int S, typename I>
__device__ __forceinline__ void stage_tile(I& a, const int stage_arg, const int step,
                                           const int tile_x, const int tile_y, const int zc) {
  constexpr int NT = TX * TY;
  constexpr int RX = TX + 4, RY = TY + 4, NR = RX * RY;
  constexpr int NRING = NR - NT;                // halo-2 ring points
  constexpr int NFX = (TX + 2) * TY;            // fD points (tile + x halo 1)
  constexpr int NGY = TX * (TY + 2);            // gD points (tile + y halo 1)
  constexpr int NZR = 2 * TY + 2 * TX;          // x- and y- ring-1 points
  static_assert(NRING <= NT, "one ring node per thread");
  // first threads of the y-ring (gD) and x-ring (fD) point-flux passes
  constexpr int kRingG0 = 64;
  constexpr int kRingF0 = 64 + ((2 * TX + 63) / 64) * 64;
  static_assert(kRingF0 + 2 * TY <= NT, "ring passes do not fit the block");

  __shared__ T sl[2][4][NR];    // u, v, w, T of two planes
  // p and mu are read only at the node whose point fluxes are computed:
  // the own nodes and the ring-1 points, two planes each
  __shared__ T spo[2][2][NT];   // p, mu of the own nodes
  __shared__ T spr[2][2][NZR];  // p, mu on the ring-1 points
  __shared__ T sub[5 * NR];     // U(5) of plane t (phase 2)
  __shared__ T sfg[4 * (NFX + NGY)];  // fD and gD of plane t-1 (phases 3, 3b)
  __shared__ T srz[3][3][NZR];  // u, v, w on the ring-1 columns, three planes

  const auto& g = a.g;
  const auto& ph = a.ph;
  {
    // a halt set during this launch holds this step, so every thread of the
    // block takes the same branch here
    const int h = *a.halt;
    if (h != 0 && h < step) return;
  }
  static_assert(sizeof(T) == 8 || S != 0, "the FP32 kernels take the stage as a constant");
  const int stage = S ? S : stage_arg;
  // Mixed precision keeps the stage-1 predictor U* in UB as FP32: stage 1
  // writes it, stage 2 reads it.  Its only use is the FP32 stencil input of
  // stage 2, which would round it to FP32 on loading anyway.
  constexpr bool kIn32 = sizeof(T) == 4 && S == 2, kOut32 = sizeof(T) == 4 && S == 1;
  using UIn = std::conditional_t<kIn32, float, double>;
  const int tid = threadIdx.x;
  const int tx = tid % TX, ty = tid / TX;
  const UIn* __restrict__ Uin = (const UIn*)(stage == 1 ? a.UA : a.UB);
  const double* Ubase = a.UA;
  double* Uout = stage == 1 ? a.UB : a.UA;
  const int i0 = tile_x * TX, j0 = tile_y * TY;
  const int kb = 1 + zc * a.kchunk;
  const int ke = min(kb + a.kchunk, g.n3 - 1);
  if (kb >= ke) return;
  const int P1 = g.P1, P2 = g.P2, n3 = g.n3;
  const long long F = g.field, PL = g.plane;
  const bool own_valid = (i0 + tx < P1) && (j0 + ty < P2);

  auto wrapP = [](int x, int P) {
    x %= P;
    return x < 0 ? x + P : x;
  };
  // region index of the thread's own node and its up to two ring nodes
  const int r_own = (ty + 2) * RX + (tx + 2);
  const int gi_own = wrapP(i0 + tx, P1), gj_own = wrapP(j0 + ty, P2);
  const int o_own = gj_own * P1 + gi_own;  // offset within a plane
  // the thread's ring node: with NRING <= NT, threads tid < NRING own one
  const bool has_ring = tid < NRING;
  int r_ring = 0, o_ring = 0;
  if (has_ring) {
    const int q = tid;
    int rx, ry;
    if (q < 2 * RX) {  // rows 0,1
      ry = q / RX;
      rx = q % RX;
    } else if (q < 4 * RX) {  // rows TY+2, TY+3
      const int m = q - 2 * RX;
      ry = TY + 2 + m / RX;
      rx = m % RX;
    } else {  // middle rows, columns 0,1,TX+2,TX+3
      const int m = q - 4 * RX;
      ry = 2 + m / 4;
      const int c = m % 4;
      rx = c < 2 ? c : TX + c;
    }
    r_ring = ry * RX + rx;
    o_ring = wrapP(j0 + ry - 2, P2) * P1 + wrapP(i0 + rx - 2, P1);
  }
  // ring-1 slot of a region index (x ring: rx in {1, TX+2}; y ring: ry in {1, TY+2})
  auto zring = [&](int rx, int ry) -> int {
    if ((rx == 1 || rx == TX + 2) && ry >= 2 && ry < TY + 2) return (rx == 1 ? 0 : TY) + (ry - 2);
    if ((ry == 1 || ry == TY + 2) && rx >= 2 && rx < TX + 2)
      return 2 * TY + (ry == 1 ? 0 : TX) + (rx - 2);
    return -1;
  };
  const int zr_ring = has_ring ? zring(r_ring % RX, r_ring / RX) : -1;
  // The thread's ring-pass point, if any, fixed for the launch: an fD point
  // on an x ring-1 column or a gD point on a y ring-1 row (region index,
  // ring-1 slot, slot in the fD/gD array).  The x-ring points go to the wave
  // after the y-ring ones, so that the two extra passes run on different
  // SIMDs (waves are dealt to the SIMDs round-robin).  Computed here rather
  // than per plane, the index keeps the LDS accesses attributable to their
  // arrays.
  // L2 warming (FP64, which has no registers to prefetch into): the cache
  // lines of the next plane, and in stage 2 of phase 4's base state, are
  // pulled into L2 while phases 2 and 3 compute, by one 4-byte load per
  // cache line of the region (own and halo rows, five fields), not one per
  // node (one load per node and field took ~8 % of the plane time).  The
  // loaded word is held in a register that nothing reads until after the
  // phase-3 barrier.  Loads into LDS (LDS-DMA) would need no register, but
  // the compiler drains every LDS-DMA load in flight before the next LDS
  // write (phase 3's fD/gD) and before the barrier: the FP64 stage ran 2.4 %
  // slower with them.  The probes lie at most 16 nodes (one 128-byte line)
  // apart along each of the TY + 4 rows, both ends included; threads past
  // the probe count idle.
  constexpr int kProbeX = (TX + 4 + 15) / 16 + 1, kProbeRow = (TY + 4) * kProbeX;
  static_assert(sizeof(T) == 4 || 5 * kProbeRow <= NT, "one probe per thread");
  bool warm_on = false;
  long long warm_e = 0;  // field and node, as an element offset within a plane
  if constexpr (sizeof(T) == 8) {
    if (tid < 5 * kProbeRow) {
      const int c = tid / kProbeRow, row = (tid % kProbeRow) / kProbeX, k = tid % kProbeX;
      const int col = k == kProbeX - 1 ? TX + 1 : 16 * k - 2;  // relative to i0
      warm_e = c * F + (long long)wrapP(j0 + row - 2, P2) * P1 + wrapP(i0 + col, P1);
      warm_on = true;
    }
  }
  // FP64 stage 2: the same for the base state's lines (own rows only)
  constexpr int kBProbeX = (TX + 15) / 16 + 1, kBProbe = 5 * TY * kBProbeX;
  static_assert(sizeof(T) == 4 || kBProbe <= NT, "one probe per thread");
  bool bwarm_on = false;
  long long bwarm_e = 0;
  if constexpr (sizeof(T) == 8) {
    if (tid < kBProbe) {
      const int c = tid / (TY * kBProbeX), row = (tid % (TY * kBProbeX)) / kBProbeX,
                k = tid % kBProbeX;
      const int col = k == kBProbeX - 1 ? TX - 1 : 16 * k;  // relative to i0
      bwarm_on = j0 + row < P2 && i0 + col < P1;
      bwarm_e = c * F + (long long)min(j0 + row, P2 - 1) * P1 + min(i0 + col, P1 - 1);
    }
  }
  const int rf_t = tid - kRingF0, rg_t = tid - kRingG0;
  const bool fring = rf_t >= 0 && rf_t < 2 * TY, gring = rg_t >= 0 && rg_t < 2 * TX;
  int rp_r = 0, rp_zr = 0, rp_li = 0;
  if (fring) {
    const int side = rf_t / TY, yy = rf_t % TY;
    const int rx = side ? TX + 2 : 1, ry = yy + 2;
    rp_r = ry * RX + rx;
    rp_zr = zring(rx, ry);
    rp_li = yy * (TX + 2) + (rx - 1);
  } else if (gring) {
    const int side = rg_t / TX, xx = rg_t % TX;
    const int ry = side ? TY + 2 : 1, rx = xx + 2;
    rp_r = ry * RX + rx;
    rp_zr = zring(rx, ry);
    rp_li = (ry - 1) * TX + xx;
  }
  // byte offsets of the own and ring nodes within a plane (input, base
  // state); each plane takes fresh copies (see the loop top)
  const unsigned bi_own0 = (unsigned)o_own * sizeof(UIn), bi_ring0 = (unsigned)o_ring * sizeof(UIn);
  const unsigned bd_own0 = (unsigned)o_own * sizeof(double);
  unsigned bi_own = bi_own0, bi_ring = bi_ring0, bd_own = bd_own0;
  constexpr bool kSaddr = sizeof(T) == 4;
  // register windows along the own column
  T Uw[5][5];       // U at planes t-4..t (index 4 = t)
  T ww[5];          // w at t-4..t
  unsigned zq = 0;  // bit b: plane t-b uses QUICK in z (Pe_z > PezVal)
  T zU[4];          // u,v,w,T at plane t-2 (t-1 and t are in the LDS slots)
  T Sg[3][5];       // accumulated flux sums of planes t-2, t-1, t
#pragma unroll
  for (int p = 0; p < 5; p++) {
    ww[p] = T(0);
#pragma unroll
    for (int c = 0; c < 5; c++) Uw[p][c] = T(0);
  }
#pragma unroll
  for (int p = 0; p < 3; p++) {
#pragma unroll
    for (int c = 0; c < 5; c++) Sg[p][c] = T(0);
  }
#pragma unroll
  for (int c = 0; c < 4; c++) zU[c] = T(0);

  const T invRe = T(ph.invRe), ggM = T(ph.ggM), half = T(0.5);
  const T t23 = T(kTwo3rd), gkRP = T(kGama * ph.invRe * kInvPr);
  const T invdx = T(ph.invdx), invdy = T(ph.invdy), invdXi = T(ph.invdXi);
  const T hx = half * invdx, hy = half * invdy;
  const T igM2 = T(1.0 / ph.gM2);

  // primitives of a loaded node; walls take the boundary values
  // (main_v1.9.f90:624-643)
  auto prim = [&](int off, int t, T& u, T& v, T& w, T& Tt, T& p, T& mu,
                  const double (&Uo)[5]) {
    if (t == 0 || t == n3 - 1) {
      const long long wo = (t == 0 ? 0 : PL) + off;
      u = T(a.uw[wo]);
      v = T(a.vw[wo]);
      w = T(0);
      Tt = T(a.tw[wo]);
      mu = T(a.mw[wo]);
      // dw: the duplicate-node correction of this node's wall-plane
      // divergence, as the pressure offset (2/3) mu/Re ddiv it makes in hD
      p = T((a.wflags >> (t != 0)) & 1 ? a.pw[wo] + a.dw[wo] : a.pw[wo]);
    } else {
      const T r = T(Uo[0]);
      const T ir = rcp<T>(r);
      u = T(Uo[1]) * ir;
      v = T(Uo[2]) * ir;
      w = T(Uo[3]) * ir;
      const T E = T(Uo[4]) * ir;
      Tt = E - ggM * half * (u * u + v * v + w * w);
      p = r * Tt * igM2;
      mu = pow07<T>(Tt);
    }
  };

  T u_t = T(0), v_t = T(0), mu_t = T(0);  // u, v, mu of the own node of plane t (phase 1 -> 2)
  const int tb = max(kb - 2, 0), tend = ke + 1;  // inclusive
  const int tlast = min(n3 - 1, tend);             // last plane loaded
  // The next plane's state is loaded a phase ahead into registers, so that
  // its latency is hidden: the mixed kernel, which has registers to spare,
  // right after the barrier that ends phase 3 (phases 3b and 4 hide it);
  // the one-instance FP64 stage-argument kernel (stage 1 in practice) at
  // the start of phase 4, whose register pressure is below phase 3's (the
  // FP64 stage-2 and batched kernels would spill: they load it in phase 1,
  // from L2, where the warming loads of phases 2-3 have put it).
  // Stage 2 reads its base state U^n into Ub: the mixed kernel after the
  // phase-1 barrier, FP64 after the phase-3 barrier (from L2, warmed after
  // the phase-1 barrier; held longer, it would not fit in 256 VGPRs).
  // Stage 1 takes it from the z window: FP64 exactly; in the mixed kernel,
  // whose window is FP32, rounded to FP32, which the FP32 predictor it
  // produces absorbs.  The FP64 stage-argument kernel (stage 1 only, in
  // practice) loads it in phase 4 when run for stage 2.
  constexpr bool kAhead = sizeof(T) == 4 || (S != 2 && std::is_same<I, Inst>::value);
  constexpr bool kUbEarly = sizeof(T) == 4;
  constexpr bool kUbLate = sizeof(T) == 8 && S == 2;
  constexpr bool kP4Early = sizeof(T) == 4;  // phase 4 before the barrier after phase 3
  UIn Un1[5], Ur1[5];
  double Ub[5];  // stage 2: base state of plane t-2
  if constexpr (kAhead) {
#pragma unroll
    for (int c = 0; c < 5; c++) Ur1[c] = UIn(0);
  }
  auto load_plane = [&](int t) {
    const long long n = (long long)t * PL;
    unsigned bi_own = bi_own0, bi_ring = bi_ring0;
    if constexpr (kSaddr) asm volatile("" : "+v"(bi_own), "+v"(bi_ring));
#pragma unroll
    for (int c = 0; c < 5; c++) Un1[c] = *at_b<kSaddr>(Uin + c * F + n, o_own, bi_own);
    if (has_ring) {
#pragma unroll
      for (int c = 0; c < 5; c++) Ur1[c] = *at_b<kSaddr>(Uin + c * F + n, o_ring, bi_ring);
    }
  };
  if constexpr (kAhead) load_plane(tb);
  // the forcing's wall-shear difference is fixed for the launch
  const double tdiff = cread(a.force, 0) - cread(a.force, 1);
  for (int t = tb; t <= tend; t++) {
    const int s = t & 1;
    // The node offsets, made opaque for this plane: otherwise the compiler
    // folds them into loop-invariant 64-bit addresses per field (20 VGPRs)
    // and adds the plane offset to each with a 64-bit vector add.
    if constexpr (kSaddr) {
      bi_own = bi_own0;
      bi_ring = bi_ring0;
      bd_own = bd_own0;
      asm volatile("" : "+v"(bi_own), "+v"(bi_ring), "+v"(bd_own));
    }
    const bool have_plane = t <= n3 - 1;
    // This iteration's metric and forcing values, loaded here and made
    // resident before the phase-1 barrier (the empty asm uses them there):
    // a scalar load still in flight would turn every later LDS wait into a
    // full drain (scalar loads complete out of order under the same
    // counter), and phases 2-4 are where the LDS waits are.  Indices past
    // the grid are clamped; those values go unused.
    const double jac_t = cread(a.jac, min(t, n3 - 1));      // plane t
    const double iJ_t = cread(a.invJac, min(t, n3 - 1));    // plane p+1 = t
    const double iJ_p = cread(a.invJac, max(t - 1, 0));     // plane p = t-1
    const double iJ_q = cread(a.invJac, max(t - 2, 0));     // plane q = t-2
    const double rb_q = cread(a.force, 2 + max(t - 2, 0));  // rho_bar(q)
    // Wave priorities.  A SIMD issues the oldest ready wave first, so of
    // two waves of a block on one SIMD the older would run ahead to each
    // barrier and the younger finish alone, its latencies unhidden.  Each
    // barrier interval is cut into segments of falling priority (s_setprio
    // at the segment starts): a wave that is behind outranks one ahead of
    // it, and the waves of a SIMD reach the barrier together.  Segments,
    // priority 2 down to 0: after barrier 1, phase 2 x | phase 2 y | phase
    // 3 (FP64), or phase 2 | phase 3 own nodes | ring nodes and phase 4
    // (FP32); after barrier 2, priority 1 then 0: phase 3b (FP64: and 4) |
    // phase 1.  In the FP32 kernels the three waves of a SIMD belong to
    // three blocks; the same order cut their stage time by 3 %.
    __builtin_amdgcn_s_setprio(0);
    // ---- phase 1: load plane t ----------------------------------------
    // plane t-2 of the own column leaves the LDS slot this iteration
#pragma unroll
    for (int c = 0; c < 4; c++) zU[c] = sl[s][c][r_own];
    if (have_plane) {
      T u, v, w, Tt, p, mu;
      double Uo[5];
      if constexpr (kAhead) {
#pragma unroll
        for (int c = 0; c < 5; c++) Uo[c] = Un1[c];
      } else {
        const long long n = (long long)t * PL;
#pragma unroll
        for (int c = 0; c < 5; c++) Uo[c] = *at_b<kSaddr>(Uin + c * F + n, o_own, bi_own);
      }
      prim(o_own, t, u, v, w, Tt, p, mu, Uo);
      u_t = u;
      v_t = v;
      mu_t = mu;
      sl[s][0][r_own] = u;
      sl[s][1][r_own] = v;
      sl[s][2][r_own] = w;
      sl[s][3][r_own] = Tt;
      spo[s][0][tid] = p;
      spo[s][1][tid] = mu;
#pragma unroll
      for (int c = 0; c < 5; c++) sub[c * NR + r_own] = T(Uo[c]);
      // shift the column windows
#pragma unroll
      for (int q = 0; q < 4; q++) {
        ww[q] = ww[q + 1];
#pragma unroll
        for (int c = 0; c < 5; c++) Uw[q][c] = Uw[q + 1][c];
      }
      ww[4] = w;
#pragma unroll
      for (int c = 0; c < 5; c++) Uw[4][c] = T(Uo[c]);
      // z scheme of this plane (main_v1.9.f90:409-416)
      const bool qz = T(ph.Re) * fabs(T(Uo[3])) * T(ph.dXi) * T(jac_t) > T(ph.Pez) * mu;
      zq = (zq << 1) | (qz ? 1u : 0u);
      if (has_ring) {
        if constexpr (kAhead) {
#pragma unroll
          for (int c = 0; c < 5; c++) Uo[c] = Ur1[c];
        } else {
          const long long n = (long long)t * PL;
#pragma unroll
          for (int c = 0; c < 5; c++) Uo[c] = *at_b<kSaddr>(Uin + c * F + n, o_ring, bi_ring);
        }
        prim(o_ring, t, u, v, w, Tt, p, mu, Uo);
        const int r = r_ring;
        const int zr = zr_ring;
        if (zr >= 0) {
          srz[t % 3][0][zr] = u;
          srz[t % 3][1][zr] = v;
          srz[t % 3][2][zr] = w;
          spr[s][0][zr] = p;
          spr[s][1][zr] = mu;
        }
        sl[s][0][r] = u;
        sl[s][1][r] = v;
        sl[s][2][r] = w;
        sl[s][3][r] = Tt;
#pragma unroll
        for (int c = 0; c < 5; c++) sub[c * NR + r] = T(Uo[c]);
      }
    } else {
      // past the top wall: shift the windows by one plane of zeros
#pragma unroll
      for (int q = 0; q < 4; q++) {
        ww[q] = ww[q + 1];
#pragma unroll
        for (int c = 0; c < 5; c++) Uw[q][c] = Uw[q + 1][c];
      }
      zq <<= 1;
    }
    // shift the flux-sum window: Sg[2] becomes plane t
#pragma unroll
    for (int c = 0; c < 5; c++) {
      Sg[0][c] = Sg[1][c];
      Sg[1][c] = Sg[2][c];
      Sg[2][c] = T(0);
    }
    asm volatile("" ::"s"(jac_t), "s"(iJ_t), "s"(iJ_p), "s"(iJ_q), "s"(rb_q));
    __syncthreads();
    __builtin_amdgcn_s_setprio(2);
    unsigned warm_w = 0, warm_b = 0;  // the warming loads' words (discarded)
    {
      if (sizeof(T) == 8 && t + 1 <= tlast && warm_on)  // plane t+1 into L2 (FP64)
        warm_w = *(gptr<const unsigned int>)(Uin + (long long)(t + 1) * PL + warm_e);
      // the base state of phase 4 (stage 2): the mixed kernel loads it into
      // registers here, FP64 warms its lines in L2
      if (stage == 2 && t - 2 >= kb && t - 2 < ke) {
        const long long nb = (long long)(t - 2) * PL;
        if constexpr (kUbEarly) {
          if (own_valid) {
#pragma unroll
            for (int c = 0; c < 5; c++) Ub[c] = *at_b<kSaddr>(Ubase + c * F + nb, o_own, bd_own);
          }
        } else if constexpr (kUbLate) {
          if (bwarm_on) warm_b = *(gptr<const unsigned int>)(Ubase + nb + bwarm_e);
        }
      }
    }

    // (the warming loads stay here: nothing would keep the scheduler from
    // sinking them towards their only use)
    __builtin_amdgcn_sched_barrier(0);
    // ---- phase 2: x, y convective fluxes of plane t (main_v1.9.f90:333-407)
    // Each direction first gathers its whole stencil from LDS, the node's own
    // state coming from the z window and phase 1's registers, and only then
    // computes: the reads go out back to back and their latencies overlap,
    // instead of one LDS round trip per stencil point (sched_barrier keeps
    // the compiler from sinking the reads back to their uses).
    if (have_plane && t >= kb && t < ke) {
      const T* U5 = sub;
      const int ro = r_own;
      // x, then y: the hybrid test (main_v1.9.f90:333-335) differs from node
      // to node within a wave, so both weight sets are formed and the node's
      // own is selected; one read of the stencil serves either scheme
      {
        const T* ul = sl[s][0];
        const bool quick = T(ph.Re) * fabs(Uw[4][1]) * T(ph.dx) > T(ph.Pex) * mu_t;
        T qx[5][4];
        const T um2 = ul[ro - 2], um1 = ul[ro - 1], up1 = ul[ro + 1], up2 = ul[ro + 2];
#pragma unroll
        for (int c = 0; c < 5; c++) {
          const T* q = U5 + c * NR + ro;
          qx[c][0] = q[-2];
          qx[c][1] = q[-1];
          qx[c][2] = q[1];
          qx[c][3] = q[2];
        }
        __builtin_amdgcn_sched_barrier(0);
        T wq[5], wc[5];
        quick_w<T>((u_t + up1) * half, (u_t + um1) * half, invdx * T(0.125), wq);
        central_w<T>(um2, um1, up1, up2, invdx * T(kOne12th), wc);
#pragma unroll
        for (int i = 0; i < 5; i++) wq[i] = quick ? wq[i] : wc[i];
#pragma unroll
        for (int c = 0; c < 5; c++)
          Sg[2][c] = apply_w<T>(wq, qx[c][0], qx[c][1], Uw[4][c], qx[c][2], qx[c][3], Sg[2][c]);
      }
      if constexpr (sizeof(T) == 8) __builtin_amdgcn_s_setprio(1);
      {
        const T* vl = sl[s][1];
        const bool quick = T(ph.Re) * fabs(Uw[4][2]) * T(ph.dy) > T(ph.Pey) * mu_t;
        T qy[5][4];
        const T vm2 = vl[ro - 2 * RX], vm1 = vl[ro - RX], vp1 = vl[ro + RX], vp2 = vl[ro + 2 * RX];
#pragma unroll
        for (int c = 0; c < 5; c++) {
          const T* q = U5 + c * NR + ro;
          qy[c][0] = q[-2 * RX];
          qy[c][1] = q[-RX];
          qy[c][2] = q[RX];
          qy[c][3] = q[2 * RX];
        }
        __builtin_amdgcn_sched_barrier(0);
        T wq[5], wc[5];
        quick_w<T>((v_t + vp1) * half, (v_t + vm1) * half, invdy * T(0.125), wq);
        central_w<T>(vm2, vm1, vp1, vp2, invdy * T(kOne12th), wc);
#pragma unroll
        for (int i = 0; i < 5; i++) wq[i] = quick ? wq[i] : wc[i];
#pragma unroll
        for (int c = 0; c < 5; c++)
          Sg[2][c] = apply_w<T>(wq, qy[c][0], qy[c][1], Uw[4][c], qy[c][2], qy[c][3], Sg[2][c]);
      }
    }
    // no barrier: phase 3 reads plane t-1's slots and the z-ring (complete
    // since phase 1) and writes only fD / gD, which phase 2 does not touch

    // ---- phase 3: point fluxes of plane p = t-1 (main_v1.9.f90:457-563)
    __builtin_amdgcn_s_setprio(sizeof(T) == 8 ? 0 : 1);
    const int p = t - 1;
    const int sp = p & 1;
    const bool need_h = p >= kb - 1 && p <= ke && p >= 0;
    const bool need_fg = p >= kb && p < ke;
    // z derivatives of a ring-1 or own node at plane p: the central form,
    // then, on a wall plane only, the one-sided one (a single uniform branch
    // per node pass; a branch per derivative made a chain of some 40 scalar
    // branches per plane)
    const T czc = half * invdXi * T(iJ_p);
    const bool pwall = (p == 0) | (p == n3 - 1);
    auto zder_wall = [&](T qm, T q0, T qp) -> T {
      return p == 0 ? (qp - q0) * T(ph.invdz1) : (q0 - qm) * T(ph.invdzN);
    };
    // derivatives at a region node of plane p; z values supplied.  The
    // node's whole x/y stencil (u, v, w, T at r, r+-1, r+-RX) and its p, mu
    // are read before anything is computed, so that the 22 LDS reads overlap
    // (see phase 2).
    struct Sten {
      T nb[4][4];  // u, v, w, T at r-1, r+1, r-RX, r+RX
      T c[4];      // u, v, w, T at r
      T pr, mu;
    };
    // (r: region index; pm0, pm1: the node's p and mu, own or ring-1)
    auto gather = [&](int r, const T* pm0, const T* pm1, Sten& G) {
#pragma unroll
      for (int f = 0; f < 4; f++) {
        const T* q = sl[sp][f];
        G.nb[f][0] = q[r - 1];
        G.nb[f][1] = q[r + 1];
        G.nb[f][2] = q[r - RX];
        G.nb[f][3] = q[r + RX];
        G.c[f] = q[r];
      }
      G.pr = *pm0;
      G.mu = *pm1;
    };
    auto derivs_of = [&](const Sten& G, const T zm[3], const T zp[3], T zTm, T zTp, Deriv<T>& D) {
      D.dudx = (G.nb[0][1] - G.nb[0][0]) * hx;
      D.dvdx = (G.nb[1][1] - G.nb[1][0]) * hx;
      D.dwdx = (G.nb[2][1] - G.nb[2][0]) * hx;
      D.dTdx = (G.nb[3][1] - G.nb[3][0]) * hx;
      D.dudy = (G.nb[0][3] - G.nb[0][2]) * hy;
      D.dvdy = (G.nb[1][3] - G.nb[1][2]) * hy;
      D.dwdy = (G.nb[2][3] - G.nb[2][2]) * hy;
      D.dTdy = (G.nb[3][3] - G.nb[3][2]) * hy;
      D.dudz = (zp[0] - zm[0]) * czc;
      D.dvdz = (zp[1] - zm[1]) * czc;
      D.dwdz = (zp[2] - zm[2]) * czc;
      D.dTdz = (zTp - zTm) * czc;
      if (pwall) {
        D.dudz = zder_wall(zm[0], G.c[0], zp[0]);
        D.dvdz = zder_wall(zm[1], G.c[1], zp[1]);
        D.dwdz = zder_wall(zm[2], G.c[2], zp[2]);
        D.dTdz = zder_wall(zTm, G.c[3], zTp);
      }
    };
    T Fx[4] = {T(0), T(0), T(0), T(0)}, Gy[4] = {T(0), T(0), T(0), T(0)};
    if (need_h) {
      // own node: all three point fluxes; hD stays in registers
      const int r = r_own;
      const T zm[3] = {zU[0], zU[1], zU[2]};
      Deriv<T> D;
      Sten G;
      gather(r, &spo[sp][0][tid], &spo[sp][1][tid], G);
      const T zp[3] = {sl[s][0][r], sl[s][1][r], sl[s][2][r]};
      const T zTp = sl[s][3][r];
      __builtin_amdgcn_sched_barrier(0);
      derivs_of(G, zm, zp, zU[3], zTp, D);
      const T u = G.c[0], v = G.c[1], w = G.c[2], pr = G.pr, mu = G.mu;
      const T divg = D.dudx + D.dvdy + D.dwdz;
      const T mR = mu * invRe;
      const T sxz = D.dwdx + D.dudz, syz = D.dwdy + D.dvdz, sxy = D.dvdx + D.dudy;
      const T dissH = ggM * invRe * mu * (-u * sxz - v * syz + t23 * w * divg - T(2) * w * D.dwdz);
      T h[4];
      h[0] = -mR * sxz;
      h[1] = -mR * syz;
      h[2] = pr + t23 * mR * divg - T(2) * mR * D.dwdz;
      h[3] = -gkRP * mu * D.dTdz + ggM * w * pr + dissH;
      // hD(p) enters the z divergence of plane p-1 (finished by this
      // iteration's phase 4: Sg[0]) and of plane p+1 (begun in this
      // iteration: Sg[2]); no plane of hD needs to be kept.  Added without
      // testing that those planes are in the chunk: the sums of planes
      // outside it are never used, and a test here would split the pass
      // into basic blocks the scheduler cannot cross.
      {
        const T cm = half * invdXi * T(iJ_q), cp = half * invdXi * T(iJ_t);
#pragma unroll
        for (int c = 0; c < 4; c++) {
          Sg[0][c + 1] += h[c] * cm;
          Sg[2][c + 1] -= h[c] * cp;
        }
      }
      if (need_fg) {
        const T dissF = ggM * invRe * mu * (t23 * u * divg - T(2) * u * D.dudx - v * sxy - w * sxz);
        const T dissG = ggM * invRe * mu * (-u * sxy + t23 * v * divg - T(2) * v * D.dvdy - w * syz);
        Fx[0] = pr + t23 * mR * divg - T(2) * mR * D.dudx;
        Fx[1] = -mR * sxy;
        Fx[2] = -mR * sxz;
        Fx[3] = -gkRP * mu * D.dTdx + ggM * u * pr + dissF;
        Gy[0] = -mR * sxy;
        Gy[1] = pr + t23 * mR * divg - T(2) * mR * D.dvdy;
        Gy[2] = -mR * syz;
        Gy[3] = -gkRP * mu * D.dTdy + ggM * v * pr + dissG;
      }
    }
    if constexpr (sizeof(T) == 4) __builtin_amdgcn_s_setprio(0);
    if (need_fg) {
      // fD on the tile and its x ring-1 columns
      T* fx = sfg;  // [4][(TX+2)*TY], index (ty)*(TX+2) + (rx-1)
      {
        const int li = ty * (TX + 2) + (tx + 1);
#pragma unroll
        for (int c = 0; c < 4; c++) fx[c * NFX + li] = Fx[c];
      }
      if (fring) {
        const int r = rp_r, zr = rp_zr;
        const int tm = (p - 1 + 3) % 3, tp = (p + 1) % 3;
        const T zm[3] = {srz[tm][0][zr], srz[tm][1][zr], srz[tm][2][zr]};
        const T zp[3] = {srz[tp][0][zr], srz[tp][1][zr], srz[tp][2][zr]};
        Deriv<T> D;
        Sten G;
        gather(r, &spr[sp][0][zr], &spr[sp][1][zr], G);
        __builtin_amdgcn_sched_barrier(0);
        derivs_of(G, zm, zp, T(0), T(0), D);
        const T u = G.c[0], v = G.c[1], w = G.c[2], pr = G.pr, mu = G.mu;
        const T divg = D.dudx + D.dvdy + D.dwdz;
        const T mR = mu * invRe;
        const T sxz = D.dwdx + D.dudz, sxy = D.dvdx + D.dudy;
        const T dissF = ggM * invRe * mu * (t23 * u * divg - T(2) * u * D.dudx - v * sxy - w * sxz);
        const int li = rp_li;
        fx[0 * NFX + li] = pr + t23 * mR * divg - T(2) * mR * D.dudx;
        fx[1 * NFX + li] = -mR * sxy;
        fx[2 * NFX + li] = -mR * sxz;
        fx[3 * NFX + li] = -gkRP * mu * D.dTdx + ggM * u * pr + dissF;
      }
    }
    if (need_fg) {
      // gD on the tile and its y ring-1 rows
      T* gy = sfg + 4 * NFX;  // [4][TX*(TY+2)], index (ry-1)*TX + tx
      {
        const int li = (ty + 1) * TX + tx;
#pragma unroll
        for (int c = 0; c < 4; c++) gy[c * NGY + li] = Gy[c];
      }
      if (gring) {
        const int r = rp_r, zr = rp_zr;
        const int tm = (p - 1 + 3) % 3, tp = (p + 1) % 3;
        const T zm[3] = {srz[tm][0][zr], srz[tm][1][zr], srz[tm][2][zr]};
        const T zp[3] = {srz[tp][0][zr], srz[tp][1][zr], srz[tp][2][zr]};
        Deriv<T> D;
        Sten G;
        gather(r, &spr[sp][0][zr], &spr[sp][1][zr], G);
        __builtin_amdgcn_sched_barrier(0);
        derivs_of(G, zm, zp, T(0), T(0), D);
        const T u = G.c[0], v = G.c[1], w = G.c[2], pr = G.pr, mu = G.mu;
        const T divg = D.dudx + D.dvdy + D.dwdz;
        const T mR = mu * invRe;
        const T syz = D.dwdy + D.dvdz, sxy = D.dvdx + D.dudy;
        const T dissG = ggM * invRe * mu * (-u * sxy + t23 * v * divg - T(2) * v * D.dvdy - w * syz);
        const int li = rp_li;
        gy[0 * NGY + li] = -mR * sxy;
        gy[1 * NGY + li] = pr + t23 * mR * divg - T(2) * mR * D.dvdy;
        gy[2 * NGY + li] = -mR * syz;
        gy[3 * NGY + li] = -gkRP * mu * D.dTdy + ggM * v * pr + dissG;
      }
    }
    // ---- phase 4: complete plane q = t-2 ------------------------------
    // It needs no other thread's values: the window, the own column's flux
    // sums (Sg[0] is complete once this thread's phase 3 has added hD of
    // plane t-1) and the base state.  The FP32 kernels run it before the
    // barrier after phase 3, where waves that finish phase 3 early would
    // otherwise wait (stage 1 6 % faster).  FP64 runs it after phase 3b: its
    // stage-2 base state is loaded after that barrier, and its stage 1 ran
    // 0.8 % slower with phase 4 moved.
    auto phase4 = [&]() {
      const int q = t - 2;
      if (q >= kb && q < ke) {
        const T iJ = T(iJ_q);
        // z convective (main_v1.9.f90:409-450); column values at q-2..q+2 are
        // window slots 0..4 (slot 2 = q)
        T H[5];
        const bool lowhigh = (q + 1 < 3) || (q + 1 > n3 - 2);
        T wz[5];
        if (lowhigh) {
          // second-order central next to the walls
          const T k = invdXi * iJ * half;
          wz[0] = T(0);
          wz[1] = -ww[1] * k;
          wz[2] = T(0);
          wz[3] = ww[3] * k;
          wz[4] = T(0);
        } else {
          T wc[5];
          quick_w<T>((ww[2] + ww[3]) * half, (ww[2] + ww[1]) * half, invdXi * iJ * T(0.125), wz);
          central_w<T>(ww[0], ww[1], ww[3], ww[4], invdXi * iJ * T(kOne12th), wc);
          const bool quick = (zq & 4u) != 0u;
#pragma unroll
          for (int i = 0; i < 5; i++) wz[i] = quick ? wz[i] : wc[i];
        }
#pragma unroll
        for (int c = 0; c < 5; c++)
          H[c] = apply_w<T>(wz, Uw[0][c], Uw[1][c], Uw[2][c], Uw[3][c], Uw[4][c], T(0));
        // (the z divergence of hD is already in Sg[0])
        double R[5];
        if constexpr (sizeof(T) == 4) {
          // The FP32 kernels' flux sums are FP32, so the forcing and buoyancy
          // terms join the residual in FP32 too.  Stage 1 also updates in
          // FP32, as U* is stored in FP32; stage 2 adds dt R to the FP64
          // base state in FP64 (below).
          const T s1 = T(0.5 * tdiff * rb_q);
          T Rt[5];
#pragma unroll
          for (int c = 0; c < 5; c++) Rt[c] = -(Sg[0][c] + H[c]);
          Rt[1] += s1;
          Rt[4] += T(ph.ggM) * zU[0] * s1;
          Rt[3] += T(-ph.Ri) * Uw[2][0];
          Rt[4] += T(-ph.ggM * ph.Ri) * Uw[2][3];
          if constexpr (kOut32) {
            if (own_valid) {
              const long long n = (long long)q * PL;
              const T hdt = T(0.5 * ph.dt);
#pragma unroll
              for (int c = 0; c < 5; c++)
                *at_b<kSaddr>((float*)Uout + c * F + n, o_own, bd_own / 2) = Uw[2][c] + hdt * Rt[c];
            }
            return;
          }
#pragma unroll
          for (int c = 0; c < 5; c++) R[c] = (double)Rt[c];
        } else {
          const double rb = rb_q;
          const double src1 = 0.5 * tdiff * rb;
          // u of plane q as phase 1 recovered it (zU holds plane t-2 = q)
          const double src4 = ph.ggM * (double)zU[0] * 0.5 * tdiff * rb;
#pragma unroll
          for (int c = 0; c < 5; c++) R[c] = -(double)(Sg[0][c] + H[c]);
          R[1] += src1;
          R[4] += src4;
          // buoyancy -Ri rho along z and its work (adds -0 when Ri = 0: no branch)
          R[3] += -ph.Ri * (double)Uw[2][0];
          R[4] += -ph.ggM * ph.Ri * (double)Uw[2][3];
        }
        if (own_valid) {
          const long long n = (long long)q * PL;
          double Un[5];
          if (stage == 1) {
  #pragma unroll
            for (int c = 0; c < 5; c++) Un[c] = (double)Uw[2][c] + 0.5 * ph.dt * R[c];
          } else if constexpr (kUbEarly || kUbLate) {
  #pragma unroll
            for (int c = 0; c < 5; c++) Un[c] = Ub[c] + ph.dt * R[c];
          } else {
  #pragma unroll
            for (int c = 0; c < 5; c++)
              Un[c] = *at_b<kSaddr>(Ubase + c * F + n, o_own, bd_own) + ph.dt * R[c];
          }
          if constexpr (kOut32) {
  #pragma unroll
            for (int c = 0; c < 5; c++)
              *at_b<kSaddr>((float*)Uout + c * F + n, o_own, bd_own / 2) = (float)Un[c];
          } else {
  #pragma unroll
            for (int c = 0; c < 5; c++) *at_b<kSaddr>(Uout + c * F + n, o_own, bd_own) = Un[c];
          }
          if (stage == 2) {  // blow-up check on the new state (main_v1.9.f90:728-740)
            // with rho > 0, T < 0 <=> rho^2 T = rho (rho E) - ggM/2 |rho u|^2 < 0
            const double r2T = Un[0] * Un[4] - ph.ggM * 0.5 * (Un[1] * Un[1] + Un[2] * Un[2] + Un[3] * Un[3]);
            if (Un[0] < 0.0 || r2T < 0.0) {
              const unsigned long long lin =
                  ((unsigned long long)q * a.n2 + (unsigned long long)(gj_own + 1)) * a.n1 + (gi_own + 1);
              atomicMin(a.blow, lin);
              atomicCAS(a.halt, 0, step);
            }
          }
        }
      }
    };
    if constexpr (kP4Early) phase4();
    __syncthreads();
    __builtin_amdgcn_s_setprio(1);
    if constexpr (sizeof(T) == 8) asm volatile("" ::"v"(warm_w), "v"(warm_b));
    if constexpr (kAhead && sizeof(T) == 4) {
      if (t + 1 <= tlast) load_plane(t + 1);
    }
    if constexpr (kUbLate) {
      if (t - 2 >= kb && t - 2 < ke && own_valid) {
        const long long nq = (long long)(t - 2) * PL;
#pragma unroll
        for (int c = 0; c < 5; c++) Ub[c] = *at_b<kSaddr>(Ubase + c * F + nq, o_own, bd_own);
      }
    }
    if (need_fg) {
      const T* fx = sfg;
      const T* gy = sfg + 4 * NFX;
      const int lf = ty * (TX + 2) + (tx + 1);
      const int lg = (ty + 1) * TX + tx;
      // all 16 reads first, one wait (see phase 2)
      T d[4][4];
#pragma unroll
      for (int c = 0; c < 4; c++) {
        d[c][0] = fx[c * NFX + lf - 1];
        d[c][1] = fx[c * NFX + lf + 1];
        d[c][2] = gy[c * NGY + lg - TX];
        d[c][3] = gy[c * NGY + lg + TX];
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int c = 0; c < 4; c++)
        Sg[1][c + 1] += (d[c][1] - d[c][0]) * hx + (d[c][3] - d[c][2]) * hy;
    }

    if constexpr (kAhead && sizeof(T) == 8) {
      if (t + 1 <= tlast) load_plane(t + 1);
    }
    if constexpr (!kP4Early) phase4();
    // No barrier here.  The next plane's phase 1 writes sl, spo, spr, srz
    // and sub, which this plane last read in phases 2 and 3, before the
    // barrier after phase 3; phases 3b and 4 read only sfg and registers,
    // and sfg is next written in phase 3, after the next phase-1 barrier.
    // So a wave that finishes phase 4 early starts the next plane's loads
    // while the others finish theirs.
  }
}

// One instance: blocks cover (tiles x z-chunks).  Blocks are dealt to the 8
// XCDs round-robin; renumber them so that each XCD walks a contiguous range
// of tiles and their shared halos stay in that XCD's L2.
// The FP32 kernels are held to 168 VGPRs, three waves per SIMD: three blocks
// of a 32x8 tile share a CU (128 KB of LDS).  FP64 runs at 256 VGPRs, one
// 512-thread block per CU, which its 137 KB of LDS allows anyway.
template <typename T, int TX, int TY, int S>
__global__ void __launch_bounds__(TX* TY) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 3 : 1)))
k_stage_fused(Inst a, int stage, int step) {
  const int nblk = a.ntx * a.nty * a.nz;
  int bid = blockIdx.x;
  if ((nblk & 7) == 0) bid = (bid & 7) * (nblk >> 3) + (bid >> 3);
  const int tile_x = bid % a.ntx;
  bid /= a.ntx;
  const int tile_y = bid % a.nty;
  const int zc = bid / a.nty;
  stage_tile<T, TX, TY, S>(a, stage, step, tile_x, tile_y, zc);
}

// Many instances in one persistent launch: each block pulls work items
// (instance, tile_x, tile_y, z-chunk) from a counter until the list is
// exhausted.  The counter only grows: every block ends with exactly one
// fetch past the list, so after a launch it stands at base + nwork + grid,
// which the host uses as the next base (no reset between launches).  FP32
// is held to three waves per SIMD, as in the one-instance kernel.
template <typename T, int TX, int TY, int S>
__global__ void __launch_bounds__(TX* TY) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 3 : 1)))
k_stage_batched(const Inst* __restrict__ insts, const int4* __restrict__ work, int nwork,
                unsigned* counter, unsigned base, int stage, int step) {
  __shared__ int s_item;
  for (;;) {
    if (threadIdx.x == 0) s_item = (int)(atomicAdd(counter, 1u) - base);
    __syncthreads();
    // the item is block-uniform: keep it (and the descriptor loads) scalar
    const int w = __builtin_amdgcn_readfirstlane(s_item);
    __syncthreads();
    if (w >= nwork) break;
    const int4 it = work[w];
    const int inst = __builtin_amdgcn_readfirstlane(it.x);
    // the descriptor through the constant address space, so that its fields
    // are scalar loads as for a kernel argument; read through a generic
    // pointer, the FP64 build spilled 68 VGPRs (now 15) and ran 2x slower
    const __attribute__((address_space(4))) Inst& ia =
        ((const __attribute__((address_space(4))) Inst*)insts)[inst];
    stage_tile<T, TX, TY, S>(ia, stage, step, __builtin_amdgcn_readfirstlane(it.y),
                          __builtin_amdgcn_readfirstlane(it.z), __builtin_amdgcn_readfirstlane(it.w));
    __syncthreads();
  }
}

// Wall values and wall-pressure extrapolation in FP64 for the fused path
// (main_v1.9.f90:622-716): the wall velocity and temperature (no-slip and
// T = 1, the Navier slip u_w = Ls du/dn, or a study wall condition, see
// wall_bc), rho_w <- gama Ma^2 p_old / T_w, mu_w = T_w^0.7, then the new
// wall pressure from the z-momentum balance at an impervious wall, with the
// gravity term -Ri rho_w when buoyancy is on.  Under a study condition the
// threads on the domain edges also update the wall's periodic duplicates
// (the reference's boundary loop covers them at their own coordinates) and
// the four corners, whose interior columns the reference never refreshes.
// One 64 x 4 block of wall nodes, run by 64 x 8 threads.  The block first
// recovers the primitives of its patch plus a 1-node x/y halo at the two
// planes next to the wall, up to two columns per thread, into LDS; the threads of
// the first four rows then compute a wall node each from the five columns it
// needs.  (With 256 threads, each recovered up to four columns in turn, at
// one wave per SIMD.)
constexpr int kWX = 64, kWY = 4, kWTY = 8, kWLX = kWX + 2, kWLY = kWY + 2;
__device__ __forceinline__ void wall_tile(const Inst& a, int stage, int step, int wall) {
  __shared__ double wq[7][2][kWLY][kWLX];  // r, u, v, w, T, p, mu
  const Geom& g = a.g;
  const Phys& ph = a.ph;
  const int i0 = blockIdx.x * kWX, j0 = blockIdx.y * kWY;
  if (i0 >= g.P1 || j0 >= g.P2) return;  // block-uniform (batches of smaller instances)
  // The halt flag, the patch's state and the node's old wall pressure are
  // loaded together (one memory latency, not three); the halt is tested once
  // they are in flight.  A halt set during this launch holds this step, so
  // the test is block-uniform.
  const int h = *a.halt;
  const int ii = i0 + threadIdx.x, jj = j0 + threadIdx.y;
  const double* __restrict__ U = stage == 1 ? a.UB : a.UA;
  // after stage 1 of a mixed instance, UB holds U* as FP32
  const float* __restrict__ Uf = (stage == 1 && a.u32) ? (const float*)a.UB : nullptr;
  double* __restrict__ pw = a.pw;
  double* __restrict__ rw = a.rw;
  const int check = stage == 2, n1 = a.n1, n2 = a.n2;
  const long long F = g.field, PL = g.plane;
  const int N = g.n3 - 1;
  const long long n2d = (long long)jj * g.P1 + ii;
  const int kw = wall ? N : 0;
  // blow-up flag at reference node (i, j) of this wall (1-based)
  auto flag = [&](int i, int j) {
    const unsigned long long lin =
        ((unsigned long long)kw * n2 + (unsigned long long)(j - 1)) * n1 + (i - 1);
    atomicMin(a.blow, lin);
    atomicCAS(a.halt, 0, step);
  };
  const int k1 = wall ? N - 1 : 1, k2 = wall ? N - 2 : 2;
  const double c0 = wall ? ph.b1d2c0 : ph.f1d2c0, c1 = wall ? ph.b1d2c1 : ph.f1d2c1,
               c2 = wall ? ph.b1d2c2 : ph.f1d2c2;
  const double e0 = wall ? ph.b2d1c0 : ph.f2d1c0, e1 = wall ? ph.b2d1c1 : ph.f2d1c1,
               e2 = wall ? ph.b2d1c2 : ph.f2d1c2;
  struct P {
    double r, u, v, w, T, p, mu;
  };
  const bool valid = threadIdx.y < kWY && ii < g.P1 && jj < g.P2;
  const double pold = valid ? pw[wall * PL + n2d] : 0.0;
  {
    // the patch's columns at planes k1, k2 (its four corners are not needed),
    // up to NPER per thread, all loads issued before any is used
    auto wrapm = [](int x, int P) {
      x %= P;
      return x < 0 ? x + P : x;
    };
    constexpr int NE = 2 * kWLX * kWLY, NT = kWX * kWTY, NPER = (NE + NT - 1) / NT;
    const int tid = threadIdx.y * kWX + threadIdx.x;
    double Uq[NPER][5];
    bool use[NPER];
#pragma unroll
    for (int m = 0; m < NPER; m++) {
      const int e = tid + m * NT;
      const int pl = e / (kWLX * kWLY), cc = e % (kWLX * kWLY);
      const int ly = cc / kWLX, lx = cc % kWLX;
      use[m] = e < NE && !((lx == 0 || lx == kWLX - 1) && (ly == 0 || ly == kWLY - 1));
      if (use[m]) {
        const int gi = wrapm(i0 + lx - 1, g.P1), gj = wrapm(j0 + ly - 1, g.P2);
        const long long n = (long long)(pl ? k2 : k1) * PL + (long long)gj * g.P1 + gi;
        if (Uf) {
#pragma unroll
          for (int c = 0; c < 5; c++) Uq[m][c] = (double)Uf[c * F + n];
        } else {
#pragma unroll
          for (int c = 0; c < 5; c++) Uq[m][c] = U[c * F + n];
        }
      }
    }
    if (h != 0 && h < step) return;
#pragma unroll
    for (int m = 0; m < NPER; m++) {
      if (!use[m]) continue;
      const int e = tid + m * NT;
      const int pl = e / (kWLX * kWLY), cc = e % (kWLX * kWLY);
      const int ly = cc / kWLX, lx = cc % kWLX;
      // the stage kernel's arithmetic (reciprocal, 1/gM2)
      const double r = Uq[m][0];
      const double ir = rcp<double>(r);
      const double u = Uq[m][1] * ir, v = Uq[m][2] * ir, w = Uq[m][3] * ir;
      const double T = Uq[m][4] * ir - ph.ggM * 0.5 * (u * u + v * v + w * w);
      wq[0][pl][ly][lx] = r;
      wq[1][pl][ly][lx] = u;
      wq[2][pl][ly][lx] = v;
      wq[3][pl][ly][lx] = w;
      wq[4][pl][ly][lx] = T;
      wq[5][pl][ly][lx] = r * T * (1.0 / ph.gM2);
      wq[6][pl][ly][lx] = pow07d(T);  // as the stage kernel evaluates mu at these nodes
    }
  }
  __syncthreads();
  if (!valid) return;
  auto prim = [&](int dx, int dy, int pl) {  // column (ii+dx, jj+dy) at plane k1 / k2
    const int lx = threadIdx.x + 1 + dx, ly = threadIdx.y + 1 + dy;
    P q;
    q.r = wq[0][pl][ly][lx];
    q.u = wq[1][pl][ly][lx];
    q.v = wq[2][pl][ly][lx];
    q.w = wq[3][pl][ly][lx];
    q.T = wq[4][pl][ly][lx];
    q.p = wq[5][pl][ly][lx];
    q.mu = wq[6][pl][ly][lx];
    return q;
  };
  const bool study = (a.wflags >> wall) & 1;
  const WallSpec& ws = a.ws[wall];
  // slip velocity of a wall node from the interior values above it
  const double sden = wall ? (1.0 + ph.Ls * c0) : (1.0 - ph.Ls * c0);
  const double ssgn = wall ? -ph.Ls : ph.Ls;
  auto slip = [&](double a1, double a2) {
    return ph.Ls == 0.0 ? 0.0 : ssgn * (c1 * a1 + c2 * a2) / sden;
  };
  // u, v, T and mu of the wall node at reference (i, j) over column q1, q2
  struct WV {
    double u, v, T, mu;
  };
  auto wallval = [&](int i, int j, const P& q1, const P& q2) {
    WV o;
    if (study) {
      wall_bc(ws, i, j, a.time, a.x1f, a.x2f, n1, n2, c0, c1, c2, q1.u, q2.u, q1.v, q2.v, q1.T,
              q2.T, o.u, o.v, o.T);
      o.mu = pow(o.T, 0.7);
    } else {
      o.u = slip(q1.u, q2.u);
      o.v = slip(q1.v, q2.v);
      o.T = 1.0;
      o.mu = 1.0;
    }
    return o;
  };
  const P xp1 = prim(1, 0, 0), xp2 = prim(1, 0, 1), xm1 = prim(-1, 0, 0), xm2 = prim(-1, 0, 1);
  const P yp1 = prim(0, 1, 0), yp2 = prim(0, 1, 1), ym1 = prim(0, -1, 0), ym2 = prim(0, -1, 1);
  const P c1p = prim(0, 0, 0), c2p = prim(0, 0, 1);
  const int ri = ii + 2, rj = jj + 2;  // reference indices of this node
  const WV w0 = wallval(ri, rj, c1p, c2p);
  const WV wxp = wallval(ri + 1, rj, xp1, xp2), wxm = wallval(ri - 1, rj, xm1, xm2);
  const WV wyp = wallval(ri, rj + 1, yp1, yp2), wym = wallval(ri, rj - 1, ym1, ym2);
  const double r = ph.gM2 * pold / w0.T;
  rw[wall * PL + n2d] = r;
  if (check && (r < 0.0 || w0.T < 0.0)) flag(ri, rj);
  a.uw[wall * PL + n2d] = w0.u;
  a.vw[wall * PL + n2d] = w0.v;
  a.tw[wall * PL + n2d] = w0.T;
  a.mw[wall * PL + n2d] = w0.mu;
  if (study) {
    // The wall-plane divergence of an edge node reads the duplicate beyond
    // the edge, which the stage kernel holds as the periodic image: dc is
    // the difference (the image is what that node's own update computes
    // from the same column).
    double dc = 0.0;
    if (ii == 0) dc += (wallval(g.P1 + 1, rj, xm1, xm2).u - wxm.u) * 0.5 * ph.invdx;
    if (ii == g.P1 - 1) dc += (wxp.u - wallval(2, rj, xp1, xp2).u) * 0.5 * ph.invdx;
    if (jj == 0) dc += (wallval(ri, g.P2 + 1, ym1, ym2).v - wym.v) * 0.5 * ph.invdy;
    if (jj == g.P2 - 1) dc += (wyp.v - wallval(ri, 2, yp1, yp2).v) * 0.5 * ph.invdy;
    // at a wall (w = 0) the divergence enters hD only as (2/3) mu/Re div
    // beside the pressure: store it as that pressure offset
    a.dw[wall * PL + n2d] = kTwo3rd * w0.mu * ph.invRe * dc;
    const int NRG = ring_size(g.P1, g.P2);
    double* rg = a.ring + (long long)wall * kRingFields * NRG;
    auto put = [&](int slot, const WV& o, int i, int j) {
      rg[slot] = o.u;
      rg[NRG + slot] = o.v;
      rg[2 * NRG + slot] = o.T;
      const double rd = ph.gM2 * rg[4 * NRG + slot] / o.T;
      rg[3 * NRG + slot] = rd;
      if (check && (rd < 0.0 || o.T < 0.0)) flag(i, j);
    };
    if (ii == 0) put(jj, wxm, 1, rj);
    if (ii == g.P1 - 1) put(g.P2 + jj, wxp, n1, rj);
    if (jj == 0) put(2 * g.P2 + ii, wym, ri, 1);
    if (jj == g.P2 - 1) put(2 * g.P2 + g.P1 + ii, wyp, ri, n2);
    const bool cx = ii == 0 || ii == g.P1 - 1, cy = jj == 0 || jj == g.P2 - 1;
    if (cx && cy) {
      // corner (ci, cj) over its stale interior column (the file's values)
      const int ci = ii == 0 ? 1 : n1, cj = jj == 0 ? 1 : n2;
      auto fv = [&](int c, int k) {
        return a.vfile[(((long long)c * g.n3 + k) * n2 + (cj - 1)) * n1 + (ci - 1)];
      };
      P q1, q2;
      q1.u = fv(1, k1); q1.v = fv(2, k1); q1.T = fv(5, k1);
      q2.u = fv(1, k2); q2.v = fv(2, k2); q2.T = fv(5, k2);
      const int slot = 2 * g.P2 + 2 * g.P1 + (ii == 0 ? 0 : 1) + (jj == 0 ? 0 : 2);
      put(slot, wallval(ci, cj, q1, q2), ci, cj);
    }
  }
  const double dudxdz = ((c0 * wxp.mu * wxp.u + c1 * xp1.mu * xp1.u + c2 * xp2.mu * xp2.u) -
                         (c0 * wxm.mu * wxm.u + c1 * xm1.mu * xm1.u + c2 * xm2.mu * xm2.u)) *
                        0.5 * ph.invdx;
  const double dvdydz = ((c0 * wyp.mu * wyp.v + c1 * yp1.mu * yp1.v + c2 * yp2.mu * yp2.v) -
                         (c0 * wym.mu * wym.v + c1 * ym1.mu * ym1.v + c2 * ym2.mu * ym2.v)) *
                        0.5 * ph.invdy;
  const double d2wdz2 = e0 * w0.mu * 0.0 + e1 * c1p.mu * c1p.w + e2 * c2p.mu * c2p.w;
  const double drhow2dz = c1 * c1p.r * c1p.w * c1p.w + c2 * c2p.r * c2p.w * c2p.w;
  double lhs = ph.invRe * (-dudxdz - dvdydz + kTwo3rd * (dudxdz + dvdydz - 2.0 * d2wdz2)) + drhow2dz;
  if (ph.Ri != 0.0) lhs += ph.Ri * r;
  pw[wall * PL + n2d] = -(lhs + c1 * c1p.p + c2 * c2p.p) / c0;
}

__global__ void __launch_bounds__(512) k_wall_fast(Inst a, int stage, int step) {
  wall_tile(a, stage, step, blockIdx.z);
}

// all instances of a batch: blockIdx.z = 2 * instance + wall
__global__ void __launch_bounds__(512) k_wall_batched(const Inst* __restrict__ insts, int stage,
                                                      int step) {
  wall_tile(insts[blockIdx.z >> 1], stage, step, blockIdx.z & 1);
}

// Tile shapes compiled for the stage kernel.  FP64 keeps 32x16 (137 KB of
// LDS, one block per CU); FP32 also has the smaller shapes the batch
// autotuner and choose_tiling choose between.
void tile_of(int variant, int& tx, int& ty) {
  static const int T[4][2] = {{32, 16}, {32, 8}, {16, 16}, {64, 8}};
  const int v = (variant >= 0 && variant < 4) ? variant : 0;
  tx = T[v][0];
  ty = T[v][1];
}

// Fill the tiling fields of an instance for a tile shape and z-chunk
// (kchunk <= 0: about 512 blocks of at least 8 planes, the batch autotuner's
// default candidate; single instances take choose_tiling's choice).
void set_tiling(Inst& a, int variant, int kchunk) {
  int TX, TY;
  tile_of(variant, TX, TY);
  a.ntx = (a.g.P1 + TX - 1) / TX;
  a.nty = (a.g.P2 + TY - 1) / TY;
  const int nint = a.g.n3 - 2;
  int kc = kchunk;
  if (kc <= 0) {
    const int tiles = a.ntx * a.nty;
    const int nzw = (512 + tiles - 1) / tiles;
    kc = (nint + nzw - 1) / nzw;
    if (kc < 8) kc = 8;
  }
  a.kchunk = kc;
  a.nz = (nint + kc - 1) / kc;
}

template <typename T, int TX, int TY, int S>
static int fused_bpc1() {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_stage_fused<T, TX, TY, S>, TX * TY, 0) !=
      hipSuccess)
    n = 1;
  return n > 0 ? n : 1;
}
// the smaller of the kernels a precision launches (FP64: the stage-argument
// one and stage 2's; FP32: one per stage)
template <typename T, int TX, int TY>
static int fused_bpc() {
  if constexpr (sizeof(T) == 8) {
    const int n0 = fused_bpc1<T, TX, TY, 0>(), n2 = fused_bpc1<T, TX, TY, 2>();
    return n0 < n2 ? n0 : n2;
  } else {
    const int n1 = fused_bpc1<T, TX, TY, 1>(), n2 = fused_bpc1<T, TX, TY, 2>();
    return n1 < n2 ? n1 : n2;
  }
}

// resident blocks per CU of the one-instance stage kernel
static int fused_blocks_per_cu(int mixed, int variant) {
  if (!mixed) return fused_bpc<double, 32, 16>();
  switch (variant) {
    case 1: return fused_bpc<float, 32, 8>();
    case 2: return fused_bpc<float, 16, 16>();
    case 3: return fused_bpc<float, 64, 8>();
    default: return fused_bpc<float, 32, 16>();
  }
}

// Tiling of one instance: the tile shape (FP32 only) and the planes per
// block that minimise the launch's critical path.  A block marches
// kchunk + 3 planes (two below its chunk, one above) at a near-constant time
// per plane, whatever the grid, and the launch runs in ceil(blocks / resident
// slots) rounds; so the cost is rounds x (kchunk + 3).  Small grids get thin
// chunks and many blocks, large ones a single round of long columns.  Ties
// go to the earlier shape in the preference order and, within a shape, to
// the longer chunk (fewer halo planes re-read).  Returns the variant.
int choose_tiling(Inst& a, int mixed, int* ncu_out, int* bpc_out) {
  int dev = 0, ncu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      ncu < 1)
    ncu = 256;
  const int nint = a.g.n3 - 2;
  static const int kOrder[4] = {1, 0, 2, 3};  // FP32: 32x8 first (three blocks per CU)
  long long best = -1;
  int bv = 0, bk = 1;
  for (int i = 0; i < (mixed ? 4 : 1); i++) {
    const int v = mixed ? kOrder[i] : 0;
    int tx, ty;
    tile_of(v, tx, ty);
    const long long tiles =
        (long long)((a.g.P1 + tx - 1) / tx) * (long long)((a.g.P2 + ty - 1) / ty);
    const long long slots = (long long)ncu * fused_blocks_per_cu(mixed, v);
    for (int kc = 1; kc <= nint; kc++) {
      const long long nz = (nint + kc - 1) / kc;
      const long long cost = ((tiles * nz + slots - 1) / slots) * (kc + 3);
      if (best < 0 || cost < best || (cost == best && v == bv && kc > bk)) {
        best = cost;
        bv = v;
        bk = kc;
      }
    }
  }
  set_tiling(a, bv, bk);
  if (ncu_out) *ncu_out = ncu;
  if (bpc_out) *bpc_out = fused_blocks_per_cu(mixed, bv);
  return bv;
}

template <typename T, int TX, int TY>
static void launch_one(const Inst& a, int stage, int step, hipStream_t s) {
  const unsigned nblk = (unsigned)(a.ntx * a.nty * a.nz);
  // FP64: stage 2 has a kernel of its own (it warms and loads the base
  // state); stage 1 keeps the kernel that takes the stage as an argument,
  // which ran 1 % faster than a stage-1 one (that spills 16 bytes at 256
  // VGPRs)
  if constexpr (sizeof(T) == 8) {
    if (stage == 2)
      hipLaunchKernelGGL((k_stage_fused<T, TX, TY, 2>), dim3(nblk), dim3(TX * TY), 0, s, a, stage, step);
    else
      hipLaunchKernelGGL((k_stage_fused<T, TX, TY, 0>), dim3(nblk), dim3(TX * TY), 0, s, a, stage, step);
  } else if (stage == 1)
    hipLaunchKernelGGL((k_stage_fused<T, TX, TY, 1>), dim3(nblk), dim3(TX * TY), 0, s, a, stage, step);
  else
    hipLaunchKernelGGL((k_stage_fused<T, TX, TY, 2>), dim3(nblk), dim3(TX * TY), 0, s, a, stage, step);
}

// UB <- UA as FP32 (mixed precision), every plane: the wall planes, which
// no stage writes, then hold the frozen wall state in UB's FP32 layout
__global__ void __launch_bounds__(256) k_predictor_f32(const double* __restrict__ UA,
                                                       float* __restrict__ UB, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) UB[i] = (float)UA[i];
}
void launch_predictor_f32(const Inst& a, hipStream_t s) {
  const long long n = 5 * a.g.field;
  hipLaunchKernelGGL(k_predictor_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a.UA,
                     (float*)a.UB, n);
}

void launch_wall_fast(const Inst& a, int stage, int step, hipStream_t s) {
  dim3 blk(kWX, kWTY);
  dim3 grd((a.g.P1 + 63) / 64, (a.g.P2 + 3) / 4, 2);
  hipLaunchKernelGGL(k_wall_fast, grd, blk, 0, s, a, stage, step);
}

void launch_stage_fused(const Inst& a, int mixed, int variant, int stage, int step,
                        hipStream_t s, hipEvent_t ev0, hipEvent_t ev1, bool wall) {
  if (ev0) (void)hipEventRecord(ev0, s);
  if (!mixed) {
    launch_one<double, 32, 16>(a, stage, step, s);
  } else {
    switch (variant) {
      case 1: launch_one<float, 32, 8>(a, stage, step, s); break;
      case 2: launch_one<float, 16, 16>(a, stage, step, s); break;
      case 3: launch_one<float, 64, 8>(a, stage, step, s); break;
      default: launch_one<float, 32, 16>(a, stage, step, s); break;
    }
  }
  if (ev1) (void)hipEventRecord(ev1, s);
  if (wall) launch_wall_fast(a, stage, step, s);
}

template <typename T, int TX, int TY, int S>
static int blocks_per_cu1() {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_stage_batched<T, TX, TY, S>, TX * TY, 0) !=
      hipSuccess)
    n = 1;
  return n > 0 ? n : 1;
}
template <typename T, int TX, int TY>
static int blocks_per_cu() {
  if constexpr (sizeof(T) == 8) {
    const int n0 = blocks_per_cu1<T, TX, TY, 0>(), n2 = blocks_per_cu1<T, TX, TY, 2>();
    return n0 < n2 ? n0 : n2;
  } else {
    const int n1 = blocks_per_cu1<T, TX, TY, 1>(), n2 = blocks_per_cu1<T, TX, TY, 2>();
    return n1 < n2 ? n1 : n2;
  }
}

int batched_blocks_per_cu(int mixed, int variant) {
  if (!mixed) return blocks_per_cu<double, 32, 16>();
  switch (variant) {
    case 1: return blocks_per_cu<float, 32, 8>();
    case 2: return blocks_per_cu<float, 16, 16>();
    case 3: return blocks_per_cu<float, 64, 8>();
    default: return blocks_per_cu<float, 32, 16>();
  }
}

template <typename T, int TX, int TY>
static void launch_b(const Inst* d_insts, const int4* d_work, int nwork, unsigned* counter,
                     unsigned base, int nblocks, int stage, int step, hipStream_t s) {
  // FP64: stage 1 takes the stage-argument kernel, stage 2 its own (as the
  // one-instance launch)
  if (stage == 2)
    hipLaunchKernelGGL((k_stage_batched<T, TX, TY, 2>), dim3(nblocks), dim3(TX * TY), 0, s, d_insts,
                       d_work, nwork, counter, base, stage, step);
  else if constexpr (sizeof(T) == 8)
    hipLaunchKernelGGL((k_stage_batched<T, TX, TY, 0>), dim3(nblocks), dim3(TX * TY), 0, s, d_insts,
                       d_work, nwork, counter, base, stage, step);
  else
    hipLaunchKernelGGL((k_stage_batched<T, TX, TY, 1>), dim3(nblocks), dim3(TX * TY), 0, s, d_insts,
                       d_work, nwork, counter, base, stage, step);
}

void launch_stage_batched(const Inst* d_insts, const int4* d_work, int nwork, unsigned* counter,
                          unsigned base, int nblocks, int mixed, int variant, int stage, int step,
                          int ninst, int maxP1, int maxP2, hipStream_t s, hipEvent_t ev0,
                          hipEvent_t ev1) {
  if (ev0) (void)hipEventRecord(ev0, s);
#define CC_LAUNCH_B(T, TX, TY) \
  launch_b<T, TX, TY>(d_insts, d_work, nwork, counter, base, nblocks, stage, step, s)
  if (!mixed) {
    CC_LAUNCH_B(double, 32, 16);
  } else {
    switch (variant) {
      case 1: CC_LAUNCH_B(float, 32, 8); break;
      case 2: CC_LAUNCH_B(float, 16, 16); break;
      case 3: CC_LAUNCH_B(float, 64, 8); break;
      default: CC_LAUNCH_B(float, 32, 16); break;
    }
  }
#undef CC_LAUNCH_B
  if (ev1) (void)hipEventRecord(ev1, s);
  dim3 blk(kWX, kWTY);
  dim3 grd((maxP1 + 63) / 64, (maxP2 + 3) / 4, 2 * ninst);
  hipLaunchKernelGGL(k_wall_batched, grd, blk, 0, s, d_insts, stage, step);
}

}  // namespace cc
