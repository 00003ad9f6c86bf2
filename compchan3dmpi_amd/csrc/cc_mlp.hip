// cc_mlp.hip -- learned subgrid closure: an 11-64-64-64-9 GELU MLP whose
// forward, backward and weight-gradient passes run on BF16 MFMA
// (v_mfma_f32_16x16x16_bf16) with FP32 accumulation, Adam on FP32 master
// weights, the 4x4x4 box filter of DNS snapshots that produces its training
// data, and the kernels that add its subgrid stresses to a coarse run.
//
// Matrix layout.  Activations are kept transposed (features x samples), so a
// layer is Z^T = W^T A^T: M = output features, K = input features, N = 16
// samples of one wave.  With the 16x16x16 MFMA lane layout (lane l, s = l%16,
// q = l/16: A[m=s][k=4q+j], B[k=4q+j][n=s], D[m=4q+j][n=s]) the output block
// of one layer is exactly the B operand of the next, so no transposes are
// needed between layers.
//
// Parameters (FP32 master, flattened): W1[16][64] (rows 11..15 zero), b1[64],
// W2[64][64], b2, W3[64][64], b3, W4[64][16] (columns 9..15 zero), b4[16].
// Features and targets are padded rows of 16 floats.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/cc_abi.h"
#include "cc_device.h"
#include "cc_kernels.h"

namespace cc {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kW1 = 0, kB1 = 1024, kW2 = 1088, kB2 = 5184, kW3 = 5248, kB3 = 9344, kW4 = 9408,
              kB4 = 10432, kNP = 10448;
constexpr int kNin = 11, kNout = 9;

__device__ __forceinline__ short f2bf(float x) {  // v_cvt_pk_bf16_f32: nearest even
  return __builtin_bit_cast(short, (__bf16)x);
}
__device__ __forceinline__ float bf2f(short b) {
  return __uint_as_float(((unsigned)(unsigned short)b) << 16);
}
__device__ __forceinline__ bf16x4 to_bf(f32x4 v) {
  bf16x4 r;
  r[0] = f2bf(v[0]);
  r[1] = f2bf(v[1]);
  r[2] = f2bf(v[2]);
  r[3] = f2bf(v[3]);
  return r;
}
// The standard normal CDF Phi(x) = (1 + erf(x/sqrt 2))/2 and e = exp(-x^2/2),
// from Abramowitz & Stegun 7.1.26: for z >= 0,
//   erfc z = t (a1 + t (a2 + t (a3 + t (a4 + t a5)))) exp(-z^2),  t = 1/(1 + p z),
// with an absolute error of at most 1.5e-7 (far below the BF16 rounding the
// activations get next).  With z = |x|/sqrt 2, exp(-z^2) = e, and Phi takes
// erfc directly on the negative side, so 1 + erf never cancels.
__device__ __forceinline__ float norm_cdf(float x, float& e) {
  // t = 1/(1 + p z) with p z = (0.3275911/sqrt 2)|x|; the polynomial carries
  // the factor 1/2 of erfc/2 (coefficients a_i/2); e = 2^(-x^2 log2(e)/2)
  const float t = __builtin_amdgcn_rcpf(fmaf(0.23164190f, fabsf(x), 1.0f));
  float p = fmaf(0.5307027145f, t, -0.7265760135f);
  p = fmaf(p, t, 0.7107068705f);
  p = fmaf(p, t, -0.142248368f);
  p = fmaf(p, t, 0.127414796f);
  e = __builtin_amdgcn_exp2f((x * x) * -0.72134752f);
  const float h = (p * t) * e;  // erfc(|x|/sqrt 2)/2
  return x < 0.f ? h : 1.0f - h;
}
// GELU(x) = x Phi(x); GELU'(x) = Phi(x) + x phi(x), phi(x) = e/sqrt(2 pi)
__device__ __forceinline__ float gelu(float x) {
  float e;
  return x * norm_cdf(x, e);
}
__device__ __forceinline__ float dgelu(float x) {
  float e;
  const float c = norm_cdf(x, e);
  return fmaf(x * 0.39894228f, e, c);
}
__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// BF16 weights stored as MFMA A operands: operand o, lane l = 16q + s holds
// its 4 values at wpk[(o*64 + l)*4].  Lane (s, q), element j of
//   a forward operand (W^T block (kb, ob)): W[kb*16 + 4q + j][ob*16 + s]
//     W1 ob -> ob; W2 (kb, ob) -> 4 + 4kb + ob; W3 -> 20 + 4kb + ob; W4 kb -> 36 + kb;
//   a backward operand (W block (ib, kb)): W[ib*16 + s][kb*16 + 4q + j]
//     W4 ib -> 40 + ib; W3 (ib, kb) -> 44 + 4ib + kb; W2 -> 60 + 4ib + kb.
// Each weight sits once in the forward operands and, W2..W4, once in the
// backward ones; together they fill all 76 operands.
constexpr int kNOps = 76;
constexpr int kOpW1 = 0, kOpW2 = 4, kOpW3 = 20, kOpW4 = 36;  // forward
constexpr int kOpB4 = 40, kOpB3 = 44, kOpB2 = 60;            // backward

__device__ __forceinline__ bf16x4 wop(const short* __restrict__ wpk, int op, int lane) {
  return *(const bf16x4*)(wpk + ((size_t)op * 64 + lane) * 4);
}

// store the BF16 value of parameter p at its operand positions (biases: none)
__device__ __forceinline__ void pack_weight(short* wpk, int p, short v) {
  int layer, r, c;
  if (p < kB1) {
    layer = 0, r = (p - kW1) / 64, c = (p - kW1) % 64;
  } else if (p >= kW2 && p < kB2) {
    layer = 1, r = (p - kW2) / 64, c = (p - kW2) % 64;
  } else if (p >= kW3 && p < kB3) {
    layer = 2, r = (p - kW3) / 64, c = (p - kW3) % 64;
  } else if (p >= kW4 && p < kB4) {
    layer = 3, r = (p - kW4) / 16, c = (p - kW4) % 16;
  } else {
    return;
  }
  const int kb = r / 16, ob = c / 16;
  const int of = layer == 0 ? kOpW1 + ob
                 : layer == 1 ? kOpW2 + 4 * kb + ob
                 : layer == 2 ? kOpW3 + 4 * kb + ob
                              : kOpW4 + kb;
  wpk[((size_t)of * 64 + ((r % 16) / 4) * 16 + (c % 16)) * 4 + (r % 4)] = v;
  if (layer == 0) return;
  const int ib = r / 16, kc = c / 16;
  const int obk = layer == 3 ? kOpB4 + ib : layer == 2 ? kOpB3 + 4 * ib + kc : kOpB2 + 4 * ib + kc;
  wpk[((size_t)obk * 64 + ((c % 16) / 4) * 16 + (r % 16)) * 4 + (c % 4)] = v;
}

// Forward pass; Z (optional) receives the pre-activations Z1|Z2|Z3 as
// [N][192]; idx (optional) gathers the samples.
__global__ void __launch_bounds__(256) k_mlp_fwd(const float* __restrict__ X,
                                                 const int* __restrict__ idx, int N,
                                                 const short* __restrict__ wpk,
                                                 const float* __restrict__ P,
                                                 float* __restrict__ Y, float* __restrict__ Z) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int s = lane & 15, q = lane >> 4;
  const int s0 = (blockIdx.x * (blockDim.x >> 6) + wave) * 16;
  if (s0 >= N) return;  // whole wave
  const int n = s0 + s;
  const bool valid = n < N;
  const int src = valid ? (idx ? idx[n] : n) : 0;
  f32x4 x = {0.f, 0.f, 0.f, 0.f};
  if (valid) x = *(const f32x4*)(X + (size_t)src * 16 + 4 * q);
  const bf16x4 b0 = to_bf(x);
  f32x4 h[4], g[4];
#pragma unroll
  for (int ob = 0; ob < 4; ob++) {
    // the bias seeds the accumulator: lane (s, q) holds features ob*16 + 4q + j
    f32x4 z = mfma(wop(wpk, kOpW1 + ob, lane), b0, *(const f32x4*)(P + kB1 + ob * 16 + 4 * q));
    if (Z && valid) *(f32x4*)(Z + (size_t)n * 192 + ob * 16 + 4 * q) = z;
#pragma unroll
    for (int j = 0; j < 4; j++) h[ob][j] = gelu(z[j]);
  }
#pragma unroll
  for (int L = 0; L < 2; L++) {
    const int wo = L == 0 ? kOpW2 : kOpW3, bo = L == 0 ? kB2 : kB3;
    bf16x4 hb[4];
#pragma unroll
    for (int kb = 0; kb < 4; kb++) hb[kb] = to_bf(h[kb]);
#pragma unroll
    for (int ob = 0; ob < 4; ob++) {
      f32x4 z = *(const f32x4*)(P + bo + ob * 16 + 4 * q);
#pragma unroll
      for (int kb = 0; kb < 4; kb++) z = mfma(wop(wpk, wo + 4 * kb + ob, lane), hb[kb], z);
      if (Z && valid) *(f32x4*)(Z + (size_t)n * 192 + (L + 1) * 64 + ob * 16 + 4 * q) = z;
#pragma unroll
      for (int j = 0; j < 4; j++) g[ob][j] = gelu(z[j]);
    }
#pragma unroll
    for (int ob = 0; ob < 4; ob++) h[ob] = g[ob];
  }
  f32x4 y = *(const f32x4*)(P + kB4 + 4 * q);
#pragma unroll
  for (int kb = 0; kb < 4; kb++) y = mfma(wop(wpk, kOpW4 + kb, lane), to_bf(h[kb]), y);
  if (valid) *(f32x4*)(Y + (size_t)n * 16 + 4 * q) = y;
}

// Backward pass of the mean squared error over the 9 outputs:
// dZ4 = 2 (Y - T) / (9 N), then dA = W dZ and dZ = dA * GELU'(Z) down to
// layer 1.  dZ receives dZ1|dZ2|dZ3|dZ4 as [N][208]; lossp[wave] the summed
// squared error of the wave's samples.
__global__ void __launch_bounds__(256) k_mlp_bwd(const float* __restrict__ T,
                                                 const int* __restrict__ idx, int N,
                                                 const short* __restrict__ wpk,
                                                 const float* __restrict__ Y,
                                                 const float* __restrict__ Z,
                                                 float* __restrict__ dZ,
                                                 double* __restrict__ lossp) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int s = lane & 15, q = lane >> 4;
  const int gw = blockIdx.x * (blockDim.x >> 6) + wave;
  const int s0 = gw * 16;
  if (s0 >= N) return;
  const int n = s0 + s;
  const bool valid = n < N;
  const int src = valid ? (idx ? idx[n] : n) : 0;
  const float scale = 2.0f / (9.0f * (float)N);
  f32x4 d4 = {0.f, 0.f, 0.f, 0.f};
  double se = 0.0;
  if (valid) {
    const f32x4 y = *(const f32x4*)(Y + (size_t)n * 16 + 4 * q);
    const f32x4 t = *(const f32x4*)(T + (size_t)src * 16 + 4 * q);
#pragma unroll
    for (int j = 0; j < 4; j++)
      if (4 * q + j < kNout) {
        const float e = y[j] - t[j];
        d4[j] = scale * e;
        se += (double)e * e;
      }
    *(f32x4*)(dZ + (size_t)n * 208 + 192 + 4 * q) = d4;
  }
  for (int off = 32; off > 0; off >>= 1) se += __shfl_xor(se, off, 64);
  if (lane == 0) lossp[gw] = se;
  // dA3 = W4 dZ4 (one K block of 16 outputs)
  f32x4 d[4];
  const bf16x4 b4 = to_bf(d4);
#pragma unroll
  for (int ib = 0; ib < 4; ib++) {
    f32x4 a = mfma(wop(wpk, kOpB4 + ib, lane), b4, f32x4{0.f, 0.f, 0.f, 0.f});
    f32x4 z = {0.f, 0.f, 0.f, 0.f};
    if (valid) z = *(const f32x4*)(Z + (size_t)n * 192 + 128 + ib * 16 + 4 * q);
#pragma unroll
    for (int j = 0; j < 4; j++) d[ib][j] = a[j] * dgelu(z[j]);
    if (valid) *(f32x4*)(dZ + (size_t)n * 208 + 128 + ib * 16 + 4 * q) = d[ib];
  }
#pragma unroll
  for (int L = 1; L >= 0; L--) {  // layer 3 weights (L=1) then layer 2 (L=0)
    const int wo = L == 1 ? kOpB3 : kOpB2;
    bf16x4 db[4];
#pragma unroll
    for (int kb = 0; kb < 4; kb++) db[kb] = to_bf(d[kb]);
    f32x4 e[4];
#pragma unroll
    for (int ib = 0; ib < 4; ib++) {
      f32x4 a = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kb = 0; kb < 4; kb++) a = mfma(wop(wpk, wo + 4 * ib + kb, lane), db[kb], a);
      f32x4 z = {0.f, 0.f, 0.f, 0.f};
      if (valid) z = *(const f32x4*)(Z + (size_t)n * 192 + L * 64 + ib * 16 + 4 * q);
#pragma unroll
      for (int j = 0; j < 4; j++) e[ib][j] = a[j] * dgelu(z[j]);
      if (valid) *(f32x4*)(dZ + (size_t)n * 208 + L * 64 + ib * 16 + 4 * q) = e[ib];
    }
#pragma unroll
    for (int ib = 0; ib < 4; ib++) d[ib] = e[ib];
  }
}

// Weight gradients dW_k[in][out] = sum_s A_k[s][in] dZ_k[s][out] over one
// chunk of samples, one wave per 16x16 block of one layer (40 blocks):
// M = in, N = out, K = samples.  The waves of the first input block of each
// layer also sum their 16 dZ columns in FP32, the bias gradients.
// part[chunk][kNP] receives the chunk sums; every entry is written.
__global__ void __launch_bounds__(64) k_mlp_dw(const float* __restrict__ X,
                                               const int* __restrict__ idx, int N,
                                               const float* __restrict__ Z,
                                               const float* __restrict__ dZ, int chunk,
                                               float* __restrict__ part) {
  const int lane = threadIdx.x, s = lane & 15, q = lane >> 4;
  int blk = blockIdx.x;
  int layer, ib, ob;
  if (blk < 4) { layer = 0; ib = 0; ob = blk; }
  else if (blk < 20) { layer = 1; ib = (blk - 4) / 4; ob = (blk - 4) % 4; }
  else if (blk < 36) { layer = 2; ib = (blk - 20) / 4; ob = (blk - 20) % 4; }
  else { layer = 3; ib = blk - 36; ob = 0; }
  const int c0 = blockIdx.y * chunk, c1 = min(N, c0 + chunk);
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  float bsum = 0.f;
  for (int s0 = c0; s0 < c1; s0 += 16) {
    bf16x4 a, b;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int n = s0 + 4 * q + j;
      float av = 0.f, bv = 0.f;
      if (n < c1) {
        if (layer == 0) {
          av = X[(size_t)(idx ? idx[n] : n) * 16 + s];
        } else {
          av = gelu(Z[(size_t)n * 192 + (layer - 1) * 64 + ib * 16 + s]);
        }
        bv = layer < 3 ? dZ[(size_t)n * 208 + layer * 64 + ob * 16 + s]
                       : dZ[(size_t)n * 208 + 192 + s];
      }
      a[j] = f2bf(av);
      b[j] = f2bf(bv);
      bsum += bv;
    }
    acc = mfma(a, b, acc);
  }
  const int woff = layer == 0 ? kW1 : layer == 1 ? kW2 : layer == 2 ? kW3 : kW4;
  const int ld = layer == 3 ? 16 : 64;
  for (int j = 0; j < 4; j++)
    part[(size_t)blockIdx.y * kNP + woff + (ib * 16 + 4 * q + j) * ld + ob * 16 + s] = acc[j];
  if (ib == 0) {  // wave-uniform
    bsum += __shfl_xor(bsum, 16, 64);
    bsum += __shfl_xor(bsum, 32, 64);
    const int boff = layer == 0 ? kB1 : layer == 1 ? kB2 : layer == 2 ? kB3 : kB4;
    if (q == 0) part[(size_t)blockIdx.y * kNP + boff + ob * 16 + s] = bsum;
  }
}

__device__ __forceinline__ bool frozen(int p) {
  if (p >= kW1 && p < kB1) return (p - kW1) / 64 >= kNin;       // W1 rows past the inputs
  if (p >= kW4 && p < kB4) return (p - kW4) % 16 >= kNout;      // W4 columns past the outputs
  if (p >= kB4) return p - kB4 >= kNout;
  return false;
}

// Sum the chunk partials in a fixed order (four interleaved slices of the
// chunks per parameter, then the slices in order); then either store the
// gradient or take an Adam step and refresh the BF16 copy.  256 threads:
// 64 parameters x 4 slices.
constexpr int kRedParams = 64;
__global__ void __launch_bounds__(256) k_mlp_reduce(const float* __restrict__ part, int nchunk,
                                                    float* __restrict__ grad,
                                                    float* __restrict__ P, float* __restrict__ m,
                                                    float* __restrict__ v,
                                                    short* __restrict__ wpk, int adam, float lr,
                                                    float c1, float c2) {
  __shared__ float sl[4][kRedParams];
  const int t = threadIdx.x & (kRedParams - 1), slice = threadIdx.x / kRedParams;
  const int p = blockIdx.x * kRedParams + t;
  float acc = 0.f;
  if (p < kNP)
    for (int c = slice; c < nchunk; c += 4) acc += part[(size_t)c * kNP + p];
  sl[slice][t] = acc;
  __syncthreads();
  if (slice != 0 || p >= kNP) return;
  float gsum = ((sl[0][t] + sl[1][t]) + sl[2][t]) + sl[3][t];
  if (frozen(p)) gsum = 0.f;
  if (!adam) {
    grad[p] = gsum;
    return;
  }
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float mm = b1 * m[p] + (1.f - b1) * gsum;
  const float vv = b2 * v[p] + (1.f - b2) * gsum * gsum;
  m[p] = mm;
  v[p] = vv;
  float w = P[p] - lr * (mm / c1) / (sqrtf(vv / c2) + eps);
  if (frozen(p)) w = 0.f;
  P[p] = w;
  pack_weight(wpk, p, f2bf(w));
}

__global__ void k_mlp_tobf(const float* __restrict__ P, short* __restrict__ wpk) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < kNP) pack_weight(wpk, p, f2bf(P[p]));
}

// ---- filtered DNS samples ---------------------------------------------------
// Box sums over 4x4x4 fine nodes of rho and the Favre products needed for the
// filtered velocity, temperature, subgrid stress and heat flux.
constexpr int kNF = 14;  // rho, ru, rv, rw, rT, ruu, ruv, ruw, rvv, rvw, rww, ruT, rvT, rwT
__global__ void k_box_sums(Geom g, Phys ph, const double* __restrict__ U,
                           const double* __restrict__ rw, const double* __restrict__ uw,
                           const double* __restrict__ vw, int bx, int by, int bz,
                           double* __restrict__ S) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= bx * by * bz) return;
  const int cx = t % bx, cy = (t / bx) % by, cz = t / (bx * by);
  double acc[kNF];
  for (int f = 0; f < kNF; f++) acc[f] = 0.0;
  const long long F = g.field;
  for (int dk = 0; dk < 4; dk++)
    for (int dj = 0; dj < 4; dj++)
      for (int di = 0; di < 4; di++) {
        const int ii = 4 * cx + di, jj = 4 * cy + dj, kk = 4 * cz + dk;
        const long long n2 = (long long)jj * g.P1 + ii;
        double r, u, v, w, T;
        if (kk == 0 || kk == g.n3 - 1) {
          const long long wo = (kk == 0 ? 0 : g.plane) + n2;
          r = rw[wo];
          u = uw[wo];
          v = vw[wo];
          w = 0.0;
          T = 1.0;
        } else {
          const long long n = (long long)kk * g.plane + n2;
          r = U[n];
          u = U[F + n] / r;
          v = U[2 * F + n] / r;
          w = U[3 * F + n] / r;
          T = U[4 * F + n] / r - ph.ggM * 0.5 * (u * u + v * v + w * w);
        }
        const double vals[kNF] = {r, r * u, r * v, r * w, r * T, r * u * u, r * u * v, r * u * w,
                                  r * v * v, r * v * w, r * w * w, r * u * T, r * v * T, r * w * T};
        for (int f = 0; f < kNF; f++) acc[f] += vals[f];
      }
  for (int f = 0; f < kNF; f++) S[(size_t)t * kNF + f] = acc[f] / 64.0;
}

// Features (filtered velocity gradients, rho, T) and targets (subgrid
// stress tau_ij, heat flux q_j) at the interior cells of the coarse lattice.
__global__ void k_box_samples(int bx, int by, int bz, const double* __restrict__ S,
                              const double* __restrict__ zc, double hx, double hy,
                              float* __restrict__ X, float* __restrict__ Tg) {
  const int ix = bx - 2, iy = by - 2, iz = bz - 2;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ix * iy * iz) return;
  const int cx = 1 + t % ix, cy = 1 + (t / ix) % iy, cz = 1 + t / (ix * iy);
  auto at = [&](int x, int y, int z) { return S + ((size_t)(z * by + y) * bx + x) * kNF; };
  auto vel = [&](const double* c, int i) { return c[1 + i] / c[0]; };
  const double* c = at(cx, cy, cz);
  const double* xp = at(cx + 1, cy, cz);
  const double* xm = at(cx - 1, cy, cz);
  const double* yp = at(cx, cy + 1, cz);
  const double* ym = at(cx, cy - 1, cz);
  const double* zp = at(cx, cy, cz + 1);
  const double* zm = at(cx, cy, cz - 1);
  const double dz = zc[cz + 1] - zc[cz - 1];
  float* x = X + (size_t)t * 16;
  for (int i = 0; i < 3; i++) {
    x[3 * i + 0] = (float)((vel(xp, i) - vel(xm, i)) / (2.0 * hx));
    x[3 * i + 1] = (float)((vel(yp, i) - vel(ym, i)) / (2.0 * hy));
    x[3 * i + 2] = (float)((vel(zp, i) - vel(zm, i)) / dz);
  }
  const double r = c[0];
  x[9] = (float)r;
  x[10] = (float)(c[4] / r);
  for (int f = 11; f < 16; f++) x[f] = 0.f;
  const double u[3] = {vel(c, 0), vel(c, 1), vel(c, 2)}, T = c[4] / r;
  float* y = Tg + (size_t)t * 16;
  y[0] = (float)(c[5] - r * u[0] * u[0]);
  y[1] = (float)(c[6] - r * u[0] * u[1]);
  y[2] = (float)(c[7] - r * u[0] * u[2]);
  y[3] = (float)(c[8] - r * u[1] * u[1]);
  y[4] = (float)(c[9] - r * u[1] * u[2]);
  y[5] = (float)(c[10] - r * u[2] * u[2]);
  y[6] = (float)(c[11] - r * u[0] * T);
  y[7] = (float)(c[12] - r * u[1] * T);
  y[8] = (float)(c[13] - r * u[2] * T);
  for (int f = 9; f < 16; f++) y[f] = 0.f;
}

// ---- coupling to a coarse run -----------------------------------------------
// Features of every interior node of the coarse state, normalised.
// IN32: U is the FP32 predictor U* (stage 2 of a mixed instance).
template <bool IN32>
__global__ void k_les_features(Geom g, Phys ph, const double* __restrict__ U,
                               const double* __restrict__ invJac, const double* __restrict__ uw,
                               const double* __restrict__ vw, const float* __restrict__ norm,
                               float* __restrict__ X) {
  // interior nodes < 2^31 (checked when the closure is attached): 32-bit
  // index arithmetic
  const int NI = (int)(g.plane * (g.n3 - 2)), PLn = (int)g.plane;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= NI) return;
  const int kk = 1 + t / PLn, rem = t - (kk - 1) * PLn, jj = rem / g.P1, ii = rem - jj * g.P1;
  const long long F = g.field;
  using UT = std::conditional_t<IN32, float, double>;
  const UT* __restrict__ Ut = (const UT*)U;
  auto ld = [&](long long i) -> double { return (double)Ut[i]; };
  auto vel = [&](int i, int j, int k, double* o) {
    if (k == 0 || k == g.n3 - 1) {
      const long long wo = (k == 0 ? 0 : g.plane) + (long long)j * g.P1 + i;
      o[0] = uw[wo];
      o[1] = vw[wo];
      o[2] = 0.0;
      return;
    }
    const long long n = (long long)k * g.plane + (long long)j * g.P1 + i;
    const double ir = 1.0 / ld(n);
    o[0] = ld(F + n) * ir;
    o[1] = ld(2 * F + n) * ir;
    o[2] = ld(3 * F + n) * ir;
  };
  double xp[3], xm[3], yp[3], ym[3], zp[3], zm[3];
  vel(wrapi(ii + 1, g.P1), jj, kk, xp);
  vel(wrapi(ii - 1, g.P1), jj, kk, xm);
  vel(ii, wrapi(jj + 1, g.P2), kk, yp);
  vel(ii, wrapi(jj - 1, g.P2), kk, ym);
  vel(ii, jj, kk + 1, zp);
  vel(ii, jj, kk - 1, zm);
  const double fz = 0.5 * ph.invdXi * invJac[kk];
  float f[11];
  for (int i = 0; i < 3; i++) {
    f[3 * i + 0] = (float)((xp[i] - xm[i]) * 0.5 * ph.invdx);
    f[3 * i + 1] = (float)((yp[i] - ym[i]) * 0.5 * ph.invdy);
    f[3 * i + 2] = (float)((zp[i] - zm[i]) * fz);
  }
  const long long n = (long long)kk * g.plane + (long long)jj * g.P1 + ii;
  const double r = ld(n), ir = 1.0 / r;
  const double u = ld(F + n) * ir, v = ld(2 * F + n) * ir, w = ld(3 * F + n) * ir;
  f[9] = (float)r;
  f[10] = (float)(ld(4 * F + n) * ir - ph.ggM * 0.5 * (u * u + v * v + w * w));
  float* x = X + (long long)t * 16;
  for (int c = 0; c < 11; c++) x[c] = (f[c] - creadf(norm, c)) / creadf(norm, 11 + c);
  for (int c = 11; c < 16; c++) x[c] = 0.f;
}

// Subtract fac*dt * div(subgrid flux) from the stage output: momentum flux
// tau_ij, energy flux q_j + ggM u_i tau_ij (zero on the walls).  IN32 / OUT32:
// the stage input / output is the FP32 predictor U* (mixed instances).
template <bool IN32, bool OUT32>
__global__ void k_les_apply(Geom g, Phys ph, const double* __restrict__ Uin,
                            const double* __restrict__ invJac, const float* __restrict__ Y,
                            const float* __restrict__ norm, double fac, double* __restrict__ Uout) {
  const int NI = (int)(g.plane * (g.n3 - 2)), PLn = (int)g.plane;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= NI) return;
  const int kk = 1 + t / PLn, rem = t - (kk - 1) * PLn, jj = rem / g.P1, ii = rem - jj * g.P1;
  const long long F = g.field;
  using UT = std::conditional_t<IN32, float, double>;
  const UT* __restrict__ Ut = (const UT*)Uin;
  auto ld = [&](long long i) -> double { return (double)Ut[i]; };
  // flux of node (i,j,k) in direction d: comps 0..2 momentum, 3 energy
  auto flux = [&](int i, int j, int k, int d, double* o) {
    if (k == 0 || k == g.n3 - 1) {
      o[0] = o[1] = o[2] = o[3] = 0.0;
      return;
    }
    const long long m = ((long long)(k - 1) * g.P2 + j) * g.P1 + i;
    const float* y = Y + m * 16;
    double tau[9], q[3];
    const int map[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};  // tau_ab from the 6 stored components
    // norm: feature mean[11], feature std[11], target mean[9], target std[9]
    for (int a = 0; a < 9; a++)
      tau[a] = (double)y[map[a]] * creadf(norm, 31 + map[a]) + creadf(norm, 22 + map[a]);
    for (int a = 0; a < 3; a++) q[a] = (double)y[6 + a] * creadf(norm, 37 + a) + creadf(norm, 28 + a);
    const long long n = (long long)k * g.plane + (long long)j * g.P1 + i;
    const double ir = 1.0 / ld(n);
    const double u[3] = {ld(F + n) * ir, ld(2 * F + n) * ir, ld(3 * F + n) * ir};
    for (int a = 0; a < 3; a++) o[a] = tau[3 * a + d];
    o[3] = q[d] + ph.ggM * (u[0] * tau[d] + u[1] * tau[3 + d] + u[2] * tau[6 + d]);
  };
  double fp[4], fm[4], div[4] = {0.0, 0.0, 0.0, 0.0};
  flux(wrapi(ii + 1, g.P1), jj, kk, 0, fp);
  flux(wrapi(ii - 1, g.P1), jj, kk, 0, fm);
  for (int c = 0; c < 4; c++) div[c] += (fp[c] - fm[c]) * 0.5 * ph.invdx;
  flux(ii, wrapi(jj + 1, g.P2), kk, 1, fp);
  flux(ii, wrapi(jj - 1, g.P2), kk, 1, fm);
  for (int c = 0; c < 4; c++) div[c] += (fp[c] - fm[c]) * 0.5 * ph.invdy;
  flux(ii, jj, kk + 1, 2, fp);
  flux(ii, jj, kk - 1, 2, fm);
  for (int c = 0; c < 4; c++) div[c] += (fp[c] - fm[c]) * 0.5 * ph.invdXi * invJac[kk];
  const long long n = (long long)kk * g.plane + (long long)jj * g.P1 + ii;
  if constexpr (OUT32) {
    float* Uf = (float*)Uout;
    for (int c = 0; c < 4; c++)
      Uf[(c + 1) * F + n] = (float)((double)Uf[(c + 1) * F + n] - fac * ph.dt * div[c]);
  } else {
    for (int c = 0; c < 4; c++) Uout[(c + 1) * F + n] -= fac * ph.dt * div[c];
  }
}

}  // namespace cc

using namespace cc;

namespace {
thread_local std::string m_err;
int mfail(int code, const std::string& msg) {
  m_err = msg;
  return code;
}
#define MCHK(x)                                                                        \
  do {                                                                                 \
    hipError_t e_ = (x);                                                               \
    if (e_ != hipSuccess) return mfail(CC_EHIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct Mlp {
  float *P = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr, *part = nullptr;
  short* wpk = nullptr;  // BF16 weights as MFMA operands (kNOps x 64 lanes x 4)
  float *X = nullptr, *T = nullptr, *Y = nullptr, *Z = nullptr, *dZ = nullptr;
  int* idx = nullptr;
  double* lossp = nullptr;
  int cap = 0, capIdx = 0, nchunk_cap = 0, loss_cap = 0, step = 0;
  hipStream_t s = nullptr;
};

int ensure(Mlp* m, int nbatch, int ndata) {
  if (nbatch > m->cap) {
    for (float** b : {&m->Y, &m->Z, &m->dZ})
      if (*b) MCHK(hipFree(*b));
    MCHK(hipMalloc((void**)&m->Y, (size_t)nbatch * 16 * sizeof(float)));
    MCHK(hipMalloc((void**)&m->Z, (size_t)nbatch * 192 * sizeof(float)));
    MCHK(hipMalloc((void**)&m->dZ, (size_t)nbatch * 208 * sizeof(float)));
    m->cap = nbatch;
  }
  if (ndata > m->capIdx) {
    for (float** b : {&m->X, &m->T})
      if (*b) MCHK(hipFree(*b));
    if (m->idx) MCHK(hipFree(m->idx));
    MCHK(hipMalloc((void**)&m->X, (size_t)ndata * 16 * sizeof(float)));
    MCHK(hipMalloc((void**)&m->T, (size_t)ndata * 16 * sizeof(float)));
    MCHK(hipMalloc((void**)&m->idx, (size_t)ndata * sizeof(int)));
    m->capIdx = ndata;
  }
  return 0;
}

// Samples per weight-gradient chunk: small enough that a batch of a few
// thousand samples still gives the gradient kernel thousands of waves, and
// at most kMaxChunks chunks so that the partials stay within 11 MB.
constexpr int kMinChunk = 128, kMaxChunks = 256;

struct Chunks {
  int n, len;
};

Chunks chunks_for(int N) {
  const int want = std::min(kMaxChunks, (N + kMinChunk - 1) / kMinChunk);
  int len = (N + want - 1) / want;
  len = (len + 15) / 16 * 16;
  return {(N + len - 1) / len, len};
}

// forward + backward + gradient partials of one batch (samples idx[b0..b0+nb)
// of the resident data set); lossp[(nb+15)/16] receives the summed squared
// error of each 16-sample tile.  One wave per block: a batch of a few
// thousand samples then still spreads over every CU.
int batch_grad(Mlp* m, const int* idx, int nb, Chunks ck, double* lossp) {
  const int nw = (nb + 15) / 16;
  hipLaunchKernelGGL(k_mlp_fwd, dim3(nw), dim3(64), 0, m->s, m->X, idx, nb, m->wpk, m->P, m->Y,
                     m->Z);
  hipLaunchKernelGGL(k_mlp_bwd, dim3(nw), dim3(64), 0, m->s, m->T, idx, nb, m->wpk, m->Y, m->Z,
                     m->dZ, lossp);
  hipLaunchKernelGGL(k_mlp_dw, dim3(40, ck.n), dim3(64), 0, m->s, m->X, idx, nb, m->Z, m->dZ,
                     ck.len, m->part);
  MCHK(hipGetLastError());
  return 0;
}

// per-tile squared errors of every batch of one epoch
int ensure_loss(Mlp* m, int count) {
  if (count > m->loss_cap) {
    if (m->lossp) MCHK(hipFree(m->lossp));
    MCHK(hipMalloc((void**)&m->lossp, (size_t)count * sizeof(double)));
    m->loss_cap = count;
  }
  return 0;
}

int ensure_part(Mlp* m, int nchunk) {
  if (nchunk > m->nchunk_cap) {
    if (m->part) MCHK(hipFree(m->part));
    MCHK(hipMalloc((void**)&m->part, (size_t)nchunk * kNP * sizeof(float)));
    MCHK(hipMemset(m->part, 0, (size_t)nchunk * kNP * sizeof(float)));
    m->nchunk_cap = nchunk;
  }
  return 0;
}

}  // namespace

namespace cc {
// used by cc_abi.hip: run the closure on a coarse instance during a stage
void closure_stage(void* mlp, const Inst& a, const float* norm, float* X, float* Y, int stage,
                   hipStream_t s) {
  Mlp* m = (Mlp*)mlp;
  const long long NI = a.g.plane * (a.g.n3 - 2);
  const double* Uin = stage == 1 ? a.UA : a.UB;
  double* Uout = stage == 1 ? a.UB : a.UA;
  // mixed instances keep the stage-1 predictor U* (in UB) as FP32
  const int in32 = a.u32 && stage == 2, out32 = a.u32 && stage == 1;
  const int B = 256;
  const dim3 grid((unsigned)((NI + B - 1) / B));
  if (in32)
    hipLaunchKernelGGL(k_les_features<true>, grid, dim3(B), 0, s, a.g, a.ph, Uin, a.invJac, a.uw,
                       a.vw, norm, X);
  else
    hipLaunchKernelGGL(k_les_features<false>, grid, dim3(B), 0, s, a.g, a.ph, Uin, a.invJac, a.uw,
                       a.vw, norm, X);
  hipLaunchKernelGGL(k_mlp_fwd, dim3((unsigned)((NI + 63) / 64)), dim3(256), 0, s, X,
                     (const int*)nullptr, (int)NI, m->wpk, m->P, Y, (float*)nullptr);
  const double fac = stage == 1 ? 0.5 : 1.0;
  if (in32)
    hipLaunchKernelGGL((k_les_apply<true, false>), grid, dim3(B), 0, s, a.g, a.ph, Uin, a.invJac, Y,
                       norm, fac, Uout);
  else if (out32)
    hipLaunchKernelGGL((k_les_apply<false, true>), grid, dim3(B), 0, s, a.g, a.ph, Uin, a.invJac, Y,
                       norm, fac, Uout);
  else
    hipLaunchKernelGGL((k_les_apply<false, false>), grid, dim3(B), 0, s, a.g, a.ph, Uin, a.invJac, Y,
                       norm, fac, Uout);
}
void box_filter(const Inst& a, int bx, int by, int bz, double* S, const double* zc, float* X,
                float* T, hipStream_t s) {
  const int B = 256;
  hipLaunchKernelGGL(k_box_sums, dim3((bx * by * bz + B - 1) / B), dim3(B), 0, s, a.g, a.ph, a.UA,
                     a.rw, a.uw, a.vw, bx, by, bz, S);
  const int ns = (bx - 2) * (by - 2) * (bz - 2);
  hipLaunchKernelGGL(k_box_samples, dim3((ns + B - 1) / B), dim3(B), 0, s, bx, by, bz, S, zc,
                     4.0 * a.ph.dx, 4.0 * a.ph.dy, X, T);
}
}  // namespace cc

extern "C" {

int cc_mlp_nparams(void) { return kNP; }

int cc_mlp_create(unsigned seed, void** out) {
  if (!out) return mfail(CC_EINVAL, "null argument");
  *out = nullptr;
  Mlp* m = new Mlp();
  // Xavier-uniform weights, zero biases
  std::vector<float> P(kNP, 0.f);
  std::mt19937 rng(seed);
  auto fill = [&](int off, int nin, int nout, int ld, int nin_used, int nout_used) {
    const float a = std::sqrt(6.0f / (float)(nin_used + nout_used));
    std::uniform_real_distribution<float> U(-a, a);
    for (int i = 0; i < nin; i++)
      for (int o = 0; o < nout; o++)
        P[off + i * ld + o] = (i < nin_used && o < nout_used) ? U(rng) : 0.f;
  };
  fill(kW1, 16, 64, 64, kNin, 64);
  fill(kW2, 64, 64, 64, 64, 64);
  fill(kW3, 64, 64, 64, 64, 64);
  fill(kW4, 64, 16, 16, 64, kNout);
  if (hipStreamCreate(&m->s) != hipSuccess ||
      hipMalloc((void**)&m->P, kNP * sizeof(float)) != hipSuccess ||
      hipMalloc((void**)&m->m, kNP * sizeof(float)) != hipSuccess ||
      hipMalloc((void**)&m->v, kNP * sizeof(float)) != hipSuccess ||
      hipMalloc((void**)&m->grad, kNP * sizeof(float)) != hipSuccess ||
      hipMalloc((void**)&m->wpk, (size_t)kNOps * 64 * 4 * sizeof(short)) != hipSuccess) {
    cc_mlp_destroy(m);
    return mfail(CC_ENOMEM, "mlp allocation failed");
  }
  if (cc_mlp_set_params(m, P.data()) != 0) {
    cc_mlp_destroy(m);
    return CC_EHIP;
  }
  *out = m;
  return 0;
}

int cc_mlp_set_params(void* mlp, const float* in) {
  Mlp* m = (Mlp*)mlp;
  if (!m || !in) return mfail(CC_EINVAL, "null argument");
  MCHK(hipMemcpy(m->P, in, kNP * sizeof(float), hipMemcpyHostToDevice));
  MCHK(hipMemset(m->m, 0, kNP * sizeof(float)));
  MCHK(hipMemset(m->v, 0, kNP * sizeof(float)));
  m->step = 0;
  hipLaunchKernelGGL(k_mlp_tobf, dim3((kNP + 255) / 256), dim3(256), 0, m->s, m->P, m->wpk);
  MCHK(hipStreamSynchronize(m->s));
  return 0;
}

int cc_mlp_get_params(void* mlp, float* out) {
  Mlp* m = (Mlp*)mlp;
  if (!m || !out) return mfail(CC_EINVAL, "null argument");
  MCHK(hipMemcpy(out, m->P, kNP * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

int cc_mlp_forward(void* mlp, const float* X, int N, float* Y) {
  Mlp* m = (Mlp*)mlp;
  if (!m || !X || !Y || N < 1) return mfail(CC_EINVAL, "bad arguments");
  if (ensure(m, N, N)) return CC_EHIP;
  MCHK(hipMemcpyAsync(m->X, X, (size_t)N * 16 * sizeof(float), hipMemcpyHostToDevice, m->s));
  hipLaunchKernelGGL(k_mlp_fwd, dim3((N + 63) / 64), dim3(256), 0, m->s, m->X, (const int*)nullptr,
                     N, m->wpk, m->P, m->Y, (float*)nullptr);
  MCHK(hipGetLastError());
  MCHK(hipMemcpyAsync(Y, m->Y, (size_t)N * 16 * sizeof(float), hipMemcpyDeviceToHost, m->s));
  MCHK(hipStreamSynchronize(m->s));
  return 0;
}

int cc_mlp_grad(void* mlp, const float* X, const float* T, int N, double* loss, float* grad) {
  Mlp* m = (Mlp*)mlp;
  if (!m || !X || !T || N < 1) return mfail(CC_EINVAL, "bad arguments");
  if (ensure(m, N, N)) return CC_EHIP;
  const Chunks ck = chunks_for(N);
  const int nw = (N + 15) / 16;
  if (ensure_part(m, ck.n) || ensure_loss(m, nw)) return CC_EHIP;
  MCHK(hipMemcpyAsync(m->X, X, (size_t)N * 16 * sizeof(float), hipMemcpyHostToDevice, m->s));
  MCHK(hipMemcpyAsync(m->T, T, (size_t)N * 16 * sizeof(float), hipMemcpyHostToDevice, m->s));
  if (batch_grad(m, nullptr, N, ck, m->lossp)) return CC_EHIP;
  hipLaunchKernelGGL(k_mlp_reduce, dim3((kNP + kRedParams - 1) / kRedParams), dim3(256), 0, m->s,
                     m->part, ck.n, m->grad, m->P, m->m, m->v, m->wpk, 0, 0.f, 1.f, 1.f);
  MCHK(hipGetLastError());
  std::vector<double> lp(nw);
  MCHK(hipMemcpyAsync(lp.data(), m->lossp, nw * sizeof(double), hipMemcpyDeviceToHost, m->s));
  if (grad) MCHK(hipMemcpyAsync(grad, m->grad, kNP * sizeof(float), hipMemcpyDeviceToHost, m->s));
  MCHK(hipStreamSynchronize(m->s));
  double sse = 0.0;
  for (double x : lp) sse += x;
  if (loss) *loss = sse / (9.0 * N);
  return 0;
}

int cc_mlp_train(void* mlp, const float* X, const float* T, int N, int epochs, int batch,
                 float lr, unsigned seed, double* loss_hist) {
  Mlp* m = (Mlp*)mlp;
  if (!m || !X || !T || N < 16 || epochs < 1 || batch < 16) return mfail(CC_EINVAL, "bad arguments");
  batch = std::min(batch, N);
  if (ensure(m, batch, N)) return CC_EHIP;
  const Chunks ck = chunks_for(batch);
  const int nw = (batch + 15) / 16, nbatches = N / batch;
  if (ensure_part(m, ck.n) || ensure_loss(m, nbatches * nw)) return CC_EHIP;
  std::vector<double> lp(loss_hist ? (size_t)nbatches * nw : 0);
  MCHK(hipMemcpyAsync(m->X, X, (size_t)N * 16 * sizeof(float), hipMemcpyHostToDevice, m->s));
  MCHK(hipMemcpyAsync(m->T, T, (size_t)N * 16 * sizeof(float), hipMemcpyHostToDevice, m->s));
  std::vector<int> perm(N);
  for (int i = 0; i < N; i++) perm[i] = i;
  std::mt19937 rng(seed);
  for (int ep = 0; ep < epochs; ep++) {
    MCHK(hipStreamSynchronize(m->s));  // the previous epoch's copy of perm has been consumed
    std::shuffle(perm.begin(), perm.end(), rng);
    MCHK(hipMemcpyAsync(m->idx, perm.data(), (size_t)N * sizeof(int), hipMemcpyHostToDevice, m->s));
    for (int bi = 0; bi < nbatches; bi++) {
      if (batch_grad(m, m->idx + (size_t)bi * batch, batch, ck, m->lossp + (size_t)bi * nw))
        return CC_EHIP;
      m->step++;
      const float c1 = 1.f - std::pow(0.9f, (float)m->step);
      const float c2 = 1.f - std::pow(0.999f, (float)m->step);
      hipLaunchKernelGGL(k_mlp_reduce, dim3((kNP + kRedParams - 1) / kRedParams), dim3(256), 0,
                         m->s, m->part, ck.n, m->grad, m->P, m->m, m->v, m->wpk, 1, lr, c1, c2);
    }
    MCHK(hipGetLastError());
    if (loss_hist) {  // one read per epoch; batch sums, then the epoch, in order
      MCHK(hipMemcpyAsync(lp.data(), m->lossp, lp.size() * sizeof(double),
                          hipMemcpyDeviceToHost, m->s));
      MCHK(hipStreamSynchronize(m->s));
      double ep_sse = 0.0;
      for (int bi = 0; bi < nbatches; bi++) {
        double t = 0.0;
        for (int w = 0; w < nw; w++) t += lp[(size_t)bi * nw + w];
        ep_sse += t;
      }
      loss_hist[ep] = ep_sse / (9.0 * (double)nbatches * batch);
    }
  }
  MCHK(hipStreamSynchronize(m->s));
  return 0;
}

void cc_mlp_destroy(void* mlp) {
  Mlp* m = (Mlp*)mlp;
  if (!m) return;
  for (float* b : {m->P, m->m, m->v, m->grad, m->part, m->X, m->T, m->Y, m->Z, m->dZ})
    if (b) (void)hipFree(b);
  if (m->wpk) (void)hipFree(m->wpk);
  if (m->idx) (void)hipFree(m->idx);
  if (m->lossp) (void)hipFree(m->lossp);
  if (m->s) (void)hipStreamDestroy(m->s);
  delete m;
}

const char* cc_mlp_last_error(void) { return m_err.c_str(); }

}  // extern "C"
