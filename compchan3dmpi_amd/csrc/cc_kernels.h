// cc_kernels.h -- host launchers of the HIP kernels (internal to the library).
#pragma once
#include <hip/hip_runtime.h>

#include "cc_device.h"

namespace cc {

// Number of plane-statistics quantities: 7 means, 7 mean squares, 9 products
// (main_v1.9.f90:993-1009).
constexpr int kNStat = 23;

// Wall pressure storage: bottom wall at pwall[0..plane), top at pwall + top_off.
// uw, vw, tw, mw: wall velocities, temperature and mu = T^0.7 [2][plane]
// (NULL in the reference-order path, whose walls are isothermal no-slip).
// ring: duplicate nodes of the wall planes [2][kRingFields][ring_size], used
// for wall w when bit w of wflags is set (a study wall condition).
struct WallP {
  double* p;
  long long top_off;
  double* uw = nullptr;
  double* vw = nullptr;
  double* tw = nullptr;
  double* mw = nullptr;
  double* ring = nullptr;
  int wflags = 0;
};

// ---- cc_strict.hip: reference-order FP64 stage ------------------------------
void launch_stage_strict(const Geom& g, const Phys& ph, const double* Uin, const double* Ubase,
                         double* Uout, double* W, double* rw, const double* jac,
                         const double* invJac, const double* force, int stage, int* halt,
                         unsigned long long* blow, int step, int n1, int n2, hipStream_t s,
                         hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);

// ev0/ev1 (optional) are recorded on the stream right before and right after
// the stage's main kernel, for per-launch timing.

// ---- cc_fused.hip: fused streaming stage (FP64 or mixed) --------------------
// One channel instance as the stage kernels see it.  Stage 1 reads UA and
// writes UB; stage 2 reads UB, takes UA as the RK base and writes UA.  With
// u32 (mixed precision) UB holds the stage-1 predictor U* as FP32, the
// precision its only uses (FP32 stencils) take; U^n and U^{n+1} in UA stay
// FP64.
struct Inst {
  Geom g;
  Phys ph;
  double* UA;
  double* UB;
  double* pw;  // wall pressure [2][plane]
  double* rw;  // wall density [2][plane]
  double* uw;  // wall velocities [2][plane]
  double* vw;
  double* tw;  // wall temperature [2][plane]
  double* mw;  // wall mu = T^0.7 [2][plane]
  double* ring;  // duplicate wall nodes [2][kRingFields][ring_size(P1, P2)]
  double* dw;    // duplicate-node correction of the wall-plane hD, as a pressure offset [2][plane]
  const double* vfile;  // the uploaded var(n1,n2,n3,7) (stale corner columns)
  const double* x1f;    // x1(n1), x2(n2) as read from the solution file
  const double* x2f;
  WallSpec ws[2];       // wall conditions (bottom, top)
  int wflags;           // bit w: wall w has a study condition
  double time;          // physical time of the step being taken
  const double* jac;
  const double* invJac;
  const double* force;  // tau_L, tau_U, rho_bar(n3)
  int* halt;
  unsigned long long* blow;
  int n1, n2;
  int kchunk, ntx, nty, nz;  // tiling of the stage kernel
  int u32;                   // UB holds U* as FP32 (mixed precision)
};
// tile variants: 0 = 32x16 (the only FP64 one), 1 = 32x8, 2 = 16x16, 3 = 64x8
void tile_of(int variant, int& tx, int& ty);
void set_tiling(Inst& a, int variant, int kchunk);
// one instance: the tile shape and planes per block of least launch time on
// the current device; returns the variant, and the CU count and resident
// blocks per CU it assumed (either pointer may be null)
int choose_tiling(Inst& a, int mixed, int* ncu_out, int* bpc_out);
void launch_stage_fused(const Inst& a, int mixed, int variant, int stage, int step,
                        hipStream_t s, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr,
                        bool wall = true);
// wall values and wall pressure after a stage (part of launch_stage_fused)
void launch_wall_fast(const Inst& a, int stage, int step, hipStream_t s);
// mixed precision: UB <- UA as FP32, so that UB's wall planes, which no
// stage writes, hold the frozen wall state in UB's FP32 layout
void launch_predictor_f32(const Inst& a, hipStream_t s);

// ---- cc_mlp.hip: learned subgrid closure ------------------------------------
// Subtract the closure's subgrid-flux divergence from a stage's output; X, Y
// hold 16 floats per interior node; norm: feature mean[11], std[11], target
// mean[9], std[9].
void closure_stage(void* mlp, const Inst& a, const float* norm, float* X, float* Y, int stage,
                   hipStream_t s);
// 4x4x4 box filter of an instance's state: S [bz][by][bx][14] sums, zc[bz]
// block-centre heights; X, T receive (bx-2)(by-2)(bz-2) feature/target rows.
void box_filter(const Inst& a, int bx, int by, int bz, double* S, const double* zc, float* X,
                float* T, hipStream_t s);
int batched_blocks_per_cu(int mixed, int variant);
// d_work: int4 (instance, tile_x, tile_y, z-chunk) items
void launch_stage_batched(const Inst* d_insts, const int4* d_work, int nwork, unsigned* counter,
                          unsigned base, int nblocks, int mixed, int variant, int stage, int step,
                          int ninst, int maxP1, int maxP2, hipStream_t s,
                          hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);

// ---- cc_common.hip ----------------------------------------------------------
// Vfull: the reference's var(n1,n2,n3,7) in Fortran order on the device.
// (dw: the fused path's wall divergence correction [2][plane], or NULL)
void launch_upload(const Geom& g, const Phys& ph, int n1, int n2, const double* Vfull, double* UA,
                   double* UB, double* W, WallP pw, double* rw, double* dw, double* consts,
                   hipStream_t s);
// wall arrays (uw, vw, tw, mw, dw) and the duplicate ring from the uploaded
// file, and the constant plane-sum terms for the current wall conditions
void launch_wall_init(const Geom& g, const Phys& ph, int n1, int n2, const double* Vfull,
                      WallP pw, double* dw, double* consts, hipStream_t s);
void launch_download(const Geom& g, const Phys& ph, int n1, int n2, const double* U,
                     WallP pw, const double* rw, const double* Vfull, double* Vout,
                     hipStream_t s);
// Bulk diagnostics on the reconstructed var (Vdev, from launch_download):
// out[16] = KE, rho_b[5], U_b[5], massflow[5]; part holds 1024 partials;
// x3f (n3) and x2e are the file's mesh values.
void launch_diag(int n1, int n2, int n3, double dy, const double* Vdev, double* part,
                 const double* x3f, double x2e, double* out, hipStream_t s);
// The same diagnostics read from the device state (node by node, no
// reconstructed copy), with part, out and x3f in the caller's scratch.
void launch_diag_state(const Geom& g, const Phys& ph, int n1, int n2, const double* U, WallP pw,
                       const double* rw, const double* Vfile, double* part, const double* x3f,
                       double x2e, double* out, hipStream_t s);
// Partial sums of the diagnostics' first pass, and the layout of the static
// scratch of the allocation-free diagnostics: [kDiagParts] partials, [16]
// results, [kDiagScratchMaxN3] x3 of the file.
constexpr int kDiagParts = 1024;
constexpr int kDiagScratchMaxN3 = 4096;
constexpr int kDiagOffOut = kDiagParts;
constexpr int kDiagOffX3 = kDiagParts + 16;
constexpr int kDiagScratchSize = kDiagOffX3 + kDiagScratchMaxN3;
constexpr int kMaxDevices = 64;
// Once per device, with that device made current: the device address of the
// static scratch (hipGetSymbolAddress); 0 or -1.  diag_scratch: that address,
// or null before the set-up.
int diag_scratch_init(int dev);
double* diag_scratch(int dev);
// tau[0], tau[1] -> force[0], force[1]; plane statistics -> stats[23][n3]
// and force[2 + kk] = mean density of plane kk.
// part: scratch of stats_scratch_size(n3) doubles
void launch_stats(const Geom& g, const Phys& ph, int n1, int n2, const double* U, WallP pw,
                  const double* rw, const double* consts, double* force, double* stats,
                  double* part, const int* halt, int step, hipStream_t s);
int stats_scratch_size(int n3);
// One member of a batched statistics step: a single instance's launch_stats
// arguments, its node ranges per plane, and where its blocks start in the
// two batched passes (blk0: first pass, fblk0: second pass).
struct StatDesc {
  Geom g;
  Phys ph;
  int n1, n2, split;
  int blk0, fblk0;
  const double* U;
  WallP pw;
  const double* rw;
  const double* C;
  double* force;
  double* stats;
  double* part;
  const int* halt;
};
// The descriptor of one instance, with the node ranges launch_stats uses
// (blk0 and fblk0 are left at 0 for the caller), and its block counts in the
// two passes.
StatDesc stats_desc(const Geom& g, const Phys& ph, int n1, int n2, const double* U, WallP pw,
                    const double* rw, const double* consts, double* force, double* stats,
                    double* part, const int* halt);
int stats_part_blocks(const StatDesc& d);
int stats_final_blocks(const StatDesc& d);
// The statistics step of nm members in two launches in all (d_desc on the
// device, blk0/fblk0 the running block counts); each member's results are
// bit-identical to its own launch_stats.
void launch_stats_batched(const StatDesc* d_desc, int nm, int part_blocks, int final_blocks,
                          int step, hipStream_t s);

// ---- cc_post.hip: post-processing ------------------------------------------
// sums[23*n3] += stats, sums[23*n3] += 1 (skipped on a halted run)
void launch_avg_accum(int n3, const double* stats, double* sums, const int* halt, int step,
                      hipStream_t s);
// out = var - mean(k) for nf fields of (n1,n2,n3)
void launch_fluc(int n1, int n2, int n3, int nf, const double* var, const double* mean,
                 double* out, hipStream_t s);
// Q-criterion of the velocity fluctuations U[3] on the mesh x, y, z
void launch_qcrit(int n1, int n2, int n3, const double* x, const double* y, const double* z,
                  const double* U, double* Q, hipStream_t s);

}  // namespace cc
