// cc_kernels.h -- host launchers of the HIP kernels (internal to the library).
#pragma once
#include <hip/hip_runtime.h>

#include "cc_device.h"

namespace cc {

// Number of plane-statistics quantities: 7 means, 7 mean squares, 9 products
// (main_v1.9.f90:993-1009).
constexpr int kNStat = 23;

// Wall pressure storage: bottom wall at pwall[0..plane), top at pwall + top_off.
struct WallP {
  double* p;
  long long top_off;
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
struct FusedCfg {
  int tx, ty;   // tile in ii, jj
  int kchunk;   // planes per block along z
};
void launch_stage_fused(const Geom& g, const Phys& ph, int mixed, const FusedCfg& cfg,
                        const double* Uin, const double* Ubase, double* Uout, double* pw,
                        double* rw, const double* jac, const double* invJac, const double* force,
                        int stage, int* halt, unsigned long long* blow, int step, int n1, int n2,
                        hipStream_t s, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);

// ---- cc_common.hip ----------------------------------------------------------
// Vfull: the reference's var(n1,n2,n3,7) in Fortran order on the device.
void launch_upload(const Geom& g, const Phys& ph, int n1, int n2, const double* Vfull, double* UA,
                   double* UB, double* W, WallP pw, double* rw, double* consts, hipStream_t s);
void launch_download(const Geom& g, const Phys& ph, int n1, int n2, const double* U,
                     WallP pw, const double* rw, const double* Vfull, double* Vout,
                     hipStream_t s);
// Bulk diagnostics on the reconstructed var (Vdev, from launch_download):
// out[16] = KE, rho_b[5], U_b[5], massflow[5]; part holds 1024 partials;
// x3f (n3) and x2e are the file's mesh values.
void launch_diag(int n1, int n2, int n3, double dy, const double* Vdev, double* part,
                 const double* x3f, double x2e, double* out, hipStream_t s);
// tau[0], tau[1] -> force[0], force[1]; plane statistics -> stats[23][n3]
// and force[2 + kk] = mean density of plane kk.
void launch_stats(const Geom& g, const Phys& ph, int n1, int n2, const double* U, WallP pw,
                  const double* rw, const double* consts, double* force, double* stats,
                  const int* halt, int step, hipStream_t s);

}  // namespace cc
