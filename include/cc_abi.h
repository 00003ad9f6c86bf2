/*
 * cc_abi.h -- C ABI of the MI355X time step of the compressible channel DNS
 * chanCompr (theparvezahmad/compChan3Dmpi).
 *
 * The reference has no callable API: its time step is the inline RK loop
 * body of v1.9/codes/main_v1.9.f90:311-706, and its intended seam is the
 * subroutine sequence of v1.10/codes/main_v1.10.f90:37-53
 * (calcFlux, calcSolV, share2P(solV), calcPrim, share2P(var), calcBC,
 * checkSoln, calcTauWall, avgSpatial).  Each entry point below names the
 * reference code it replaces.  A Fortran driver binds these with
 * ISO_C_BINDING (bind(C) interfaces; see compchan3dmpi_amd/fortran).
 *
 * Conventions: plain pointers and sizes; arrays in the reference's Fortran
 * order; every function returns 0 on success, a positive step number when
 * the blow-up check fired, or a negative error code (cc_last_error() gives
 * the message).  One handle = one channel instance = one HIP stream on the
 * current device; a handle is not re-entrant.
 */
#ifndef CC_ABI_H
#define CC_ABI_H

#ifdef __cplusplus
extern "C" {
#endif

#define CC_ABI_VERSION 1

/* precision modes */
#define CC_FP64_STRICT 0 /* FP64, reference evaluation order, no FMA contraction */
#define CC_FP64 1        /* FP64, fused streaming kernel */
#define CC_MIXED 2       /* FP32 flux/stencil arithmetic, FP64 state, RK and reductions */

/* error codes */
#define CC_EINVAL (-1)
#define CC_EHIP (-2)
#define CC_ENOMEM (-3)
#define CC_ESTATE (-4)

/* Run parameters.  n1,n2,n3 are the reference's nP = [streamwise, spanwise,
 * wall-normal] node counts including both periodic endpoints
 * (v1.9/codes/comdata.f90:4); Re, Mac, dt, istat, PexVal..PezVal are the
 * param.dat values read by comdata.f90:31-68.  gama = 1.4 and Pr = 0.71 are
 * compile-time constants in the reference (comdata.f90:11-12) and are not
 * parameters here either.  Ri_b and slip_len select physics the reference
 * does not have (buoyancy source, Navier wall slip); both 0 reproduce it. */
typedef struct {
  int n1, n2, n3;
  int istat;
  int precision;
  int reserved;
  double Re, Mac, dt;
  double PexVal, PeyVal, PezVal;
  double Ri_b;
  double slip_len;
} cc_params;

/* Mesh, Jacobian and one-sided wall coefficients (replaces comdata.f90:70-134
 * setupMesh and main_v1.9.f90:52-76), device buffers and stream. */
int cc_create(const cc_params* p, void** handle);

/* Initial state from var(n1,n2,n3,7) = (rho,u,v,w,E,T,p) as read from the
 * solution file (replaces main_v1.9.f90:210-301: conserved variables, mu,
 * kt and the first halo exchange). */
int cc_upload(void* handle, const double* var, int nstep, double time);

/* nsteps time steps: nstep += 1, time += dt, two RK stages each, the blow-up
 * check, and the wall-shear and plane statistics every istat steps that feed
 * the bulk forcing (replaces the loop body main_v1.9.f90:320-1025 minus its
 * file output).  Returns 0, or the step at which rho<0 or T<0 appeared (the
 * state is then that step's, as when the reference stops at :742). */
int cc_step(void* handle, int nsteps);

/* Statistics of the last istat step: tau[2] = tauxzL, tauxzU
 * (main_v1.9.f90:929-988); stat1, stat2 (n3,7) and corr (n3,9) in the
 * Fortran layout of stat1Spc, stat2Spc, corrSpc (:993-1025).  Any pointer
 * may be NULL. */
int cc_stats(void* handle, double* tau, double* stat1, double* stat2, double* corr);

/* Bulk diagnostics of main_v1.9.f90:799-906 on the current state:
 * out[0] = KE, out[1..5] rho_b, out[6..10] U_b, out[11..15] mass flow for the
 * planes [5,10,70,100,125].  x3_file (n3) and x2_end are the wall-normal
 * coordinates and x2(n2) as read from the solution file (F10.6): the reader
 * overwrites the analytic mesh with them (:238) and the diagnostics use
 * them (:821-830). */
int cc_diag(void* handle, const double* x3_file, double x2_end, double* out);

/* Current state as var(n1,n2,n3,7), every halo value as the reference holds
 * it (periodic images, stale corners and wall rings). */
int cc_download(void* handle, double* var, int* nstep, double* time);

/* First offending node (1-based i,j,k) of the last blow-up, or zeros. */
int cc_blown(void* handle, int* ijk);

/* Per-launch timing of each RK stage's main kernel (no counterpart in the
 * reference, which times 10 steps with cpu_time at main_v1.9.f90:322,1042):
 * mode 1 resets and enables HIP-event timing on the handle's stream, mode 0
 * disables it, any other mode only queries.  out[0] = summed kernel time in
 * ms, out[1] = number of timed launches since the reset (out may be NULL). */
int cc_profile(void* handle, int mode, double* out);

void cc_destroy(void* handle);

/* ---- batched ensembles (no counterpart in the reference, whose ensembles are
 * separate runs, v1.9_STWNV/bin/startAll.sh:4-18) ------------------------------
 * n instances of possibly different grids, Ri_b and slip lengths, advanced
 * together: each RK stage is one persistent launch that pulls (instance, tile,
 * z-chunk) work items from a queue.  Every member computes exactly what it
 * computes alone with cc_step.  All members share one precision mode (FP64 or
 * mixed). */
int cc_batch_create(const cc_params* p, int n, void** batch);
/* Handle of member i, for cc_upload, cc_stats, cc_diag, cc_download and
 * cc_blown (not cc_step, and not cc_destroy: the batch owns it). */
void* cc_batch_member(void* batch, int i);
/* nsteps steps of every member that has not stopped; all must stand at the
 * same step.  Returns 0, or the earliest step at which a member blew up (that
 * member stops there, the others go on). */
int cc_batch_step(void* batch, int nsteps);
/* Autotuner: times trial_steps batched steps for every compiled tile shape
 * and z-chunk, keeps the fastest, and restores the members' state.
 * out[0] = best ms per step, out[1..2] = tile (x, y), out[3] = z-chunk. */
int cc_batch_tune(void* batch, int trial_steps, double* out);
void cc_batch_destroy(void* batch);
const char* cc_last_error(void);
int cc_abi_version(void);

#ifdef __cplusplus
}
#endif
#endif
