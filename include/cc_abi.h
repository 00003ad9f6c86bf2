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

/* The same diagnostics with no device allocation (no counterpart in the
 * reference; for callers that cannot allocate device memory: code under HIP
 * graph capture, tight serving loops, harnesses that forbid it).  cc_diag
 * allocates, on its first call, a copy of var (7 n1 n2 n3 doubles) and its
 * partial sums; this variant reads the device state node by node and keeps
 * its partial sums, results and x3_file in a static device array of 41 KB
 * per GPU, held until the process ends.  Limit: n3 <= 4096 (CC_EINVAL
 * beyond).  Calls on one GPU share that array: calls that overlap on one GPU
 * race.  Same results as cc_diag. */
int cc_diag_noalloc(void* handle, const double* x3_file, double x2_end, double* out);

/* Current state as var(n1,n2,n3,7), every halo value as the reference holds
 * it (periodic images, stale corners and wall rings). */
int cc_download(void* handle, double* var, int* nstep, double* time);

/* First offending node (1-based i,j,k) of the last blow-up, or zeros. */
int cc_blown(void* handle, int* ijk);

/* Per-launch timing of each RK stage's main kernel (no counterpart in the
 * reference, which times 10 steps with cpu_time at main_v1.9.f90:322,1042):
 * mode 1 resets and enables HIP-event timing on the handle's stream for the
 * stage launches of every step; mode -k (k >= 2) does the same for the steps
 * whose number is a multiple of k (a sample that leaves the other steps
 * without timing events); mode 0 disables it; any other mode only queries.
 * out[0] = summed kernel time in ms, out[1] = number of timed launches since
 * the reset (out may be NULL). */
int cc_profile(void* handle, int mode, double* out);

/* Tiling the fused stage kernel uses on this handle (no counterpart in the
 * reference, which decomposes the domain over MPI ranks at
 * main_v1.9.f90:153-200): out[0], out[1] = tile nodes in x and y, out[2] =
 * planes per block, out[3] = blocks per launch, out[4] = the device's CUs,
 * out[5] = resident blocks per CU.  Chosen at cc_create to minimise
 * ceil(blocks / (out[4] * out[5])) * (planes per block + 3); ignored by
 * fp64-strict. */
int cc_tiling(void* handle, int* out);

/* Override that tiling (no counterpart in the reference): kchunk planes per
 * block and, in the mixed mode, the tile shape (variant 0 = 32x16, 1 = 32x8,
 * 2 = 16x16, 3 = 64x8; -1 keeps the current shape; FP64 has 0 only).
 * kchunk <= 0 restores the automatic choice.  Results do not depend on the
 * tiling: every choice computes the same values, which is what the override
 * lets a test check.  CC_EINVAL for fp64-strict and batch members. */
int cc_set_tiling(void* handle, int variant, int kchunk);

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

/* Wall conditions of the reference's parameter studies.  Each study is a
 * copy of the v1.9 solver whose bottom-wall block of the boundary update is
 * changed (top wall unchanged):
 *   s1..s6   v1.9_STWNV/v1.9_s1/codes/main_v1.9_s1.f90:609 (s3: :609-610),
 *            spanwise travelling wave v = amp sin(kx x + omega t), on odd j
 *            with phase_odd added;
 *   rd1..rd3 v1.9_RD/v1.9_rd1/codes/main_v1.9_rd1.f90:611-640, three discs
 *            of radius R = x2(n2)/10 at x = 2R, y = 2R, 5R, 8R rotating at
 *            disc_omega;
 *   g1, gst* v1.9_groov/v1.9_g1/codes/main_v1.9_g1.f90:609-618 and
 *            v1.9_groov_strat/v1.9_gst1..5: stripes (bands mod(6j/n2,2)==0, diagonal
 *            bands, or wavy bands) where the one-sided du/dz = dv/dz = grad,
 *            no-slip elsewhere;
 *   st1..st4 v1.9_strat/v1.9_st1/codes/main_v1.9_st1.f90:612: fixed wall
 *            temperature Tw, or (st3, st4) a fixed one-sided dT/dz.
 * Like the reference, the condition is applied at every wall node including
 * the periodic duplicates, each at its own file coordinates.  A zero-filled
 * cc_wall is the reference's isothermal no-slip wall. */
#define CC_WALL_NOSLIP 0
#define CC_WALL_WAVE 1
#define CC_WALL_DISCS 2
#define CC_WALL_STRIPES 3
#define CC_STRIPES_BANDS 0
#define CC_STRIPES_DIAGONAL 1
#define CC_STRIPES_WAVY 2
#define CC_WALLT_ONE 0   /* T = 1 */
#define CC_WALLT_FIXED 1 /* T = Tw */
#define CC_WALLT_GRAD 2  /* one-sided dT/dz = dTdz */
typedef struct {
  int vel;      /* CC_WALL_* */
  int stripes;  /* CC_STRIPES_*, for CC_WALL_STRIPES */
  int temp;     /* CC_WALLT_* */
  int reserved;
  double amp, kx, omega, phase_odd; /* travelling wave */
  double disc_omega;                /* rotating discs */
  double grad;                      /* stripes: du/dz = dv/dz on the stripes */
  double Tw, dTdz;                  /* wall temperature or its gradient */
} cc_wall;

/* Set the condition of wall 0 (bottom, k = 1) or 1 (top, k = n3).  x1 (n1)
 * and x2 (n2) are the coordinates read from the solution file (the studies
 * evaluate their conditions at those F10.6 values); NULL keeps the analytic
 * mesh.  Call before cc_upload, or right after it before any step.  The
 * FP64 and mixed modes support it; the reference-order mode, batch members
 * and a non-zero slip length do not (CC_EINVAL). */
int cc_set_wall(void* handle, int wall, const cc_wall* w, const double* x1, const double* x2);

/* The condition a named study applies ("s1".."s6", "rd1".."rd3", "g1",
 * "gst1".."gst5", "st1".."st4"; "" or "ref" = the reference wall) and the
 * wall it applies to.  Host-only; needs no device. */
int cc_wall_preset(const char* name, int* wall, cc_wall* w);

/* ---- post-processing (v1.9/codes/avg_v1.7.f90, v1.9/codes/fluc.f90,
 * v1.8_validation/codes/qInvariant.f90) on the device ---------------------- */

/* Time averages of the plane statistics.  mode 1 restarts the accumulation,
 * 2 stops it, 0 only queries.  While on, every statistics step (istat) adds
 * the device's full-precision stat1Spc/stat2Spc/corrSpc to running sums, so
 * no moment history has to go through text; mean, rms (n3,7) and covar (n3,9)
 * are then computed as avg_v1.7.f90:239-283 does (mean = sum/count, rms =
 * sqrt(<q^2> - <q>^2), covariances <ab> - <a><b> in its order rhow, rhoT,
 * uv, uw, vw, vT, Tw, pw, pT).  Any output pointer may be NULL. */
int cc_average(void* handle, int mode, double* mean, double* rms, double* covar, int* count);

/* Fluctuation field var - mean(k) of all 7 variables (fluc.f90:55-62) and
 * the Q-criterion of the velocity fluctuations (qInvariant.f90:26-81) from
 * an instance's current state, computed on the device.  mean (n3,7); NULL =
 * the accumulated time average.  fluc (n1,n2,n3,7) and Q (n1,n2,n3) may be
 * NULL; Q needs the mesh x1, x2, x3 (the solution file's values, as the
 * reference program reads them from fluc.dat); Q = 0 on boundary nodes. */
int cc_postprocess(void* handle, const double* mean, const double* x1, const double* x2,
                   const double* x3, double* fluc, double* Q);

/* The same two computations on host arrays (for fields read from files):
 * var, out (n1,n2,n3,7); U (n1,n2,n3,3) = u', v', w'; Q (n1,n2,n3). */
int cc_pp_fluctuation(int n1, int n2, int n3, const double* var, const double* mean,
                      double* out);
int cc_pp_qcriterion(int n1, int n2, int n3, const double* x, const double* y, const double* z,
                     const double* U, double* Q);

/* ---- formatted point files (host, multi-threaded; nthreads <= 0: all cores
 * up to 16) --------------------------------------------------------------
 * Byte-identical to the Fortran runtime's output of the reference's writers:
 * records (3(1X,F10.6),nf(1X,E22.15)[,1X,I1]) in k, j, i order. */

/* A solution file as main_v1.9.f90:759-773 writes it (title with time and
 * nstep, variable names, ZONE line, a blank record before every k and j);
 * pattern (n1,n2,n3) adds the study programs' "Pattern" column (NULL: none).
 * Replaces the reference's record-by-record Fortran write. */
int cc_write_solution(const char* path, double time, int nstep, int n1, int n2, int n3,
                      const double* x1, const double* x2, const double* x3, const double* var,
                      const int* pattern, int nthreads);
/* Reads one (replaces main_v1.9.f90:210-251): time, nstep from the title
 * line, x1, x2, x3 and var (n1,n2,n3,7); any output pointer may be NULL.
 * A "Pattern" column (the halo-corrected solver's files) is skipped. */
int cc_read_solution(const char* path, int n1, int n2, int n3, double* x1, double* x2,
                     double* x3, double* var, double* time, int* nstep, int nthreads);
/* Grid of a solution file from its ZONE k=,j=,i= line. */
int cc_solution_dims(const char* path, int* n1, int* n2, int* n3);
/* Any point file: header (the caller's lines, each ending in '\n'), nf
 * fields (n1,n2,n3,nf), blank_records != 0 for the blank records before each
 * k and j (fluc.dat) or 0 (qInvariant.dat). */
int cc_write_points(const char* path, const char* header, int n1, int n2, int n3,
                    const double* x1, const double* x2, const double* x3, int nf,
                    const double* fields, const int* pattern, int blank_records, int nthreads);
/* Reads one: skips nheader lines, then n1*n2*n3 records of 3 + nf (+ 1 when
 * has_pattern) numbers, blank records ignored; x1..x3 and fields may be NULL. */
int cc_read_points(const char* path, int nheader, int n1, int n2, int n3, int nf, int has_pattern,
                   double* x1, double* x2, double* x3, double* fields, int nthreads);
/* One value as the writers format it: kind 0 = E22.15, 1 = F10.6; out
 * receives the field and a NUL (at least 23 bytes). */
int cc_format_field(double x, int kind, char* out);
const char* cc_io_last_error(void);

/* ---- learned subgrid closure (no counterpart in the reference, which has no
 * LES) ----------------------------------------------------------------------
 * An 11-64-64-64-9 GELU MLP: inputs the filtered velocity-gradient tensor
 * (9, row-major du_i/dx_j), filtered density and temperature; outputs the
 * subgrid stress (tau_11, tau_12, tau_13, tau_22, tau_23, tau_33) and heat
 * flux (q_1, q_2, q_3).  Forward, backward and weight-gradient passes run on
 * BF16 MFMA with FP32 accumulation; Adam keeps FP32 master weights.  Feature
 * and target rows are 16 floats (zero-padded).  Parameters are flattened as
 * W1[16][64], b1[64], W2[64][64], b2[64], W3[64][64], b3[64], W4[64][16],
 * b4[16]. */
int cc_mlp_nparams(void);
int cc_mlp_create(unsigned seed, void** mlp);
int cc_mlp_set_params(void* mlp, const float* params);
int cc_mlp_get_params(void* mlp, float* params);
int cc_mlp_forward(void* mlp, const float* X, int N, float* Y);
/* loss = mean over N samples and 9 outputs of (Y - T)^2, and its gradient. */
int cc_mlp_grad(void* mlp, const float* X, const float* T, int N, double* loss, float* grad);
/* Adam over shuffled mini-batches; loss_hist[epochs] (may be NULL). */
int cc_mlp_train(void* mlp, const float* X, const float* T, int N, int epochs, int batch,
                 float lr, unsigned seed, double* loss_hist);
void cc_mlp_destroy(void* mlp);
const char* cc_mlp_last_error(void);
/* Training samples from an instance's current state: 4x4x4 box filter
 * (Favre averages), features and subgrid targets at the interior cells of
 * the coarse lattice.  With X or T NULL only *M (the sample count) is set. */
int cc_filter_samples(void* handle, float* X, float* T, int* M);
/* Add the closure's subgrid fluxes (momentum tau_ij, energy q_j + ggM u_i
 * tau_ij) to every stage of cc_step; norm[40] = feature mean[11], std[11],
 * target mean[9], std[9].  mlp NULL removes it. */
int cc_set_closure(void* handle, void* mlp, const float* norm);
const char* cc_last_error(void);
int cc_abi_version(void);

#ifdef __cplusplus
}
#endif
#endif
