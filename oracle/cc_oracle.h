/*
 * cc_oracle.h -- CPU restatement of one time step of the compressible channel
 * DNS chanCompr (reference: v1.10/codes/main_v1.9.f90, the halo-corrected copy
 * of v1.9/codes/main_v1.9.f90, plus v1.9/codes/comdata.f90).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product path (compchan3dmpi_amd, the HIP kernels) never links or calls it.
 *
 * Layout and evaluation order follow the reference exactly: Fortran
 * column-major arrays, solV(0:n1+1,0:n2+1,0:n3+1,5,2), var(0:n1+1,...,7),
 * point fluxes fD/gD/hD(n1,n2,n3,5), loops k-j-i with i innermost, and every
 * arithmetic expression evaluated left to right as the Fortran source groups
 * it.  Build with -ffp-contract=off so no multiply-add is fused.
 */
#ifndef CC_ORACLE_H
#define CC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc orc_t;

/* Mesh, one-sided coefficients and constants (comdata.f90:70-134,
 * main_v1.9.f90:52-76).  Pex/Pey/Pez are the cell-Peclet thresholds. */
orc_t *orc_create(int n1, int n2, int n3, double Re, double Mac, double dt,
                  int istat, double Pex, double Pey, double Pez);
void orc_destroy(orc_t *o);

/* Initial state from a primitive field var(n1,n2,n3,7) in Fortran order
 * (rho,u,v,w,E,T,p), as the reader leaves it (main_v1.9.f90:227-251), then
 * conserved variables, mu, kt (:284-294) and the first halo fill (:297-313). */
int orc_load(orc_t *o, const double *var, int nstep, double time);

/* x3 and x2(n2) as read back from the solution file (F10.6): the reader
 * overwrites the analytic mesh with them (main_v1.9.f90:238) and only the
 * bulk diagnostics use them (:821-830).  Default: the analytic mesh. */
void orc_set_x3_file(orc_t *o, const double *x3, double x2end);

/* Physics outside the reference (both 0 = reference): buoyancy body force
 * -Ri rho along z (work -ggM Ri rho w in the energy equation, and the
 * gravity term in the wall-pressure balance), and Navier slip of length Ls
 * on both walls: u_w = Ls du/dn with the one-sided coefficients of
 * main_v1.9.f90:59-83, so Ls -> inf is the zero-gradient wall of
 * v1.9_groov/v1.9_g1/codes/main_v1.9_g1.f90:611-612.  With Ls != 0 the
 * wall-plane halo ring holds periodic images (a stale u = 0 ring would put a
 * jump at the periodic seam).  Parity unpinned against the reference (it has
 * neither); pinned by the closed forms in tests/test_oracle_ext.py. */
void orc_set_ext(orc_t *o, double Ri, double Ls);

/* Wall conditions of the reference's parameter studies, one per wall
 * (wall 0 = bottom k=1, 1 = top k=n3; every study changes the bottom wall
 * only).  A zero-filled struct is the reference's isothermal no-slip wall.
 *   vel = ORC_WALL_WAVE: v = amp*sin(kx*x1 + omega*time [+ phase_odd on odd j])
 *         (v1.9_STWNV/v1.9_s1/codes/main_v1.9_s1.f90:609, ..._s3.f90:609-610);
 *   vel = ORC_WALL_DISCS: three discs of radius x2(n2)/10 centred at
 *         x = 2R, y = 2R, 5R, 8R turning at disc_omega
 *         (v1.9_RD/v1.9_rd1/codes/main_v1.9_rd1.f90:611-640).  Parity
 *         unpinned on a disc's axis: the reference's angle function
 *         (v1.9_RD/.../comdata.f90:180-193) leaves its result unset when a
 *         coordinate offset is exactly 0; this takes the axis angle there,
 *         and no tested grid has a node on an axis;
 *   vel = ORC_WALL_STRIPES: on the stripes u, v satisfy the one-sided
 *         du/dz = dv/dz = grad, no-slip elsewhere; stripes = 0 bands
 *         mod(6j/n2,2)==0 (v1.9_groov/v1.9_g1/codes/main_v1.9_g1.f90:609-618),
 *         1 diagonal bands (v1.9_groov_strat/v1.9_gst3/codes/main_v1.9_gst3.f90),
 *         2 wavy bands (..._gst5.f90);
 *   temp = ORC_WALLT_FIXED: T = Tw (v1.9_strat/v1.9_st1/codes/main_v1.9_st1.f90:612);
 *   temp = ORC_WALLT_GRAD: the one-sided dT/dz = dTdz
 *         (v1.9_strat/v1.9_st3/codes/main_v1.9_st3.f90:612).
 * x1, x2: the coordinates read from the solution file (the BC loop sees the
 * file's F10.6 values).  The loop covers every node 1..n1 x 1..n2 of the
 * wall, periodic duplicates included, each at its own coordinates, so the
 * duplicates can differ from their periodic images. */
enum { ORC_WALL_NOSLIP = 0, ORC_WALL_WAVE = 1, ORC_WALL_DISCS = 2, ORC_WALL_STRIPES = 3 };
enum { ORC_WALLT_ONE = 0, ORC_WALLT_FIXED = 1, ORC_WALLT_GRAD = 2 };
typedef struct {
  int vel, stripes, temp, reserved;
  double amp, kx, omega, phase_odd;
  double disc_omega, grad, Tw, dTdz;
} orc_wall;
void orc_set_wall(orc_t *o, int wall, const orc_wall *w, const double *x1, const double *x2);

/* Forcing state (tauxzL, tauxzU, stat1Spc(:,1)); zero until the first
 * statistics step, exactly as the reference's static initialisation. */
void orc_set_forcing(orc_t *o, double tauL, double tauU, const double *rhobar);

/* One RK stage (iter = 1 or 2): calcFlux, calcSolV, share2P(solV), calcPrim,
 * share2P(var), calcBC (main_v1.10.f90:37-44). */
void orc_stage(orc_t *o, int iter);

/* nsteps full time steps including the blow-up check and the istat-cadence
 * wall-shear and plane statistics.  Returns 0, or 1 when rho<0 or T<0 was
 * found (the step at which the reference would stop). */
int orc_step(orc_t *o, int nsteps);

/* Accessors.  var/solv are returned over 1..n in Fortran order
 * var(n1,n2,n3,7), U(n1,n2,n3,5) (current solution, level 1). */
void orc_get_var(const orc_t *o, double *var);
void orc_get_solv(const orc_t *o, int level, double *U);
void orc_get_mu(const orc_t *o, double *mu);
int orc_get_nstep(const orc_t *o);
double orc_get_time(const orc_t *o);
/* tau[2] = tauxzL, tauxzU; stat1/stat2 (n3,7), corr (n3,9), Fortran order. */
void orc_get_stats(const orc_t *o, double *tau, double *stat1, double *stat2,
                   double *corr);
/* coef[12] = f1d2c0..2, f2d1c0..2, b1d2c0..2, b2d1c0..2; mesh[n3*4] =
 * x3, jac, invJac, Xi; scal[8] = dx, dy, dXi, dz1, dzN, ggM, invRe, Pi. */
void orc_get_mesh(const orc_t *o, double *coef, double *mesh, double *scal);
/* Residual R(n1,n2,n3,5) of the last stage (updated nodes only, 0 elsewhere). */
void orc_get_residual(const orc_t *o, double *R);
/* Threads the node loops run on (OpenMP; 1 without it). */
int orc_threads(void);
/* Blow-up location of the last orc_step that returned 1: i,j,k (1-based). */
void orc_get_blown(const orc_t *o, int *ijk);

/* Bulk diagnostics of main_v1.9.f90:799-906 on the current var:
 * out[0] = KE, out[1..5] rho_b, out[6..10] U_b, out[11..15] massflow for the
 * planes [5,10,70,100,125]; a plane outside 2 < p < n1-1 gives 0. */
void orc_diag(const orc_t *o, double *out);

/* Run the reference's plane-statistics / wall-shear routines on demand. */
void orc_calc_tauwall(orc_t *o);
void orc_avg_spatial(orc_t *o);

#ifdef __cplusplus
}
#endif
#endif
