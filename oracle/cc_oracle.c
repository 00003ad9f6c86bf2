/*
 * cc_oracle.c -- CPU restatement (FP64) of the chanCompr time step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP kernels, never shipped
 * on the product path.  Every routine cites the reference lines it follows;
 * "main:" is v1.10/codes/main_v1.9.f90 (the halo-corrected copy of v1.9,
 * which is the canonical physics), "comdata:" is v1.9/codes/comdata.f90.
 *
 * The loop nests over nodes are independent per node and run on OpenMP
 * threads when built with -fopenmp (the results do not depend on the thread
 * count); the reductions (wall shear, plane statistics, diagnostics) stay
 * sequential in the reference's order.
 *
 * One MPI rank with topology 1x1x1 is restated: periodic x/y exchanges become
 * self-copies over the ranges the reference's derived datatypes cover
 * (main:297-313), the z exchange is a no-op (MPI_PROC_NULL), and the
 * reductions are the identity.
 */
#include "cc_oracle.h"

#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <stdlib.h>
#include <string.h>

struct orc {
  int n1, n2, n3;
  int NX, NY, NZ;          /* halo extents n+2 */
  double Re, invRe, Mac, dt, PexVal, PeyVal, PezVal, ggM;
  int istat;
  double Pi, dx, dy, dXi, invdx, invdy, invdXi;
  double *x1, *x2, *x3, *x3f, *Xi, *jac, *invJac;
  double x2end;  /* x2(n2) as read from the solution file */
  double dz1, dzN;
  double f1d2c0, f1d2c1, f1d2c2, b1d2c0, b1d2c1, b1d2c2;
  double f2d1c0, f2d1c1, f2d1c2, b2d1c0, b2d1c1, b2d1c2;
  double *solV;            /* (0:n1+1,0:n2+1,0:n3+1,5,2) */
  double *var;             /* (0:n1+1,0:n2+1,0:n3+1,7)   */
  double *fD, *gD, *hD;    /* (n1,n2,n3,5)                */
  double *F1, *G1, *H1, *F2, *G2, *H2, *R;
  double *mu, *kt;         /* (n1,n2,n3)                  */
  double tauL, tauU;
  double *stat1, *stat2, *corr; /* (n3,7), (n3,7), (n3,9) */
  int nstep;
  double time;
  int blown[3];
  /* physics the reference does not have (0 = reference): bulk Richardson
   * number of a buoyancy body force -Ri rho along z, and the Navier slip
   * length of both walls */
  double Ri, Ls;
  /* wall conditions of the parameter studies (orc_set_wall) */
  orc_wall wall[2];
  double *x1f, *x2f;  /* x1, x2 as read from the solution file */
};

static const double one = 1.0, two = 2.0, zer = 0.0, quar = 0.25, haf = 0.5;
static const double gama = 1.4;
#define TWO3RD (2.0 / 3.0)
#define ONE12TH (1.0 / 12.0)
#define INVPR (1.0 / 0.71) /* comdata:12 */

/* ---- index helpers (Fortran column-major) ---------------------------- */
/* halo arrays: i in 0..n1+1 etc. */
static inline size_t ih(const orc_t *o, int i, int j, int k) {
  return (size_t)i + (size_t)o->NX * ((size_t)j + (size_t)o->NY * (size_t)k);
}
static inline size_t hsz(const orc_t *o) {
  return (size_t)o->NX * o->NY * o->NZ;
}
#define SV(c, s, i, j, k) \
  o->solV[ih(o, i, j, k) + hsz(o) * ((size_t)(c) + 5 * (size_t)(s))]
#define VR(c, i, j, k) o->var[ih(o, i, j, k) + hsz(o) * (size_t)(c)]
/* interior arrays: i in 1..n1 */
static inline size_t in3(const orc_t *o, int i, int j, int k) {
  return (size_t)(i - 1) +
         (size_t)o->n1 * ((size_t)(j - 1) + (size_t)o->n2 * (size_t)(k - 1));
}
static inline size_t nsz(const orc_t *o) {
  return (size_t)o->n1 * o->n2 * o->n3;
}
#define A5(arr, c, i, j, k) (arr)[in3(o, i, j, k) + nsz(o) * (size_t)(c)]
#define MU(i, j, k) o->mu[in3(o, i, j, k)]
#define KT(i, j, k) o->kt[in3(o, i, j, k)]
/* components are 0-based here: c = 0..4 is Fortran 1..5, var c = 0..6 */

/* ---- setup: comdata:70-134 and main:52-76 ------------------------------ */
orc_t *orc_create(int n1, int n2, int n3, double Re, double Mac, double dt,
                  int istat, double Pex, double Pey, double Pez) {
  if (n1 < 6 || n2 < 6 || n3 < 6) return NULL;
  orc_t *o = (orc_t *)calloc(1, sizeof(orc_t));
  o->n1 = n1; o->n2 = n2; o->n3 = n3;
  o->NX = n1 + 2; o->NY = n2 + 2; o->NZ = n3 + 2;
  o->Re = Re; o->Mac = Mac; o->dt = dt; o->istat = istat;
  o->PexVal = Pex; o->PeyVal = Pey; o->PezVal = Pez;
  o->invRe = 1.0 / Re;                       /* comdata:59 */
  o->ggM = gama * (gama - 1) * (Mac * Mac);  /* comdata:60 */

  o->x1 = (double *)calloc(n1, sizeof(double));
  o->x2 = (double *)calloc(n2, sizeof(double));
  o->x3 = (double *)calloc(n3, sizeof(double));
  o->x3f = (double *)calloc(n3, sizeof(double));
  o->Xi = (double *)calloc(n3, sizeof(double));
  o->jac = (double *)calloc(n3, sizeof(double));
  o->invJac = (double *)calloc(n3, sizeof(double));

  o->Pi = 4.0 * atan(1.0);                   /* comdata:74 */
  o->x1[n1 - 1] = 4.0 * o->Pi;               /* comdata:81-83 */
  o->x2[n2 - 1] = 4.0 * o->Pi / 3.0;
  o->x3[n3 - 1] = 2.0;
  o->dx = o->x1[n1 - 1] / (double)(n1 - 1);  /* comdata:85-87 */
  o->dy = o->x2[n2 - 1] / (double)(n2 - 1);
  o->dXi = o->x3[n3 - 1] / (double)(n3 - 1);
  o->invdx = 1.0 / o->dx;
  o->invdy = 1.0 / o->dy;
  o->invdXi = 1.0 / o->dXi;
  for (int i = 0; i < n1 - 1; i++) o->x1[i + 1] = o->x1[i] + o->dx;
  for (int j = 0; j < n2 - 1; j++) o->x2[j + 1] = o->x2[j] + o->dy;
  for (int k = 0; k < n3 - 1; k++) o->Xi[k + 1] = o->Xi[k] + o->dXi;
  for (int k = 0; k < n3; k++) {             /* comdata:105-109 */
    o->x3[k] = 1.0 - cos(0.50 * o->Xi[k] * o->Pi);
    o->jac[k] = 0.50 * o->Pi * sin(0.50 * o->Xi[k] * o->Pi);
  }
  for (int k = 1; k < n3; k++) o->invJac[k] = 1.0 / o->jac[k]; /* :110 */
  memcpy(o->x3f, o->x3, n3 * sizeof(double));
  o->x2end = o->x2[n2 - 1];
  o->x1f = (double *)calloc(n1, sizeof(double));
  o->x2f = (double *)calloc(n2, sizeof(double));
  memcpy(o->x1f, o->x1, n1 * sizeof(double));
  memcpy(o->x2f, o->x2, n2 * sizeof(double));

  /* one-sided coefficients, main:59-83 */
  const double *x3 = o->x3;
  double dz1 = x3[1] - x3[0];
  double alpha = (x3[2] - x3[1]) / dz1;
  double t = dz1 * alpha * (alpha + one);
  o->f1d2c0 = (one - (alpha + one) * (alpha + one)) / t;
  o->f1d2c1 = ((alpha + one) * (alpha + one)) / t;
  o->f1d2c2 = -one / t;
  t = haf * (dz1 * dz1) * alpha * (alpha + one);
  o->f2d1c0 = alpha / t;
  o->f2d1c1 = -(alpha + one) / t;
  o->f2d1c2 = one / t;
  double dzN = x3[n3 - 1] - x3[n3 - 2];
  alpha = (x3[n3 - 2] - x3[n3 - 3]) / dzN;
  t = dzN * alpha * (alpha + one);
  o->b1d2c0 = ((alpha + one) * (alpha + one) - one) / t;
  o->b1d2c1 = -((alpha + one) * (alpha + one)) / t;
  o->b1d2c2 = one / t;
  t = haf * (dzN * dzN) * alpha * (alpha + one);
  o->b2d1c0 = alpha / t;
  o->b2d1c1 = -(alpha + one) / t;
  o->b2d1c2 = one / t;
  o->dz1 = dz1;
  o->dzN = dzN;

  size_t H = hsz(o), N = nsz(o);
  o->solV = (double *)calloc(H * 10, sizeof(double));
  o->var = (double *)calloc(H * 7, sizeof(double));
  o->fD = (double *)calloc(N * 5, sizeof(double));
  o->gD = (double *)calloc(N * 5, sizeof(double));
  o->hD = (double *)calloc(N * 5, sizeof(double));
  o->F1 = (double *)calloc(N * 5, sizeof(double));
  o->G1 = (double *)calloc(N * 5, sizeof(double));
  o->H1 = (double *)calloc(N * 5, sizeof(double));
  o->F2 = (double *)calloc(N * 5, sizeof(double));
  o->G2 = (double *)calloc(N * 5, sizeof(double));
  o->H2 = (double *)calloc(N * 5, sizeof(double));
  o->R = (double *)calloc(N * 5, sizeof(double));
  o->mu = (double *)calloc(N, sizeof(double));
  o->kt = (double *)calloc(N, sizeof(double));
  o->stat1 = (double *)calloc((size_t)n3 * 7, sizeof(double));
  o->stat2 = (double *)calloc((size_t)n3 * 7, sizeof(double));
  o->corr = (double *)calloc((size_t)n3 * 9, sizeof(double));
  if (!o->solV || !o->var || !o->fD || !o->gD || !o->hD || !o->F1 || !o->G1 ||
      !o->H1 || !o->F2 || !o->G2 || !o->H2 || !o->R || !o->mu || !o->kt) {
    orc_destroy(o);
    return NULL;
  }
  return o;
}

void orc_destroy(orc_t *o) {
  if (!o) return;
  free(o->x1); free(o->x2); free(o->x3); free(o->x3f); free(o->x1f); free(o->x2f); free(o->Xi);
  free(o->jac); free(o->invJac);
  free(o->solV); free(o->var); free(o->fD); free(o->gD); free(o->hD);
  free(o->F1); free(o->G1); free(o->H1); free(o->F2); free(o->G2);
  free(o->H2); free(o->R); free(o->mu); free(o->kt);
  free(o->stat1); free(o->stat2); free(o->corr);
  free(o);
}

void orc_set_x3_file(orc_t *o, const double *x3, double x2end) {
  memcpy(o->x3f, x3, (size_t)o->n3 * sizeof(double));
  o->x2end = x2end;
}

void orc_set_wall(orc_t *o, int wall, const orc_wall *w, const double *x1, const double *x2) {
  if (wall < 0 || wall > 1) return;
  if (w) o->wall[wall] = *w;
  else memset(&o->wall[wall], 0, sizeof(orc_wall));
  if (x1) memcpy(o->x1f, x1, o->n1 * sizeof(double));
  if (x2) memcpy(o->x2f, x2, o->n2 * sizeof(double));
}

void orc_set_ext(orc_t *o, double Ri, double Ls) {
  o->Ri = Ri;
  o->Ls = Ls;
}

void orc_set_forcing(orc_t *o, double tauL, double tauU, const double *rb) {
  o->tauL = tauL;
  o->tauU = tauU;
  for (int k = 0; k < o->n3; k++) o->stat1[k] = rb ? rb[k] : 0.0;
}

/* ---- halo fills at one rank: main:297-313, 589-596, 612-619 ------------ */
/* Two-plane periodic copy of ncomp components of a halo array, over the
 * ranges of the yz/xz datatypes: x for j in bn2..en2, k in bn3..en3; then y
 * for i in bn1..en1.  Edge/corner columns are never written. */
static void share2P(orc_t *o, double *a, int ncomp) {
  const int n1 = o->n1, n2 = o->n2, n3 = o->n3;
  size_t H = hsz(o);
  for (int c = 0; c < ncomp; c++) {
    double *b = a + H * c;
    for (int k = 2; k <= n3 - 1; k++)
      for (int j = 2; j <= n2 - 1; j++) {
        /* recv at bn1-2..bn1-1 = 0..1 from en1-1..en1 = n1-2..n1-1 */
        b[ih(o, 0, j, k)] = b[ih(o, n1 - 2, j, k)];
        b[ih(o, 1, j, k)] = b[ih(o, n1 - 1, j, k)];
        /* recv at en1+1..en1+2 = n1..n1+1 from bn1..bn1+1 = 2..3 */
        b[ih(o, n1, j, k)] = b[ih(o, 2, j, k)];
        b[ih(o, n1 + 1, j, k)] = b[ih(o, 3, j, k)];
      }
    for (int k = 2; k <= n3 - 1; k++)
      for (int i = 2; i <= n1 - 1; i++) {
        b[ih(o, i, 0, k)] = b[ih(o, i, n2 - 2, k)];
        b[ih(o, i, 1, k)] = b[ih(o, i, n2 - 1, k)];
        b[ih(o, i, n2, k)] = b[ih(o, i, 2, k)];
        b[ih(o, i, n2 + 1, k)] = b[ih(o, i, 3, k)];
      }
    /* z: MPI_PROC_NULL on a non-periodic single rank -> no-op */
  }
}

/* One-plane copy of the viscous point fluxes, main:543-548:
 * fD x-halo (i=1 <- n1-1, i=n1 <- 2), gD y-halo; hD z is a no-op. */
static void share1P_visc(orc_t *o) {
  const int n1 = o->n1, n2 = o->n2, n3 = o->n3;
  for (int c = 0; c < 5; c++) {
    for (int k = 2; k <= n3 - 1; k++)
      for (int j = 2; j <= n2 - 1; j++) {
        A5(o->fD, c, 1, j, k) = A5(o->fD, c, n1 - 1, j, k);
        A5(o->fD, c, n1, j, k) = A5(o->fD, c, 2, j, k);
      }
    for (int k = 2; k <= n3 - 1; k++)
      for (int i = 2; i <= n1 - 1; i++) {
        A5(o->gD, c, i, 1, k) = A5(o->gD, c, i, n2 - 1, k);
        A5(o->gD, c, i, n2, k) = A5(o->gD, c, i, 2, k);
      }
  }
}

/* ---- initial solution: main:284-313 ------------------------------------ */
int orc_load(orc_t *o, const double *v, int nstep, double time) {
  const int n1 = o->n1, n2 = o->n2, n3 = o->n3;
  size_t N = nsz(o);
  for (int c = 0; c < 7; c++)
    for (int k = 1; k <= n3; k++)
      for (int j = 1; j <= n2; j++)
        for (int i = 1; i <= n1; i++)
          VR(c, i, j, k) = v[in3(o, i, j, k) + N * c];
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++) {
        SV(0, 0, i, j, k) = VR(0, i, j, k);
        for (int c = 1; c < 5; c++) SV(c, 0, i, j, k) = VR(0, i, j, k) * VR(c, i, j, k);
        MU(i, j, k) = pow(VR(5, i, j, k), 0.7);
        KT(i, j, k) = pow(VR(5, i, j, k), 0.7);
      }
  share2P(o, o->solV, 5);
  share2P(o, o->var, 7);
  o->nstep = nstep;
  o->time = time;
  return 0;
}

/* ---- convective flux: main:329-454 -------------------------------------- */
/* QUICK face pair of main:339-361 for one component: upwind-biased face
 * values at e (between c and p) and w (between m and c).  With a NaN face
 * velocity none of the reference's four cases applies and its F keeps the
 * value of the previous evaluation (F is a persistent array); this returns
 * 0 there.  The two agree on every state without NaN; how NaN spreads after
 * its first appearance is parity unpinned. */
static inline double quick(double ue, double uw, double qmm, double qm,
                           double qc, double qp, double qpp) {
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

static void calcFluxConv(orc_t *o) {
  const int n1 = o->n1, n2 = o->n2, n3 = o->n3;
  #pragma omp parallel for schedule(static)
  for (int k = 2; k <= n3 - 1; k++)
    for (int j = 2; j <= n2 - 1; j++)
      for (int i = 2; i <= n1 - 1; i++) {
        /* x (main:333-369) */
        double Pe = o->Re * fabs(SV(1, 0, i, j, k)) * o->dx / MU(i, j, k);
        if (Pe > o->PexVal) {
          double t1 = (VR(1, i, j, k) + VR(1, i + 1, j, k)) * haf;
          double t2 = (VR(1, i, j, k) + VR(1, i - 1, j, k)) * haf;
          for (int c = 0; c < 5; c++)
            A5(o->F1, c, i, j, k) =
                quick(t1, t2, SV(c, 0, i - 2, j, k), SV(c, 0, i - 1, j, k),
                      SV(c, 0, i, j, k), SV(c, 0, i + 1, j, k),
                      SV(c, 0, i + 2, j, k)) * o->invdx / 8.0;
        } else {
          for (int c = 0; c < 5; c++)
            A5(o->F1, c, i, j, k) =
                (-VR(1, i + 2, j, k) * SV(c, 0, i + 2, j, k) +
                 8.0 * VR(1, i + 1, j, k) * SV(c, 0, i + 1, j, k) -
                 8.0 * VR(1, i - 1, j, k) * SV(c, 0, i - 1, j, k) +
                 VR(1, i - 2, j, k) * SV(c, 0, i - 2, j, k)) * o->invdx * ONE12TH;
        }
        /* y (main:371-407) */
        Pe = o->Re * fabs(SV(2, 0, i, j, k)) * o->dy / MU(i, j, k);
        if (Pe > o->PeyVal) {
          double t1 = (VR(2, i, j, k) + VR(2, i, j + 1, k)) * haf;
          double t2 = (VR(2, i, j, k) + VR(2, i, j - 1, k)) * haf;
          for (int c = 0; c < 5; c++)
            A5(o->G1, c, i, j, k) =
                quick(t1, t2, SV(c, 0, i, j - 2, k), SV(c, 0, i, j - 1, k),
                      SV(c, 0, i, j, k), SV(c, 0, i, j + 1, k),
                      SV(c, 0, i, j + 2, k)) * o->invdy / 8.0;
        } else {
          for (int c = 0; c < 5; c++)
            A5(o->G1, c, i, j, k) =
                (-VR(2, i, j + 2, k) * SV(c, 0, i, j + 2, k) +
                 8.0 * VR(2, i, j + 1, k) * SV(c, 0, i, j + 1, k) -
                 8.0 * VR(2, i, j - 1, k) * SV(c, 0, i, j - 1, k) +
                 VR(2, i, j - 2, k) * SV(c, 0, i, j - 2, k)) * o->invdy * ONE12TH;
        }
        /* z (main:409-450) */
        Pe = o->Re * fabs(SV(3, 0, i, j, k)) * o->dXi * o->jac[k - 1] / MU(i, j, k);
        const double iJ = o->invJac[k - 1];
        if (k < 3 || k > n3 - 2) {
          for (int c = 0; c < 5; c++)
            A5(o->H1, c, i, j, k) =
                (VR(3, i, j, k + 1) * SV(c, 0, i, j, k + 1) -
                 VR(3, i, j, k - 1) * SV(c, 0, i, j, k - 1)) * o->invdXi * iJ * haf;
        } else if (Pe > o->PezVal) {
          double t1 = (VR(3, i, j, k) + VR(3, i, j, k + 1)) * haf;
          double t2 = (VR(3, i, j, k) + VR(3, i, j, k - 1)) * haf;
          for (int c = 0; c < 5; c++)
            A5(o->H1, c, i, j, k) =
                quick(t1, t2, SV(c, 0, i, j, k - 2), SV(c, 0, i, j, k - 1),
                      SV(c, 0, i, j, k), SV(c, 0, i, j, k + 1),
                      SV(c, 0, i, j, k + 2)) * o->invdXi * iJ / 8.0;
        } else {
          for (int c = 0; c < 5; c++)
            A5(o->H1, c, i, j, k) =
                (-VR(3, i, j, k + 2) * SV(c, 0, i, j, k + 2) +
                 8.0 * VR(3, i, j, k + 1) * SV(c, 0, i, j, k + 1) -
                 8.0 * VR(3, i, j, k - 1) * SV(c, 0, i, j, k - 1) +
                 VR(3, i, j, k - 2) * SV(c, 0, i, j, k - 2)) * o->invdXi * iJ * ONE12TH;
        }
      }
}

/* ---- viscous/pressure point fluxes: main:457-538 ------------------------ */
static void calcFluxVisc(orc_t *o) {
  const int n1 = o->n1, n2 = o->n2, n3 = o->n3;
  const double invRe = o->invRe, ggM = o->ggM;
  const double invdx = o->invdx, invdy = o->invdy, invdXi = o->invdXi;
  #pragma omp parallel for schedule(static)
  for (int k = 1; k <= n3; k++)
    for (int j = 2; j <= n2 - 1; j++)
      for (int i = 2; i <= n1 - 1; i++) {
        double Ut = VR(1, i, j, k), Vt = VR(2, i, j, k), Wt = VR(3, i, j, k);
        double dudx = (VR(1, i + 1, j, k) - VR(1, i - 1, j, k)) * haf * invdx;
        double dudy = (VR(1, i, j + 1, k) - VR(1, i, j - 1, k)) * haf * invdy;
        double dvdx = (VR(2, i + 1, j, k) - VR(2, i - 1, j, k)) * haf * invdx;
        double dvdy = (VR(2, i, j + 1, k) - VR(2, i, j - 1, k)) * haf * invdy;
        double dwdx = (VR(3, i + 1, j, k) - VR(3, i - 1, j, k)) * haf * invdx;
        double dwdy = (VR(3, i, j + 1, k) - VR(3, i, j - 1, k)) * haf * invdy;
        double dTdx = (VR(5, i + 1, j, k) - VR(5, i - 1, j, k)) * haf * invdx;
        double dTdy = (VR(5, i, j + 1, k) - VR(5, i, j - 1, k)) * haf * invdy;
        double dudz, dvdz, dwdz, dTdz;
        if (k == 1) {
          dudz = (VR(1, i, j, 2) - VR(1, i, j, 1)) / o->dz1;
          dvdz = (VR(2, i, j, 2) - VR(2, i, j, 1)) / o->dz1;
          dwdz = (VR(3, i, j, 2) - VR(3, i, j, 1)) / o->dz1;
          dTdz = (VR(5, i, j, 2) - VR(5, i, j, 1)) / o->dz1;
        } else if (k == n3) {
          dudz = (VR(1, i, j, n3) - VR(1, i, j, n3 - 1)) / o->dzN;
          dvdz = (VR(2, i, j, n3) - VR(2, i, j, n3 - 1)) / o->dzN;
          dwdz = (VR(3, i, j, n3) - VR(3, i, j, n3 - 1)) / o->dzN;
          dTdz = (VR(5, i, j, n3) - VR(5, i, j, n3 - 1)) / o->dzN;
        } else {
          const double iJ = o->invJac[k - 1];
          dudz = (VR(1, i, j, k + 1) - VR(1, i, j, k - 1)) * haf * invdXi * iJ;
          dvdz = (VR(2, i, j, k + 1) - VR(2, i, j, k - 1)) * haf * invdXi * iJ;
          dwdz = (VR(3, i, j, k + 1) - VR(3, i, j, k - 1)) * haf * invdXi * iJ;
          dTdz = (VR(5, i, j, k + 1) - VR(5, i, j, k - 1)) * haf * invdXi * iJ;
        }
        double divg = dudx + dvdy + dwdz;
        double muT = MU(i, j, k), ktT = KT(i, j, k), PressT = VR(6, i, j, k);
        double dissF = ggM * invRe *
                       (TWO3RD * muT * Ut * divg - two * muT * Ut * dudx -
                        muT * Vt * (dvdx + dudy) - muT * Wt * (dwdx + dudz));
        double dissG = ggM * invRe *
                       (-muT * Ut * (dvdx + dudy) + TWO3RD * muT * Vt * divg -
                        two * muT * Vt * dvdy - muT * Wt * (dwdy + dvdz));
        double dissH = ggM * invRe *
                       (-muT * Ut * (dwdx + dudz) - muT * Vt * (dwdy + dvdz) +
                        TWO3RD * muT * Wt * divg - two * muT * Wt * dwdz);
        A5(o->fD, 0, i, j, k) = zer;
        A5(o->fD, 1, i, j, k) = PressT + TWO3RD * muT * invRe * divg - two * muT * invRe * dudx;
        A5(o->fD, 2, i, j, k) = -muT * invRe * (dvdx + dudy);
        A5(o->fD, 3, i, j, k) = -muT * invRe * (dwdx + dudz);
        A5(o->fD, 4, i, j, k) = -gama * ktT * invRe * INVPR * dTdx + ggM * Ut * PressT + dissF;
        A5(o->gD, 0, i, j, k) = zer;
        A5(o->gD, 1, i, j, k) = -muT * invRe * (dvdx + dudy);
        A5(o->gD, 2, i, j, k) = PressT + TWO3RD * muT * invRe * divg - two * muT * invRe * dvdy;
        A5(o->gD, 3, i, j, k) = -muT * invRe * (dwdy + dvdz);
        A5(o->gD, 4, i, j, k) = -gama * ktT * invRe * INVPR * dTdy + ggM * Vt * PressT + dissG;
        A5(o->hD, 0, i, j, k) = zer;
        A5(o->hD, 1, i, j, k) = -muT * invRe * (dwdx + dudz);
        A5(o->hD, 2, i, j, k) = -muT * invRe * (dwdy + dvdz);
        A5(o->hD, 3, i, j, k) = PressT + TWO3RD * muT * invRe * divg - two * muT * invRe * dwdz;
        A5(o->hD, 4, i, j, k) = -gama * ktT * invRe * INVPR * dTdz + ggM * Wt * PressT + dissH;
      }
}

/* ---- divergence, source, residual, RK2: main:551-586 ------------------- */
static void calcSolV(orc_t *o, int iter) {
  const int n1 = o->n1, n2 = o->n2, n3 = o->n3;
  #pragma omp parallel for schedule(static)
  for (int k = 2; k <= n3 - 1; k++)
    for (int j = 2; j <= n2 - 1; j++)
      for (int i = 2; i <= n1 - 1; i++)
        for (int c = 0; c < 5; c++) {
          A5(o->F2, c, i, j, k) = (A5(o->fD, c, i + 1, j, k) - A5(o->fD, c, i - 1, j, k)) * haf * o->invdx;
          A5(o->G2, c, i, j, k) = (A5(o->gD, c, i, j + 1, k) - A5(o->gD, c, i, j - 1, k)) * haf * o->invdy;
          A5(o->H2, c, i, j, k) = (A5(o->hD, c, i, j, k + 1) - A5(o->hD, c, i, j, k - 1)) * haf * o->invdXi * o->invJac[k - 1];
        }
  const double tdiff = o->tauL - o->tauU;
  #pragma omp parallel for schedule(static)
  for (int k = 2; k <= n3 - 1; k++)
    for (int j = 2; j <= n2 - 1; j++)
      for (int i = 2; i <= n1 - 1; i++) {
        double src[5];
        src[0] = zer;
        src[1] = haf * tdiff * o->stat1[k - 1];
        src[2] = zer;
        src[3] = zer;
        src[4] = o->ggM * VR(1, i, j, k) * haf * tdiff * o->stat1[k - 1];
        if (o->Ri != 0.0) { /* buoyancy: -Ri rho along z and its work */
          src[3] = -o->Ri * SV(0, 0, i, j, k);
          src[4] = src[4] - o->ggM * o->Ri * SV(3, 0, i, j, k);
        }
        for (int c = 0; c < 5; c++) {
          double r = src[c] - (A5(o->F1, c, i, j, k) + A5(o->F2, c, i, j, k) +
                               A5(o->G1, c, i, j, k) + A5(o->G2, c, i, j, k) +
                               A5(o->H1, c, i, j, k) + A5(o->H2, c, i, j, k));
          A5(o->R, c, i, j, k) = r;
          if (iter == 1) {
            SV(c, 1, i, j, k) = SV(c, 0, i, j, k);
            SV(c, 0, i, j, k) = SV(c, 0, i, j, k) + haf * o->dt * r;
          } else {
            SV(c, 0, i, j, k) = SV(c, 1, i, j, k) + o->dt * r;
          }
        }
      }
}

/* ---- primitive recovery: main:599-609 ----------------------------------- */
static void calcPrim(orc_t *o) {
  const int n1 = o->n1, n2 = o->n2, n3 = o->n3;
  #pragma omp parallel for schedule(static)
  for (int k = 2; k <= n3 - 1; k++)
    for (int j = 2; j <= n2 - 1; j++)
      for (int i = 2; i <= n1 - 1; i++) {
        double rho = SV(0, 0, i, j, k);
        VR(0, i, j, k) = rho;
        for (int c = 1; c < 5; c++) VR(c, i, j, k) = SV(c, 0, i, j, k) / rho;
        double u = VR(1, i, j, k), v = VR(2, i, j, k), w = VR(3, i, j, k);
        VR(5, i, j, k) = VR(4, i, j, k) - o->ggM * haf * (u * u + v * v + w * w);
        VR(6, i, j, k) = VR(0, i, j, k) * VR(5, i, j, k) / (gama * (o->Mac * o->Mac));
      }
}

/* Wall-plane halo ring from its periodic images (slip walls only: with
 * u_w != 0 a stale ring would put a jump at the periodic seam). */
static void wall_ring_images(orc_t *o, int c0, int c1) {
  const int n1 = o->n1, n2 = o->n2, n3 = o->n3;
  for (int kk = 0; kk < 2; kk++) {
    const int k = kk ? n3 : 1;
    for (int c = c0; c <= c1; c++) {
      for (int j = 2; j <= n2 - 1; j++) {
        VR(c, 1, j, k) = VR(c, n1 - 1, j, k);
        VR(c, n1, j, k) = VR(c, 2, j, k);
      }
      for (int i = 1; i <= n1; i++) {
        VR(c, i, 1, k) = VR(c, i, n2 - 1, k);
        VR(c, i, n2, k) = VR(c, i, 2, k);
      }
    }
  }
}

/* Polar angle in [0, 2pi) as the rotating-disc studies compute it
 * (v1.9_RD/v1.9_rd1/codes/comdata.f90:180-193).  That function leaves its
 * result unset when either argument is zero; the restatement takes the
 * angle of the axis there (0, pi/2, pi, 3pi/2; 0 at the centre, where the
 * radius multiplying it is 0). */
static double myatand(double A1, double A2) {
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

/* Stripe predicate of the groove studies at node (i,j) (1-based). */
static int on_stripe(const orc_t *o, const orc_wall *w, int i, int j) {
  const int n1 = o->n1, n2 = o->n2;
  const double x = o->x1f[i - 1], y = o->x2f[j - 1];
  if (w->stripes == 0) /* main_v1.9_g1.f90:609 */
    return ((6 * j / n2) % 2) == 0;
  if (w->stripes == 1) { /* main_v1.9_gst3.f90: diagonal bands */
    const double r = (o->x1f[n1 - 1] + o->x2f[n2 - 1]) / 8.0;
    return ((int)floor((x - y + o->x2f[n2 - 1]) / r) % 2) == 0;
  }
  { /* main_v1.9_gst5.f90: wavy bands, every third */
    const double r = o->x2f[n2 - 1] / 8.0;
    return (((int)floor((y - r * sin(x)) / r) + 1) % 3) == 0;
  }
}

/* u, v, T of wall node (i,j) under the wall condition of wall kk, from the
 * current primitives of the two planes above it (the BC blocks of the
 * parameter studies; see orc_set_wall). */
static void wall_mode_values(const orc_t *o, int kk, int i, int j, double *u, double *v,
                             double *T) {
  const orc_wall *w = &o->wall[kk];
  const int n3 = o->n3;
  const int ka = kk ? n3 - 1 : 2, kb = kk ? n3 - 2 : 3;
  const double c0 = kk ? o->b1d2c0 : o->f1d2c0, c1 = kk ? o->b1d2c1 : o->f1d2c1,
               c2 = kk ? o->b1d2c2 : o->f1d2c2;
  const double x = o->x1f[i - 1], y = o->x2f[j - 1];
  *u = 0.0;
  *v = 0.0;
  if (w->vel == ORC_WALL_WAVE) {
    const double ph = (j % 2 == 1) ? w->phase_odd : 0.0;
    *v = w->amp * sin(w->kx * x + w->omega * o->time + ph);
  } else if (w->vel == ORC_WALL_DISCS) {
    const double R = o->x2f[o->n2 - 1] / 10.0;
    const double cy[3] = {2.0, 5.0, 8.0};
    for (int d = 0; d < 3; d++) {
      const double ax = x - two * R, ay = y - cy[d] * R;
      const double r = sqrt(ax * ax + ay * ay);
      if (r < R) {
        const double th = myatand(ax, ay);
        *u = -(r * w->disc_omega * sin(th));
        *v = r * w->disc_omega * cos(th);
        break;
      }
    }
  } else if (w->vel == ORC_WALL_STRIPES) {
    if (on_stripe(o, w, i, j)) {
      *u = (w->grad - c1 * VR(1, i, j, ka) - c2 * VR(1, i, j, kb)) / c0;
      *v = (w->grad - c1 * VR(2, i, j, ka) - c2 * VR(2, i, j, kb)) / c0;
    }
  }
  if (w->temp == ORC_WALLT_FIXED)
    *T = w->Tw;
  else if (w->temp == ORC_WALLT_GRAD)
    *T = (w->dTdz - c1 * VR(5, i, j, ka) - c2 * VR(5, i, j, kb)) / c0;
  else
    *T = 1.0;
}

static int wall_active(const orc_t *o, int kk) {
  return o->wall[kk].vel != ORC_WALL_NOSLIP || o->wall[kk].temp != ORC_WALLT_ONE;
}

/* ---- wall values, mu, wall pressure, kt: main:622-720 ------------------- */
static void calcBC(orc_t *o) {
  const int n1 = o->n1, n2 = o->n2, n3 = o->n3;
  const double gM2 = gama * (o->Mac * o->Mac);
  for (int j = 1; j <= n2; j++)
    for (int i = 1; i <= n1; i++) {
      for (int kk = 0; kk < 2; kk++) {
        int k = kk ? n3 : 1;
        VR(1, i, j, k) = 0.0; VR(2, i, j, k) = 0.0; VR(3, i, j, k) = 0.0;
        const int upd = i > 1 && i < n1 && j > 1 && j < n2;
        if (wall_active(o, kk)) {
          double u, v, T;
          wall_mode_values(o, kk, i, j, &u, &v, &T);
          VR(1, i, j, k) = u;
          VR(2, i, j, k) = v;
          VR(5, i, j, k) = T;
          VR(4, i, j, k) = VR(5, i, j, k) + o->ggM * haf * (u * u + v * v + 0.0 * 0.0);
          VR(0, i, j, k) = gM2 * VR(6, i, j, k) / VR(5, i, j, k);
          continue;
        }
        if (o->Ls != 0.0 && upd) {
          /* Navier slip u_w = Ls du/dn with the one-sided coefficients;
           * Ls -> inf is the zero-gradient wall of main_v1.9_g1.f90:611 */
          for (int c = 1; c <= 2; c++) {
            if (!kk)
              VR(c, i, j, 1) = o->Ls * (o->f1d2c1 * VR(c, i, j, 2) + o->f1d2c2 * VR(c, i, j, 3)) /
                               (1.0 - o->Ls * o->f1d2c0);
            else
              VR(c, i, j, n3) = -o->Ls * (o->b1d2c1 * VR(c, i, j, n3 - 1) + o->b1d2c2 * VR(c, i, j, n3 - 2)) /
                                (1.0 + o->Ls * o->b1d2c0);
          }
        }
        VR(5, i, j, k) = 1.0;
        double u = VR(1, i, j, k), v = VR(2, i, j, k), w = VR(3, i, j, k);
        VR(4, i, j, k) = VR(5, i, j, k) + o->ggM * haf * (u * u + v * v + w * w);
        VR(0, i, j, k) = gM2 * VR(6, i, j, k) / VR(5, i, j, k);
      }
    }
  if (o->Ls != 0.0) wall_ring_images(o, 0, 6);
  /* mu over bn-1..en+1 = 1..n (x, y) and all k (main:643) */
  #pragma omp parallel for schedule(static)
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++) MU(i, j, k) = pow(VR(5, i, j, k), 0.7);
  const double invRe = o->invRe;
  for (int j = 2; j <= n2 - 1; j++)
    for (int i = 2; i <= n1 - 1; i++) {
      /* bottom, main:649-681 */
      {
        double dudxdz = ((o->f1d2c0 * MU(i + 1, j, 1) * VR(1, i + 1, j, 1) +
                          o->f1d2c1 * MU(i + 1, j, 2) * VR(1, i + 1, j, 2) +
                          o->f1d2c2 * MU(i + 1, j, 3) * VR(1, i + 1, j, 3)) -
                         (o->f1d2c0 * MU(i - 1, j, 1) * VR(1, i - 1, j, 1) +
                          o->f1d2c1 * MU(i - 1, j, 2) * VR(1, i - 1, j, 2) +
                          o->f1d2c2 * MU(i - 1, j, 3) * VR(1, i - 1, j, 3))) *
                        haf * o->invdx;
        double dvdydz = ((o->f1d2c0 * MU(i, j + 1, 1) * VR(2, i, j + 1, 1) +
                          o->f1d2c1 * MU(i, j + 1, 2) * VR(2, i, j + 1, 2) +
                          o->f1d2c2 * MU(i, j + 1, 3) * VR(2, i, j + 1, 3)) -
                         (o->f1d2c0 * MU(i, j - 1, 1) * VR(2, i, j - 1, 1) +
                          o->f1d2c1 * MU(i, j - 1, 2) * VR(2, i, j - 1, 2) +
                          o->f1d2c2 * MU(i, j - 1, 3) * VR(2, i, j - 1, 3))) *
                        haf * o->invdy;
        double d2wdz2 = o->f2d1c0 * MU(i, j, 1) * VR(3, i, j, 1) +
                        o->f2d1c1 * MU(i, j, 2) * VR(3, i, j, 2) +
                        o->f2d1c2 * MU(i, j, 3) * VR(3, i, j, 3);
        double w1 = VR(3, i, j, 1), w2 = VR(3, i, j, 2), w3 = VR(3, i, j, 3);
        double drhow2dz = o->f1d2c0 * VR(0, i, j, 1) * (w1 * w1) +
                          o->f1d2c1 * VR(0, i, j, 2) * (w2 * w2) +
                          o->f1d2c2 * VR(0, i, j, 3) * (w3 * w3);
        double lhs = invRe * (-dudxdz - dvdydz + TWO3RD * (dudxdz + dvdydz - two * d2wdz2)) + drhow2dz;
        if (o->Ri != 0.0) lhs = lhs + o->Ri * VR(0, i, j, 1);
        VR(6, i, j, 1) = -(lhs + o->f1d2c1 * VR(6, i, j, 2) + o->f1d2c2 * VR(6, i, j, 3)) / o->f1d2c0;
      }
      /* top, main:683-716 */
      {
        const int N = n3;
        double dudxdz = ((o->b1d2c0 * MU(i + 1, j, N) * VR(1, i + 1, j, N) +
                          o->b1d2c1 * MU(i + 1, j, N - 1) * VR(1, i + 1, j, N - 1) +
                          o->b1d2c2 * MU(i + 1, j, N - 2) * VR(1, i + 1, j, N - 2)) -
                         (o->b1d2c0 * MU(i - 1, j, N) * VR(1, i - 1, j, N) +
                          o->b1d2c1 * MU(i - 1, j, N - 1) * VR(1, i - 1, j, N - 1) +
                          o->b1d2c2 * MU(i - 1, j, N - 2) * VR(1, i - 1, j, N - 2))) *
                        haf * o->invdx;
        double dvdydz = ((o->b1d2c0 * MU(i, j + 1, N) * VR(2, i, j + 1, N) +
                          o->b1d2c1 * MU(i, j + 1, N - 1) * VR(2, i, j + 1, N - 1) +
                          o->b1d2c2 * MU(i, j + 1, N - 2) * VR(2, i, j + 1, N - 2)) -
                         (o->b1d2c0 * MU(i, j - 1, N) * VR(2, i, j - 1, N) +
                          o->b1d2c1 * MU(i, j - 1, N - 1) * VR(2, i, j - 1, N - 1) +
                          o->b1d2c2 * MU(i, j - 1, N - 2) * VR(2, i, j - 1, N - 2))) *
                        haf * o->invdy;
        double d2wdz2 = o->b2d1c0 * MU(i, j, N) * VR(3, i, j, N) +
                        o->b2d1c1 * MU(i, j, N - 1) * VR(3, i, j, N - 1) +
                        o->b2d1c2 * MU(i, j, N - 2) * VR(3, i, j, N - 2);
        double w1 = VR(3, i, j, N), w2 = VR(3, i, j, N - 1), w3 = VR(3, i, j, N - 2);
        double drhow2dz = o->b1d2c0 * VR(0, i, j, N) * (w1 * w1) +
                          o->b1d2c1 * VR(0, i, j, N - 1) * (w2 * w2) +
                          o->b1d2c2 * VR(0, i, j, N - 2) * (w3 * w3);
        double lhs = invRe * (-dudxdz - dvdydz + TWO3RD * (dudxdz + dvdydz - two * d2wdz2)) + drhow2dz;
        if (o->Ri != 0.0) lhs = lhs + o->Ri * VR(0, i, j, N);
        VR(6, i, j, N) = -(lhs + o->b1d2c1 * VR(6, i, j, N - 1) + o->b1d2c2 * VR(6, i, j, N - 2)) / o->b1d2c0;
      }
    }
  if (o->Ls != 0.0) wall_ring_images(o, 6, 6);
  #pragma omp parallel for schedule(static)
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++) KT(i, j, k) = pow(VR(5, i, j, k), 0.7);
}

void orc_stage(orc_t *o, int iter) {
  calcFluxConv(o);
  calcFluxVisc(o);
  share1P_visc(o);
  calcSolV(o, iter);
  share2P(o, o->solV, 5);
  calcPrim(o);
  share2P(o, o->var, 7);
  calcBC(o);
}

/* ---- wall shear: main:929-988 ------------------------------------------ */
void orc_calc_tauwall(orc_t *o) {
  const int n1 = o->n1, n2 = o->n2, n3 = o->n3;
  double s = 0.0;
  for (int j = 2; j <= n2 - 1; j++)
    for (int i = 2; i <= n1 - 1; i++) {
      double dwdx = MU(i, j, 1) * (VR(3, i + 1, j, 1) - VR(3, i - 1, j, 1)) * haf * o->invdx;
      double dudz = MU(i, j, 1) * (o->f1d2c0 * VR(1, i, j, 1) + o->f1d2c1 * VR(1, i, j, 2) +
                                   o->f1d2c2 * VR(1, i, j, 3));
      s = s + (dwdx + dudz) * o->invRe;
    }
  o->tauL = s / (double)((n1 - 2) * (n2 - 2));
  s = 0.0;
  for (int j = 2; j <= n2 - 1; j++)
    for (int i = 2; i <= n1 - 1; i++) {
      double dwdx = MU(i, j, n3) * (VR(3, i + 1, j, n3) - VR(3, i - 1, j, n3)) * haf * o->invdx;
      double dudz = MU(i, j, n3) * (o->b1d2c0 * VR(1, i, j, n3) + o->b1d2c1 * VR(1, i, j, n3 - 1) +
                                    o->b1d2c2 * VR(1, i, j, n3 - 2));
      s = s + (dwdx + dudz) * o->invRe;
    }
  o->tauU = s / (double)((n1 - 2) * (n2 - 2));
}

/* ---- plane statistics: main:993-1025 ------------------------------------ */
void orc_avg_spatial(orc_t *o) {
  const int n1 = o->n1, n2 = o->n2, n3 = o->n3;
  const double nXY = (double)(n1 * n2);
  static const int cp[9][2] = {{0, 3}, {0, 5}, {1, 2}, {1, 3}, {2, 3},
                               {2, 5}, {5, 3}, {6, 3}, {6, 5}};
  for (int k = 1; k <= n3; k++) {
    for (int p = 0; p < 7; p++) {
      double s1 = 0.0, s2 = 0.0;
      for (int j = 1; j <= n2; j++)
        for (int i = 1; i <= n1; i++) {
          double x = VR(p, i, j, k);
          s1 = s1 + x;
          s2 = s2 + x * x;
        }
      o->stat1[(k - 1) + (size_t)n3 * p] = s1 / nXY;
      o->stat2[(k - 1) + (size_t)n3 * p] = s2 / nXY;
    }
    for (int q = 0; q < 9; q++) {
      double s = 0.0;
      for (int j = 1; j <= n2; j++)
        for (int i = 1; i <= n1; i++)
          s = s + VR(cp[q][0], i, j, k) * VR(cp[q][1], i, j, k);
      o->corr[(k - 1) + (size_t)n3 * q] = s / nXY;
    }
  }
}

/* ---- blow-up check: main:728-766 ---------------------------------------- */
static int checkSoln(orc_t *o) {
  const int n1 = o->n1, n2 = o->n2, n3 = o->n3;
  int found = 0;
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++)
        if (VR(0, i, j, k) < 0.0 || VR(5, i, j, k) < 0.0) {
          if (!found) { o->blown[0] = i; o->blown[1] = j; o->blown[2] = k; }
          found = 1;
        }
  return found;
}

/* ---- time loop body: main:320-1051 -------------------------------------- */
int orc_step(orc_t *o, int nsteps) {
  for (int s = 0; s < nsteps; s++) {
    o->nstep = o->nstep + 1;
    o->time = o->time + o->dt;
    orc_stage(o, 1);
    orc_stage(o, 2);
    if (checkSoln(o)) return 1;
    if (o->nstep % o->istat == 0) {
      orc_calc_tauwall(o);
      orc_avg_spatial(o);
    }
  }
  return 0;
}

/* ---- bulk diagnostics: main:799-906 ------------------------------------- */
void orc_diag(const orc_t *o, double *out) {
  const int n1 = o->n1, n2 = o->n2, n3 = o->n3;
  static const int plane[5] = {5, 10, 70, 100, 125};
  double s = 0.0;
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++) {
        double u = VR(1, i, j, k), v = VR(2, i, j, k), w = VR(3, i, j, k);
        s = s + (u * u + v * v + w * w);
      }
  out[0] = s / (double)(n1 * n2 * n3);
  const double *x3 = o->x3f;
  double Ayz = x3[n3 - 1] * o->x2end;
  for (int q = 0; q < 3; q++)
    for (int m = 0; m < 5; m++) {
      int p = plane[m];
      double t = 0.0;
      if (p > 2 && p < n1 - 1) {
        for (int k = 1; k <= n3 - 1; k++)
          for (int j = 1; j <= n2 - 1; j++) {
            double dAyz = o->dy * (x3[k] - x3[k - 1]);
            double a, b, c, d;
            if (q == 0) {
              a = VR(0, p, j, k); b = VR(0, p, j + 1, k);
              c = VR(0, p, j, k + 1); d = VR(0, p, j + 1, k + 1);
            } else if (q == 1) {
              a = VR(1, p, j, k); b = VR(1, p, j + 1, k);
              c = VR(1, p, j, k + 1); d = VR(1, p, j + 1, k + 1);
            } else {
              a = VR(0, p, j, k) * VR(1, p, j, k);
              b = VR(0, p, j + 1, k) * VR(1, p, j + 1, k);
              c = VR(0, p, j, k + 1) * VR(1, p, j, k + 1);
              d = VR(0, p, j + 1, k + 1) * VR(1, p, j + 1, k + 1);
            }
            t = t + quar * (a + b + c + d) * dAyz;
          }
      }
      out[1 + 5 * q + m] = t / Ayz;
    }
}

/* ---- accessors ----------------------------------------------------------- */
void orc_get_var(const orc_t *o, double *v) {
  size_t N = nsz(o);
  for (int c = 0; c < 7; c++)
    for (int k = 1; k <= o->n3; k++)
      for (int j = 1; j <= o->n2; j++)
        for (int i = 1; i <= o->n1; i++)
          v[in3(o, i, j, k) + N * c] = VR(c, i, j, k);
}
void orc_get_solv(const orc_t *o, int level, double *U) {
  size_t N = nsz(o);
  for (int c = 0; c < 5; c++)
    for (int k = 1; k <= o->n3; k++)
      for (int j = 1; j <= o->n2; j++)
        for (int i = 1; i <= o->n1; i++)
          U[in3(o, i, j, k) + N * c] = SV(c, level, i, j, k);
}
void orc_get_mu(const orc_t *o, double *m) { memcpy(m, o->mu, nsz(o) * sizeof(double)); }
int orc_get_nstep(const orc_t *o) { return o->nstep; }
double orc_get_time(const orc_t *o) { return o->time; }
void orc_get_stats(const orc_t *o, double *tau, double *s1, double *s2, double *cr) {
  if (tau) { tau[0] = o->tauL; tau[1] = o->tauU; }
  if (s1) memcpy(s1, o->stat1, (size_t)o->n3 * 7 * sizeof(double));
  if (s2) memcpy(s2, o->stat2, (size_t)o->n3 * 7 * sizeof(double));
  if (cr) memcpy(cr, o->corr, (size_t)o->n3 * 9 * sizeof(double));
}
void orc_get_mesh(const orc_t *o, double *coef, double *mesh, double *scal) {
  if (coef) {
    double c[12] = {o->f1d2c0, o->f1d2c1, o->f1d2c2, o->f2d1c0, o->f2d1c1, o->f2d1c2,
                    o->b1d2c0, o->b1d2c1, o->b1d2c2, o->b2d1c0, o->b2d1c1, o->b2d1c2};
    memcpy(coef, c, sizeof c);
  }
  if (mesh) {
    int n3 = o->n3;
    memcpy(mesh, o->x3, n3 * sizeof(double));
    memcpy(mesh + n3, o->jac, n3 * sizeof(double));
    memcpy(mesh + 2 * n3, o->invJac, n3 * sizeof(double));
    memcpy(mesh + 3 * n3, o->Xi, n3 * sizeof(double));
  }
  if (scal) {
    double s[8] = {o->dx, o->dy, o->dXi, o->dz1, o->dzN, o->ggM, o->invRe, o->Pi};
    memcpy(scal, s, sizeof s);
  }
}
void orc_get_residual(const orc_t *o, double *R) { memcpy(R, o->R, nsz(o) * 5 * sizeof(double)); }
void orc_get_blown(const orc_t *o, int *ijk) { memcpy(ijk, o->blown, sizeof o->blown); }

int orc_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
