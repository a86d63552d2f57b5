/*
 * pcom_oracle.c - CPU FP64 restatement of the PCOM v2.2 time-stepping path.
 *
 * TEST INFRASTRUCTURE ONLY (see pcom_oracle.h).  It is the checker for the
 * MI355X library and is never linked into it.
 *
 * Conventions
 *  - Arrays are the reference's halo'd arrays, Fortran column-major,
 *    1-based in the index macros below: (ni, nj[, nd3]).
 *  - Each routine mirrors the reference routine of the same name; the
 *    expression order of every floating-point statement follows the
 *    Fortran source left to right, so that results are bit-for-bit what a
 *    non-contracting IEEE FP64 evaluation of that source gives.
 *  - Default-REAL literals of the active build (gfortran, no -r8) are
 *    float32 values widened to FP64: written here as float literals.
 *  - One MPI rank: the halo swap is the zonally periodic copy of
 *    mod_mympi.f90:669-746; the wall rows are left untouched.
 */
#include "pcom_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---- constants, mod_con.f90:16-41 (default-REAL literals are float) ---- */
#define RHO0   1029.0
#define A0     (1.0 / RHO0)
#define GRAV   ((double)9.8f)
#define ARAD   6370e3
#define PI_F   ((double)3.141592653589793f)
#define OMEGA  (PI_F / 43082.0)
#define TORAD  (PI_F / 180.0)
#define GAMMA_T (GRAV * 40.0 / 3901.0)
#define GAMMA_S ((double)(1.0f / (float)(120 * 24 * 60 * 60)))
#define GAMMA_B ((double)0.2f)
#define DELTAT ((double)1.0e-2f)
#define DELTAS ((double)1.0e-3f)
#define CDBOT  ((double)2.6f)
#define TICE   (-1.5)
#define KM_C   ((double)1.0e-5f)
#define AM_C   1.0e3
#define SMAG_C ((double)0.56f)
#define AH_C   1.0e3

static const char *k_const_names[] = {
  "rho0", "a0", "g", "a", "pi", "omega", "torad", "gamma_t", "gamma_s",
  "gamma_b", "deltat", "deltas", "cdbot", "tice", "km_c", "am_c", "smag_c",
  "ah_c"};
static const double k_const_values[] = {
  RHO0, A0, GRAV, ARAD, PI_F, OMEGA, TORAD, GAMMA_T, GAMMA_S, GAMMA_B,
  DELTAT, DELTAS, CDBOT, TICE, KM_C, AM_C, SMAG_C, AH_C};

#ifdef _OPENMP
#include <omp.h>
int orc_threads(void) { return omp_get_max_threads(); }
void orc_set_threads(int n) { if (n > 0) omp_set_num_threads(n); }
#else
int orc_threads(void) { return 1; }
void orc_set_threads(int n) { (void)n; }
#endif

int orc_constants(const char ***names, const double **values) {
  *names = k_const_names;
  *values = k_const_values;
  return (int)(sizeof(k_const_values) / sizeof(k_const_values[0]));
}

/* ---- stagger tables, mod_arrays.f90:148-166, 524-552 ---------------- */
static const int GOI[5] = {0, 0, 1, 1, 0};   /* g%i */
static const int GOJ[5] = {0, 0, 0, 1, 1};   /* g%j */
static const int GEW[5] = {0, 2, 1, 4, 3};   /* g%ew */
static const int GNS[5] = {0, 4, 3, 2, 1};   /* g%ns */
static const int GDI[5] = {0, 3, 4, 1, 2};   /* g%di */

typedef struct { int y, m, d, h, mi, s, mm; } otime;

struct orc {
  int glo_ni, glo_nj, ni, nj, nk, nkp, nim, njm;
  size_t n2, n3, n3p;
  int dtbt, dtbc, rst;
  /* horizontal grids g1j..g4j: rh, tn, dx%x(1), dx%x(2) */
  double *rh[5], *tn[5], *dx1[5], *dx2[5];
  double *hposx, *hposy;   /* (ni,nj,4) */
  double *lon, *lat;       /* (ni), (nj) */
  /* vertical grids gi1 (nkp), gi2 (nk) */
  double *z1, *pr1, *dz1, *dpr1, *z2, *pr2, *dz2, *dpr2;
  /* 3-D grids g12 (T) and g32 (U) */
  int *msk_t, *msk_u, *lev_t, *lev_u;
  double *phih_t, *prh_t, *phih_u, *prh_u;
  double *cv1, *cv2, *cor, *graphihx, *graphihy;
  double *km, *kh;         /* (ni,nj,nkp) */
  double *frc[6];          /* taux tauy t s pa fw, (ni,nj,12) */
  double *bnd[6];          /* (ni,nj) */
  /* prognostic / diagnostic state */
  double *tc, *sc, *tp, *sp;
  double *uc, *vc, *up, *vp, *auc, *avc, *aup, *avp, *aupp, *avpp;
  double *fx, *fy, *am, *alpha, *adp, *wm, *prho1, *prho2;
  double *act, *acs, *acr, *acu, *acv, *acw;
  double *chc, *chp, *mch, *mch2, *acch, *acssh, *ssh;
  double *buc, *bvc, *bup, *bvp, *but, *bvt, *pax, *pay;
  double *bphi[4], *bgra[4];   /* xn xs ye yw */
  int n_ts, n_uv, n_w, n_ch, n_ssh;
  long long step_i;            /* tctr%i */
  long long nsteps;
  long long forcing_taken;     /* step whose daily refresh the caller applied */
  otime ct, pt;
};

#define NI (o->ni)
#define NJ (o->nj)
#define NK (o->nk)
#define I2(i, j) ((size_t)((i) - 1) + (size_t)o->ni * (size_t)((j) - 1))
#define I3(i, j, k) (I2(i, j) + o->n2 * (size_t)((k) - 1))

static double *dalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }
static int *ialloc(size_t n) { return (int *)calloc(n ? n : 1, sizeof(int)); }
static void dfill(double *a, size_t n, double v) {
#pragma omp parallel for schedule(static)
  for (size_t p = 0; p < n; ++p) a[p] = v;
}

/* ---- halo swap, one rank: mod_mympi.f90:669-746 ---------------------- */
static void swp(const orc *o, double *a, int nd3) {
  for (int k = 1; k <= nd3; ++k)
    for (int j = 1; j <= NJ; ++j) a[I3(1, j, k)] = a[I3(NI - 1, j, k)];
  for (int k = 1; k <= nd3; ++k)
    for (int j = 1; j <= NJ; ++j) a[I3(NI, j, k)] = a[I3(2, j, k)];
}
static void swpi(const orc *o, int *a, int nd3) {
  for (int k = 1; k <= nd3; ++k)
    for (int j = 1; j <= NJ; ++j) a[I3(1, j, k)] = a[I3(NI - 1, j, k)];
  for (int k = 1; k <= nd3; ++k)
    for (int j = 1; j <= NJ; ++j) a[I3(NI, j, k)] = a[I3(2, j, k)];
}

/* ---- mod_den.f90:57-118 --------------------------------------------- */
/* LIT(x) is a default-REAL literal of the active build (float32 widened).
 * Built with -DPCOM_R8 the EOS literals are FP64 instead, as the
 * reference's ifort "-r8" Makefile (tools/Makefile_ali_cluster_ifc:6)
 * makes them; that variant exists only to check the EOS against the
 * notes.rst:20-26 statement about the JM95 check value. */
#ifdef PCOM_R8
#define LIT(x) (x)
#else
#define LIT(x) ((double)x##f)
#endif
double orc_den_rho(double t, double s, double p0) {
  double p = p0 * LIT(1e-5);
  double p2 = p * p;
  double t2 = t * t;
  double t3 = t2 * t;
  double t4 = t3 * t;
  double t5 = t4 * t;
  double s32 = s * sqrt(s);
  double s2 = s * s;
  double r0 = LIT(999.842594) + LIT(6.793952e-2) * t - LIT(9.095290e-3) * t2
            + LIT(1.001685e-4) * t3 - LIT(1.120083e-6) * t4 + LIT(6.536336e-9) * t5;
  double a = LIT(8.24493e-1) - LIT(4.0899e-3) * t + LIT(7.6438e-5) * t2
           - LIT(8.2467e-7) * t3 + LIT(5.3875e-9) * t4;
  double b = LIT(-5.72466e-3) + LIT(1.0227e-4) * t - LIT(1.6546e-6) * t2;
  double r = r0 + a * s + b * s32 + LIT(4.8314e-4) * s2;
  double k = LIT(1.965933e4) + LIT(1.444304e2) * t - LIT(1.706103) * t2
           + LIT(9.648704e-3) * t3 - LIT(4.190253e-5) * t4
           + (LIT(5.284855e1) - LIT(3.101089e-1) * t + LIT(6.283263e-3) * t2
              - LIT(5.084188e-5) * t3) * s
           + (LIT(3.886640e-1) + LIT(9.085835e-3) * t - LIT(4.619924e-4) * t2) * s32
           + (LIT(3.186519) + LIT(2.212276e-2) * t - LIT(2.984642e-4) * t2
              + LIT(1.956415e-6) * t3) * p
           + (LIT(6.704388e-3) - LIT(1.847318e-4) * t + LIT(2.059331e-7) * t2) * s * p
           + LIT(1.480266e-4) * s32 * p
           + (LIT(2.102898e-4) - LIT(1.202016e-5) * t + LIT(1.394680e-7) * t2) * p2
           + (LIT(-2.040237e-6) + LIT(6.128773e-8) * t + LIT(6.207323e-10) * t2) * s * p2;
  return r / (1.0 - p / k);
}

void orc_den_rho_n(size_t n, const double *t, const double *s,
                   const double *p, double *rho) {
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < n; ++q) rho[q] = orc_den_rho(t[q], s[q], p[q]);
}

/* mod_den.f90:33-55 */
static void den_alpha(orc *o, double *alpha, const double *t, const double *s,
                      const double *ch, const int *mask) {
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k <= NK; ++k)
    for (int j = 1; j <= NJ; ++j)
      for (int i = 1; i <= NI; ++i)
        if (mask[I3(i, j, k)] == 1)
          alpha[I3(i, j, k)] = 1.0 / orc_den_rho(t[I3(i, j, k)], s[I3(i, j, k)],
                                                 ch[I2(i, j)] * o->pr2[k - 1]);
}

/* mod_den.f90:120-153 */
void orc_den_prho(orc *o) {
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k <= NK; ++k)
    for (int j = 1; j <= NJ; ++j)
      for (int i = 1; i <= NI; ++i)
        if (o->msk_t[I3(i, j, k)] > 0) {
          double t = o->tc[I3(i, j, k)], s = o->sc[I3(i, j, k)];
          double p = o->chc[I2(i, j)] * o->pr2[k - 1];
          double ra = orc_den_rho(t + DELTAT, s, p);
          double rb = orc_den_rho(t - DELTAT, s, p);
          o->prho1[I3(i, j, k)] = (ra - rb) / (2 * DELTAT);
          ra = orc_den_rho(t, s + DELTAS, p);
          rb = orc_den_rho(t, s - DELTAS, p);
          o->prho2[I3(i, j, k)] = (ra - rb) / (2 * DELTAS);
        }
}

void orc_den_alpha_main(orc *o) { den_alpha(o, o->alpha, o->tc, o->sc, o->chc, o->msk_t); }

/* ---- mod_op.f90 ------------------------------------------------------- */

/* ter_r3d / ter_r2d, mod_op.f90:361-427 */
static void ter(orc *o, double *ans, const double *var, int nd3, int ga, int gb, double dft) {
  dfill(ans, o->n2 * (size_t)nd3, dft);
  int oi = GOI[ga], oj = GOJ[ga];
  if (gb == GEW[ga]) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 1; k <= nd3; ++k)
      for (int j = 1; j <= NJ; ++j)
        for (int i = 1 + oi; i <= o->nim + oi; ++i)
          ans[I3(i, j, k)] = 0.5 * (var[I3(i - oi, j, k)] + var[I3(i - oi + 1, j, k)]);
  } else if (gb == GNS[ga]) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 1; k <= nd3; ++k)
      for (int j = 1 + oj; j <= o->njm + oj; ++j)
        for (int i = 1; i <= NI; ++i)
          ans[I3(i, j, k)] = 0.5 * (var[I3(i, j - oj, k)] + var[I3(i, j - oj + 1, k)]);
  } else if (gb == GDI[ga]) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 1; k <= nd3; ++k)
      for (int j = 1 + oj; j <= o->njm + oj; ++j)
        for (int i = 1 + oi; i <= o->nim + oi; ++i)
          ans[I3(i, j, k)] = (var[I3(i - oi, j - oj, k)] + var[I3(i - oi, j - oj + 1, k)] +
                              var[I3(i - oi + 1, j - oj, k)] + var[I3(i - oi + 1, j - oj + 1, k)]) * 0.25;
  } else {
    abort();  /* same grid: the reference stops (mod_op.f90:374-375) */
  }
  swp(o, ans, nd3);
}

/* p1_r3d / p1_r3d_b / p1_r2d, mod_op.f90:496-527, 561-620 */
static void p1(orc *o, double *ans, const double *var, int nd3, int ga, int gb, const int *mask) {
  size_t n = o->n2 * (size_t)nd3;
  double *tmp = NULL;
  const double *t = var;
  if (GEW[gb] != ga) {
    tmp = dalloc(n);
    ter(o, tmp, var, nd3, ga, GEW[gb], 0.0);
    t = tmp;
  }
  dfill(ans, n, 0.0);
  int oi = GOI[ga];
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k <= nd3; ++k)
    for (int j = 1; j <= NJ; ++j)
      for (int i = 1 + oi; i <= o->nim + oi; ++i) {
        double d = t[I3(i - oi + 1, j, k)] - t[I3(i - oi, j, k)];
        if (mask) d = d * (double)mask[I3(i - oi + 1, j, k)] * (double)mask[I3(i - oi, j, k)];
        ans[I3(i, j, k)] = d;
      }
  const double *dx = o->dx1[gb];
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k <= nd3; ++k)
    for (size_t p = 0; p < o->n2; ++p) ans[p + o->n2 * (size_t)(k - 1)] /= dx[p];
  swp(o, ans, nd3);
  free(tmp);
}

/* p2_r3d / p2_r3d_b / p2_r2d, mod_op.f90:622-713 */
static void p2(orc *o, double *ans, const double *var, int nd3, int ga, int gb, const int *mask) {
  size_t n = o->n2 * (size_t)nd3;
  double *tmp = NULL;
  const double *t = var;
  if (GNS[gb] != ga) {
    tmp = dalloc(n);
    ter(o, tmp, var, nd3, ga, GNS[gb], 0.0);
    t = tmp;
  }
  dfill(ans, n, 0.0);
  int oj = GOJ[ga];
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k <= nd3; ++k)
    for (int j = 1 + oj; j <= o->njm + oj; ++j)
      for (int i = 1; i <= NI; ++i) {
        double d = t[I3(i, j - oj + 1, k)] - t[I3(i, j - oj, k)];
        if (mask) d = d * (double)mask[I3(i, j - oj + 1, k)] * (double)mask[I3(i, j - oj, k)];
        ans[I3(i, j, k)] = d;
      }
  const double *dy = o->dx2[gb];
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k <= nd3; ++k)
    for (size_t p = 0; p < o->n2; ++p) ans[p + o->n2 * (size_t)(k - 1)] /= dy[p];
  swp(o, ans, nd3);
  free(tmp);
}

/* p3_r3d, mod_op.f90:715-735: var (nkp) -> ans (nk) */
static void p3(orc *o, double *ans, const double *var) {
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k <= NK; ++k)
    for (size_t p = 0; p < o->n2; ++p)
      ans[p + o->n2 * (k - 1)] = (var[p + o->n2 * (k - 1)] - var[p + o->n2 * k]) / o->dpr2[k - 1];
}

/* pz_r3d, mod_op.f90:762-790: var (nk) -> ans (nkp), height derivative */
static void pz(orc *o, double *ans, const double *var, const int *msk) {
  dfill(ans, o->n3p, 0.0);
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 2; k <= NK; ++k)
    for (size_t p = 0; p < o->n2; ++p)
      ans[p + o->n2 * (k - 1)] = (var[p + o->n2 * (k - 2)] - var[p + o->n2 * (k - 1)]) *
                                 (double)msk[p + o->n2 * (k - 2)] * (double)msk[p + o->n2 * (k - 1)] /
                                 o->dz1[k - 1];
}

/* dx3_r3d_b, mod_op.f90:792-811 */
static void dx3(orc *o, double *ans, const double *var) {
  dfill(ans, o->n3p, 0.0);
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 2; k <= NK; ++k)
    for (size_t p = 0; p < o->n2; ++p)
      ans[p + o->n2 * (k - 1)] = var[p + o->n2 * (k - 2)] - var[p + o->n2 * (k - 1)];
}

/* vter_r3d, mod_op.f90:327-359. from_gi = 2: nk -> nkp; 1: nkp -> nk */
static void vter(orc *o, double *ans, const double *var, int from_gi) {
  if (from_gi == 2) {
    dfill(ans, o->n3p, 0.0);
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 2; k <= NK; ++k)
      for (size_t p = 0; p < o->n2; ++p)
        ans[p + o->n2 * (k - 1)] = (var[p + o->n2 * (k - 2)] + var[p + o->n2 * (k - 1)]) * 0.5;
  } else {
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 1; k <= NK; ++k)
      for (size_t p = 0; p < o->n2; ++p)
        ans[p + o->n2 * (k - 1)] = (var[p + o->n2 * (k - 1)] + var[p + o->n2 * k]) * 0.5;
  }
}

/* px_r3d_center, mod_op.f90:529-559 (hg = guj = g3, hg%ew = g4) */
static void px_center(orc *o, double *ans, const double *var, int nd3) {
  dfill(ans, o->n2 * (size_t)nd3, 0.0);
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k <= nd3; ++k)
    for (int j = 1; j <= NJ; ++j)
      for (int i = 2; i <= NI - 1; ++i) {
        double dx = ARAD * o->rh[4][I2(i, j)] * (o->dx1[4][I2(i, j)] + o->dx1[4][I2(i + 1, j)]);
        ans[I3(i, j, k)] = (var[I3(i + 1, j, k)] - var[I3(i - 1, j, k)]) / dx;
      }
  swp(o, ans, nd3);
}

/* div_r3d / div_r2d, mod_op.f90:813-848 */
static void divg(orc *o, double *ans, const double *va, const double *vb, int nd3, int ga, int gb, int gc) {
  size_t n = o->n2 * (size_t)nd3;
  double *wa = dalloc(n), *wb = dalloc(n), *vr = dalloc(n);
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < nd3; ++k)
    for (size_t p = 0; p < o->n2; ++p) vr[p + o->n2 * k] = vb[p + o->n2 * k] * o->rh[gb][p];
  p1(o, wa, va, nd3, ga, gc, NULL);
  p2(o, wb, vr, nd3, gb, gc, NULL);
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < nd3; ++k)
    for (size_t p = 0; p < o->n2; ++p) {
      size_t q = p + o->n2 * k;
      ans[q] = (wa[q] + wb[q]) / (ARAD * o->rh[gc][p]);
    }
  free(wa); free(wb); free(vr);
}

/* gra_r2d, gra_r2d_b, gra_r3d, gra_r3d_b, mod_op.f90:850-902 */
static void gra(orc *o, double *ax, double *ay, const double *var, int nd3, int ga, int gb, int gc) {
  p1(o, ax, var, nd3, ga, gb, NULL);
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < nd3; ++k)
    for (size_t p = 0; p < o->n2; ++p) ax[p + o->n2 * k] = ax[p + o->n2 * k] / (ARAD * o->rh[gb][p]);
  p2(o, ay, var, nd3, ga, gc, NULL);
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < o->n2 * (size_t)nd3; ++q) ay[q] = ay[q] / ARAD;
}

/* op_lap, mod_op.f90:904-925 (2-D) */
static void lap(orc *o, double *ans, const double *var, int ga, int gb) {
  double *wa = dalloc(o->n2), *wb = dalloc(o->n2);
  int ge = GEW[gb], gn = GNS[gb];
  p1(o, wa, var, 1, ga, ge, NULL);
#pragma omp parallel for schedule(static)
  for (size_t p = 0; p < o->n2; ++p) wa[p] = wa[p] / o->rh[ge][p];
  p1(o, wb, wa, 1, ge, gb, NULL);
#pragma omp parallel for schedule(static)
  for (size_t p = 0; p < o->n2; ++p) ans[p] = wb[p];
  p2(o, wa, var, 1, ga, gn, NULL);
#pragma omp parallel for schedule(static)
  for (size_t p = 0; p < o->n2; ++p) wa[p] = wa[p] * o->rh[gn][p];
  p2(o, wb, wa, 1, gn, gb, NULL);
#pragma omp parallel for schedule(static)
  for (size_t p = 0; p < o->n2; ++p) ans[p] = ans[p] + wb[p];
#pragma omp parallel for schedule(static)
  for (size_t p = 0; p < o->n2; ++p) ans[p] = ans[p] / (ARAD * ARAD * o->rh[gb][p]);
  free(wa); free(wb);
}

/* int_tobot, mod_op.f90:927-944 (grd = g12: dpr = gi2%dpr) */
static void int_tobot(orc *o, double *ans, const double *var, const int *msk) {
  double *t = dalloc(o->n3);
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k <= NK; ++k)
    for (size_t p = 0; p < o->n2; ++p) {
      size_t q = p + o->n2 * (k - 1);
      t[q] = var[q] * o->dpr2[k - 1] * (double)msk[q];
    }
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k <= NK - 1; ++k)
    for (size_t p = 0; p < o->n2; ++p) {
      double s = 0.0;
      for (int kk = k + 1; kk <= NK; ++kk) s += t[p + o->n2 * (kk - 1)];
      ans[p + o->n2 * (k - 1)] = t[p + o->n2 * (k - 1)] * 0.5 + s;
    }
#pragma omp parallel for schedule(static)
  for (size_t p = 0; p < o->n2; ++p) ans[p + o->n2 * (NK - 1)] = t[p + o->n2 * (NK - 1)] * 0.5;
  free(t);
}

/* op_vint_ns (dir = 2) / op_vint_ew (dir = 1), mod_op.f90:946-981:
 * ra uses g32 msk/prh, rb the same cshift'ed by -1 along dir. */
static void vint_face(orc *o, const double *var, double *ra, double *rb, int dir) {
#pragma omp parallel for collapse(2) schedule(static)
  for (int j = 1; j <= NJ; ++j)
    for (int i = 1; i <= NI; ++i) {
      int is = i, js = j;
      if (dir == 1) is = (i == 1) ? NI : i - 1;
      else js = (j == 1) ? NJ : j - 1;
      double sa = 0.0, sb = 0.0;
      for (int k = 1; k <= NK; ++k) {
        double vd = var[I3(i, j, k)] * o->dpr2[k - 1];
        sa += vd * (double)o->msk_u[I3(i, j, k)];
        sb += vd * (double)o->msk_u[I3(is, js, k)];
      }
      ra[I2(i, j)] = sa / o->prh_u[I2(i, j)];
      rb[I2(i, j)] = sb / o->prh_u[I2(is, js)];
    }
  swp(o, ra, 1);
  swp(o, rb, 1);
}

/* vint_r3d, mod_op.f90:983-995 */
static void vint(orc *o, double *ans, const double *var, const int *msk, const double *prh) {
#pragma omp parallel for schedule(static)
  for (size_t p = 0; p < o->n2; ++p) {
    double s = 0.0;
    for (int k = 1; k <= NK; ++k) {
      size_t q = p + o->n2 * (k - 1);
      s += var[q] * o->dpr2[k - 1] * (double)msk[q];
    }
    ans[p] = s / prh[p];
  }
}

/* fri_r3d, mod_op.f90:103-167 */
static void fri_r3d(orc *o, double *ans, const double *sch, const double *v,
                    const double *vp, const double *vc, const double *tau) {
  size_t n3 = o->n3, n2 = o->n2;
  double *mx = dalloc(n3), *my = dalloc(n3), *wka = dalloc(n3), *wkp = dalloc(o->n3p);
  double *cha = dalloc(n2), *chb = dalloc(n2), *va = dalloc(n3), *vb = dalloc(n3);
  gra(o, mx, my, v, NK, 3, 4, 2);
  ter(o, cha, o->chc, 1, 1, 4, 0.0);
  ter(o, chb, o->chc, 1, 1, 2, 0.0);
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < NK; ++k)
    for (size_t p = 0; p < n2; ++p) {
      va[p + n2 * k] = cha[p] * mx[p + n2 * k];
      vb[p + n2 * k] = chb[p] * my[p + n2 * k];
    }
  divg(o, wka, va, vb, NK, 4, 2, 3);
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < NK; ++k)
    for (size_t p = 0; p < n2; ++p) {
      size_t q = p + n2 * k;
      ans[q] = o->am[q] / sch[p] * wka[q];
    }
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < NK; ++k)
    for (size_t p = 0; p < n2; ++p) {
      size_t q = p + n2 * k;
      ans[q] = ans[q] + o->am[q] * o->cv1[p] * vp[q];
    }
  px_center(o, wka, vc, NK);
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < NK; ++k)
    for (size_t p = 0; p < n2; ++p) {
      size_t q = p + n2 * k;
      ans[q] = ans[q] + o->am[q] * o->cv2[p] * sch[p] * wka[q];
    }
  pz(o, wkp, vp, o->msk_u);
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < NK; ++k)
    for (size_t p = 0; p < n2; ++p) {
      size_t q = p + n2 * k;   /* km%v(:,:,1:nk) paired element-wise */
      wkp[q] = GRAV * RHO0 * o->km[q] * wkp[q] / (sch[p] * sch[p]);
    }
#pragma omp parallel for schedule(static)
  for (size_t p = 0; p < n2; ++p) wkp[p] = GRAV * tau[p] / sch[p];
#pragma omp parallel for collapse(2) schedule(static)
  for (int j = 1; j <= NJ; ++j)
    for (int i = 1; i <= NI; ++i) {
      int k = o->lev_u[I2(i, j)];
      if (k > 0) {
        double vv = v[I3(i, j, k)], cc = vc[I3(i, j, k)];
        double mag = sqrt(vv * vv + cc * cc);
        wkp[I3(i, j, k + 1)] = vv * GRAV * CDBOT * mag / sch[I2(i, j)];
      }
    }
  p3(o, wka, wkp);
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < n3; ++q) ans[q] = ans[q] * (double)o->msk_u[q] + wka[q];
  free(mx); free(my); free(wka); free(wkp); free(cha); free(chb); free(va); free(vb);
}

/* op_fri, mod_op.f90:169-184 */
static void op_fri(orc *o, const double *sch) {
  size_t n3 = o->n3, n2 = o->n2;
  double *u = dalloc(n3), *v = dalloc(n3), *mv = dalloc(n3);
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < NK; ++k)
    for (size_t p = 0; p < n2; ++p) {
      size_t q = p + n2 * k;
      u[q] = o->uc[q] / sch[p];
      v[q] = o->vc[q] / sch[p];
      mv[q] = -v[q];
    }
  fri_r3d(o, o->fx, sch, u, o->uc, mv, o->bnd[0]);
  fri_r3d(o, o->fy, sch, v, o->vc, u, o->bnd[1]);
  free(u); free(v); free(mv);
}

/* op_dif, mod_op.f90:186-222 (var on g1) */
static void op_dif(orc *o, double *ans, const double *ch, const double *var) {
  size_t n3 = o->n3, n2 = o->n2;
  double *mx = dalloc(n3), *my = dalloc(n3), *wk = dalloc(n3), *wkp = dalloc(o->n3p);
  double *cha = dalloc(n2), *chb = dalloc(n2), *va = dalloc(n3), *vb = dalloc(n3);
  p1(o, mx, var, NK, 1, 2, o->msk_t);
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < NK; ++k)
    for (size_t p = 0; p < n2; ++p) mx[p + n2 * k] = mx[p + n2 * k] / (ARAD * o->rh[2][p]);
  p2(o, my, var, NK, 1, 4, o->msk_t);
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < n3; ++q) my[q] = my[q] / ARAD;
  ter(o, cha, ch, 1, 1, 2, 0.0);
  ter(o, chb, ch, 1, 1, 4, 0.0);
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < NK; ++k)
    for (size_t p = 0; p < n2; ++p) {
      va[p + n2 * k] = cha[p] * mx[p + n2 * k];
      vb[p + n2 * k] = chb[p] * my[p + n2 * k];
    }
  divg(o, wk, va, vb, NK, 2, 4, 1);
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < n3; ++q) ans[q] = AH_C * wk[q];
  pz(o, wkp, var, o->msk_t);
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < o->n3p; ++q) wkp[q] = GRAV * RHO0 * o->kh[q] * wkp[q];
  p3(o, wk, wkp);
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < n3; ++q) ans[q] = ans[q] + wk[q];
  free(mx); free(my); free(wk); free(wkp); free(cha); free(chb); free(va); free(vb);
}

/* op_adv, mod_op.f90:224-278 (var on g3) */
static void op_adv(orc *o, double *ans, const double *var, const double *sch) {
  size_t n3 = o->n3, n2 = o->n2, n3p = o->n3p;
  double *wka = dalloc(n3), *wkb = dalloc(n3), *u = dalloc(n3), *v = dalloc(n3);
  double *m1 = dalloc(n3), *m2 = dalloc(n3);
  double *wkap = dalloc(n3p), *wkbp = dalloc(n3p), *w = dalloc(n3p);
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < NK; ++k)
    for (size_t p = 0; p < n2; ++p) wka[p + n2 * k] = o->uc[p + n2 * k] / sch[p];
  ter(o, u, wka, NK, 3, 4, 0.0);
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < NK; ++k)
    for (size_t p = 0; p < n2; ++p) {
      size_t q = p + n2 * k;
      wka[q] = o->vc[q] / sch[p];
      wka[q] = wka[q] * o->rh[3][p];
    }
  ter(o, v, wka, NK, 3, 2, 0.0);
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < o->nkp; ++k)
    for (size_t p = 0; p < n2; ++p) wkap[p + n2 * k] = o->wm[p + n2 * k] / o->chc[p];
  ter(o, w, wkap, o->nkp, 1, 3, 0.0);
  ter(o, wka, var, NK, 3, 4, 0.0);
  ter(o, wkb, var, NK, 3, 2, 0.0);
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < n3; ++q) { m1[q] = wka[q] * u[q]; m2[q] = wkb[q] * v[q]; }
  p1(o, wka, m1, NK, 4, 3, NULL);
  p2(o, wkb, m2, NK, 2, 3, NULL);
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < NK; ++k)
    for (size_t p = 0; p < n2; ++p) {
      size_t q = p + n2 * k;
      ans[q] = (wka[q] + wkb[q]) / (ARAD * o->rh[3][p]);
    }
  vter(o, wkap, var, 2);
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < n3p; ++q) wkap[q] = wkap[q] * w[q];
#pragma omp parallel for schedule(static)
  for (size_t p = 0; p < n2; ++p) { wkbp[p] = 0.0; wkbp[p + n2 * NK] = 0.0; }
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 2; k <= NK; ++k)
    for (size_t p = 0; p < n2; ++p)
      wkbp[p + n2 * (k - 1)] = wkap[p + n2 * (k - 1)] * (double)o->msk_u[p + n2 * (k - 2)] *
                               (double)o->msk_u[p + n2 * (k - 1)];
  p3(o, wka, wkbp);
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < n3; ++q) ans[q] = ans[q] + wka[q];
  p1(o, m1, u, NK, 4, 3, NULL);
  p2(o, m2, v, NK, 2, 3, NULL);
  p3(o, wka, w);
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < NK; ++k)
    for (size_t p = 0; p < n2; ++p) {
      size_t q = p + n2 * k;
      ans[q] = ans[q] - 0.5 * var[q] * ((m1[q] + m2[q]) / (ARAD * o->rh[3][p]) + wka[q]);
    }
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < n3; ++q) ans[q] = ans[q] * (double)o->msk_u[q];
  free(wka); free(wkb); free(u); free(v); free(m1); free(m2); free(wkap); free(wkbp); free(w);
}

/* op_adv_ts, mod_op.f90:280-325 (var on g1) */
static void op_adv_ts(orc *o, double *ans, const double *var, const double *wm) {
  size_t n3 = o->n3, n2 = o->n2, n3p = o->n3p;
  double *sch = dalloc(n2), *um1 = dalloc(n3), *um2 = dalloc(n3), *m1 = dalloc(n3), *m2 = dalloc(n3);
  double *wka = dalloc(n3), *wkb = dalloc(n3), *tmp = dalloc(n3), *wkp = dalloc(n3p);
  ter(o, sch, o->chc, 1, 1, 3, 1.0);
#pragma omp parallel for schedule(static)
  for (size_t p = 0; p < n2; ++p) sch[p] = sqrt(sch[p]);
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < NK; ++k)
    for (size_t p = 0; p < n2; ++p) {
      size_t q = p + n2 * k;
      um1[q] = o->uc[q] * sch[p];
      um2[q] = o->vc[q] * sch[p] * o->rh[3][p];
    }
  p1(o, m1, var, NK, 1, 2, NULL);
  p2(o, m2, var, NK, 1, 4, NULL);
  ter(o, wka, um1, NK, 3, 2, 0.0);
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < n3; ++q) tmp[q] = wka[q] * m1[q];
  ter(o, wkb, tmp, NK, 2, 1, 0.0);
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < NK; ++k)
    for (size_t p = 0; p < n2; ++p) ans[p + n2 * k] = wkb[p + n2 * k] / (ARAD * o->rh[1][p]);
  ter(o, wka, um2, NK, 3, 4, 0.0);
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < n3; ++q) tmp[q] = wka[q] * m2[q];
  ter(o, wkb, tmp, NK, 4, 1, 0.0);
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < NK; ++k)
    for (size_t p = 0; p < n2; ++p) {
      size_t q = p + n2 * k;
      ans[q] = ans[q] + wkb[q] / (ARAD * o->rh[1][p]);
    }
  dx3(o, wkp, var);
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < n3p; ++q) wkp[q] = wm[q] * wkp[q];
  vter(o, wka, wkp, 1);
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < NK; ++k)
    for (size_t p = 0; p < n2; ++p) wka[p + n2 * k] = wka[p + n2 * k] / o->dpr2[k];
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < n3; ++q) ans[q] = ans[q] + wka[q];
  free(sch); free(um1); free(um2); free(m1); free(m2); free(wka); free(wkb); free(tmp); free(wkp);
}

/* ---- mod_int.f90 ------------------------------------------------------ */

/* int_pgra, mod_int.f90:64-88 */
void orc_int_pgra(orc *o) {
  size_t n3 = o->n3, n2 = o->n2;
  double *wk = dalloc(n3), *wkb = dalloc(n3), *wkc = dalloc(n3), *wkx = dalloc(n3), *wky = dalloc(n3);
  int_tobot(o, o->adp, o->alpha, o->msk_t);
  ter(o, wkb, o->adp, NK, 1, 2, 0.0);
  ter(o, wkc, o->alpha, NK, 1, 2, 0.0);
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < NK; ++k)
    for (size_t p = 0; p < n2; ++p) {
      size_t q = p + n2 * k;
      wk[q] = wkb[q] + o->pr2[k] * wkc[q];
    }
  vint_face(o, wk, o->bphi[0], o->bphi[1], 2);
  ter(o, wkb, o->adp, NK, 1, 4, 0.0);
  ter(o, wkc, o->alpha, NK, 1, 4, 0.0);
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < NK; ++k)
    for (size_t p = 0; p < n2; ++p) {
      size_t q = p + n2 * k;
      wk[q] = wkb[q] + o->pr2[k] * wkc[q];
    }
  vint_face(o, wk, o->bphi[2], o->bphi[3], 1);
  gra(o, wkx, wky, o->adp, NK, 1, 2, 4);
  vint_face(o, wkx, o->bgra[0], o->bgra[1], 2);
  vint_face(o, wky, o->bgra[2], o->bgra[3], 1);
  free(wk); free(wkb); free(wkc); free(wkx); free(wky);
}

/* upwelling, mod_int.f90:626-647 */
static void upwelling(orc *o) {
  size_t n3 = o->n3, n2 = o->n2;
  double *sch = dalloc(n2), *va = dalloc(n3), *vb = dalloc(n3), *wk3 = dalloc(n3), *wk2 = dalloc(n2);
  ter(o, sch, o->chp, 1, 1, 3, 0.0);
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < NK; ++k)
    for (size_t p = 0; p < n2; ++p) {
      size_t q = p + n2 * k;
      double s3 = sqrt(sch[p]);
      va[q] = s3 * o->uc[q];
      vb[q] = s3 * o->vc[q];
    }
  divg(o, wk3, va, vb, NK, 3, 3, 1);
  vint(o, wk2, wk3, o->msk_t, o->prh_t);
#pragma omp parallel for schedule(static)
  for (size_t p = 0; p < n2; ++p) wk2[p] = -wk2[p];
  for (int k = 2; k <= NK; ++k)
#pragma omp parallel for schedule(static)
    for (size_t p = 0; p < n2; ++p) {
      size_t q = p + n2 * (k - 1), qm = p + n2 * (k - 2);
      o->wm[q] = o->wm[qm] + o->dpr2[k - 2] * (wk2[p] + wk3[qm]);
      o->wm[q] = o->wm[q] * (double)o->msk_t[q];
    }
  free(sch); free(va); free(vb); free(wk3); free(wk2);
}

static void copy3(const orc *o, double *dst, const double *src) { memcpy(dst, src, o->n3 * sizeof(double)); }

/* int_readyc, mod_int.f90:90-151 */
void orc_int_readyc(orc *o) {
  size_t n2 = o->n2;
  double *sch = dalloc(n2), *wkx = dalloc(n2), *wky = dalloc(n2);
  gra(o, wkx, wky, o->bnd[4], 1, 1, 2, 4);
  ter(o, o->pax, wkx, 1, 2, 3, 0.0);
  ter(o, o->pay, wky, 1, 4, 3, 0.0);
  ter(o, sch, o->chc, 1, 1, 3, 1.0);
#pragma omp parallel for schedule(static)
  for (size_t p = 0; p < n2; ++p) sch[p] = sqrt(sch[p]);
#pragma omp parallel for schedule(static)
  for (size_t p = 0; p < n2; ++p) if (o->lev_u[p] == 0) sch[p] = 1.0;
  upwelling(o);
  copy3(o, o->aupp, o->aup);
  copy3(o, o->avpp, o->avp);
  copy3(o, o->aup, o->auc);
  copy3(o, o->avp, o->avc);
  op_adv(o, o->auc, o->uc, sch);
  op_adv(o, o->avc, o->vc, sch);
  if (o->rst != 1) {
    if (o->step_i == 1) {
      copy3(o, o->aupp, o->auc); copy3(o, o->avpp, o->avc);
      copy3(o, o->aup, o->auc);  copy3(o, o->avp, o->avc);
    } else if (o->step_i == 2) {
      copy3(o, o->aupp, o->aup); copy3(o, o->avpp, o->avp);
    }
  }
  op_fri(o, sch);
  dfill(o->but, n2, 0.0);
  dfill(o->bvt, n2, 0.0);
  for (int k = 0; k < NK; ++k)
#pragma omp parallel for schedule(static)
    for (size_t p = 0; p < n2; ++p) {
      size_t q = p + n2 * k;
      double x = -o->auc[q] - sch[p] * A0 * o->pax[p] + o->fx[q];
      double y = -o->avc[q] - sch[p] * A0 * o->pay[p] + o->fy[q];
      o->but[p] = o->but[p] + x * o->dpr2[k] * (double)o->msk_u[q];
      o->bvt[p] = o->bvt[p] + y * o->dpr2[k] * (double)o->msk_u[q];
    }
#pragma omp parallel for schedule(static)
  for (size_t p = 0; p < n2; ++p) {
    o->but[p] = o->but[p] / o->prh_u[p];
    o->bvt[p] = o->bvt[p] / o->prh_u[p];
  }
  free(sch); free(wkx); free(wky);
}

/* int_trop, mod_int.f90:153-296.  The automatic arrays grachint and
 * chgraint are never written on row nj (x(1)); the reference reads them
 * there uninitialised and multiplies by seau = 0; here they start at 0. */
void orc_int_trop(orc *o) {
  size_t n2 = o->n2;
  double *sch = dalloc(n2), *wk = dalloc(n2), *s1 = dalloc(n2), *s2 = dalloc(n2);
  double *upb1 = dalloc(n2), *upb2 = dalloc(n2), *l1 = dalloc(n2), *l2 = dalloc(n2);
  double *vs1 = dalloc(n2), *vs2 = dalloc(n2), *vz1 = dalloc(n2), *vz2 = dalloc(n2);
  double *gx = dalloc(n2), *gy = dalloc(n2), *gi1 = dalloc(n2), *gi2 = dalloc(n2);
  double *ci1 = dalloc(n2), *ci2 = dalloc(n2), *pg1 = dalloc(n2), *pg2 = dalloc(n2);
  double *pu1 = dalloc(n2), *pu2 = dalloc(n2), *wm1 = dalloc(n2), *wm2 = dalloc(n2);
  const int *seau = o->msk_u;  /* gu%msk(:,:,1) */
  double bt = (double)o->dtbt;
#pragma omp parallel for schedule(static)
  for (size_t p = 0; p < n2; ++p) { upb1[p] = o->buc[p]; upb2[p] = o->bvc[p]; }
#pragma omp parallel for schedule(static)
  for (size_t p = 0; p < n2; ++p) { o->mch[p] = o->chc[p]; o->mch2[p] = o->chc[p]; }
  int nstep = o->dtbc / o->dtbt * 2;
  for (int nt = 1; nt <= nstep; ++nt) {
    dfill(sch, n2, 1.0);
    ter(o, wk, o->chp, 1, 1, 3, 0.0);
#pragma omp parallel for schedule(static)
    for (size_t p = 0; p < n2; ++p) if (o->lev_u[p] > 0) sch[p] = sqrt(wk[p]);
#pragma omp parallel for schedule(static)
    for (size_t p = 0; p < n2; ++p) {
      s1[p] = o->prh_u[p] * sch[p] * o->bup[p];
      s2[p] = o->prh_u[p] * sch[p] * o->bvp[p];
    }
    divg(o, wk, s1, s2, 1, 3, 3, 1);
#pragma omp parallel for schedule(static)
    for (size_t p = 0; p < n2; ++p) {
      wk[p] = o->chp[p] - wk[p] * bt * GAMMA_B / o->prh_t[p];
      if (o->lev_t[p] > 0) o->chc[p] = wk[p];
    }
    ter(o, wk, o->chc, 1, 1, 3, 0.0);
#pragma omp parallel for schedule(static)
    for (size_t p = 0; p < n2; ++p) if (o->lev_u[p] > 0) sch[p] = sqrt(wk[p]);
    lap(o, l1, o->bup, 3, 3);
    lap(o, l2, o->bvp, 3, 3);
#pragma omp parallel for schedule(static)
    for (size_t p = 0; p < n2; ++p) {
      vs1[p] = (double)seau[p] * o->am[p] * l1[p];
      vs2[p] = (double)seau[p] * o->am[p] * l2[p];
    }
    if (nt == 1) for (size_t p = 0; p < n2; ++p) { vz1[p] = vs1[p]; vz2[p] = vs2[p]; }
    gra(o, gx, gy, o->chc, 1, 1, 2, 4);
#pragma omp parallel for collapse(2) schedule(static)
    for (int j = 1; j <= o->njm; ++j)
      for (int i = 1; i <= NI; ++i)
        gi1[I2(i, j)] = 0.5 * (gx[I2(i, j)] * o->bphi[0][I2(i, j)] + gx[I2(i, j + 1)] * o->bphi[1][I2(i, j + 1)]);
#pragma omp parallel for collapse(2) schedule(static)
    for (int j = 1; j <= NJ; ++j)
      for (int i = 1; i <= o->nim; ++i)
        gi2[I2(i, j)] = 0.5 * (gy[I2(i, j)] * o->bphi[2][I2(i, j)] + gy[I2(i + 1, j)] * o->bphi[3][I2(i + 1, j)]);
    swp(o, gi1, 1);
    swp(o, gi2, 1);
    ter(o, wk, o->chc, 1, 1, 2, 0.0);
#pragma omp parallel for collapse(2) schedule(static)
    for (int j = 1; j <= o->njm; ++j)
      for (int i = 1; i <= NI; ++i)
        ci1[I2(i, j)] = 0.5 * (wk[I2(i, j)] * o->bgra[0][I2(i, j)] + wk[I2(i, j + 1)] * o->bgra[1][I2(i, j + 1)]);
    ter(o, wk, o->chc, 1, 1, 4, 0.0);
#pragma omp parallel for collapse(2) schedule(static)
    for (int j = 1; j <= NJ; ++j)
      for (int i = 1; i <= o->nim; ++i)
        ci2[I2(i, j)] = 0.5 * (wk[I2(i, j)] * o->bgra[2][I2(i, j)] + wk[I2(i + 1, j)] * o->bgra[3][I2(i + 1, j)]);
    swp(o, ci1, 1);
    swp(o, ci2, 1);
#pragma omp parallel for schedule(static)
    for (size_t p = 0; p < n2; ++p) {
      double su = (double)seau[p];
      pg1[p] = (gi1[p] + ci1[p] + o->graphihx[p]) * su * sch[p];
      pg2[p] = (gi2[p] + ci2[p] + o->graphihy[p]) * su * sch[p];
      pu1[p] = su * (vs1[p] - vz1[p] + o->but[p] - pg1[p] + 0.5 * o->cor[p] * o->bvp[p]);
      pu2[p] = su * (vs2[p] - vz2[p] + o->bvt[p] - pg2[p] - 0.5 * o->cor[p] * o->bup[p]);
      wm1[p] = pu1[p] * bt + o->bup[p];
      wm2[p] = pu2[p] * bt + o->bvp[p];
      double x = 0.5 * o->cor[p] * bt;
      o->buc[p] = (wm1[p] + 0.5 * o->cor[p] * bt * wm2[p]) / (1 + x * x);
      o->bvc[p] = (wm2[p] - 0.5 * o->cor[p] * bt * wm1[p]) / (1 + x * x);
      o->buc[p] = 0.5 * (o->buc[p] + o->bup[p]);
      o->bvc[p] = 0.5 * (o->bvc[p] + o->bvp[p]);
      pu1[p] = su * (vs1[p] - vz1[p] - pg1[p] + o->but[p] + o->cor[p] * o->bvc[p]);
      pu2[p] = su * (vs2[p] - vz2[p] - pg2[p] + o->bvt[p] - o->cor[p] * o->buc[p]);
      o->buc[p] = pu1[p] * bt + o->bup[p];
      o->bvc[p] = pu2[p] * bt + o->bvp[p];
      o->bup[p] = o->buc[p];
      o->bvp[p] = o->bvc[p];
    }
#pragma omp parallel for schedule(static)
    for (size_t p = 0; p < n2; ++p) {
      s1[p] = o->prh_u[p] * sch[p] * o->buc[p];
      s2[p] = o->prh_u[p] * sch[p] * o->bvc[p];
    }
    divg(o, wk, s1, s2, 1, 3, 3, 1);
#pragma omp parallel for schedule(static)
    for (size_t p = 0; p < n2; ++p) {
      o->chc[p] = o->chp[p] - (double)o->msk_t[p] * wk[p] * bt / o->prh_t[p];
      if (nt <= nstep / 2) o->mch[p] = o->mch[p] + o->chc[p];
      o->chp[p] = o->chc[p];
      upb1[p] = upb1[p] + o->buc[p];
      upb2[p] = upb2[p] + o->bvc[p];
      o->mch2[p] = o->mch2[p] + o->chc[p];
    }
  }
#pragma omp parallel for schedule(static)
  for (size_t p = 0; p < n2; ++p) {
    o->buc[p] = upb1[p] / (double)(nstep + 1);
    o->bvc[p] = upb2[p] / (double)(nstep + 1);
    o->bup[p] = o->buc[p];
    o->bvp[p] = o->bvc[p];
    o->mch[p] = o->mch[p] / (double)(nstep / 2 + 1);
    o->mch2[p] = o->mch2[p] / (double)(nstep + 1);
    o->chc[p] = o->mch2[p];
    o->chp[p] = o->chc[p];
    o->acch[p] = o->acch[p] + o->chc[p];
  }
  o->n_ch += 1;
  free(sch); free(wk); free(s1); free(s2); free(upb1); free(upb2); free(l1); free(l2);
  free(vs1); free(vs2); free(vz1); free(vz2); free(gx); free(gy); free(gi1); free(gi2);
  free(ci1); free(ci2); free(pg1); free(pg2); free(pu1); free(pu2); free(wm1); free(wm2);
}

/* smag, mod_int.f90:403-440 */
static void smag(orc *o, const double *u, const double *v) {
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k <= NK; ++k)
    for (int j = 2; j <= NJ - 1; ++j)
      for (int i = 2; i <= NI - 1; ++i) {
        double dx = ARAD * (o->rh[4][I2(i, j)] * o->dx1[4][I2(i, j)] +
                            o->rh[4][I2(i + 1, j)] * o->dx1[4][I2(i + 1, j)]);
        double dy = ARAD * (o->dx2[2][I2(i, j)] + o->dx2[2][I2(i, j + 1)]);
        double tn = o->tn[3][I2(i, j)];
        double dt = 4 * (u[I3(i + 1, j, k)] - u[I3(i - 1, j, k)]) / dx -
                    4 * (v[I3(i, j + 1, k)] - v[I3(i, j - 1, k)]) / dy +
                    v[I3(i, j, k)] * tn / ARAD;
        double ds = 4 * (v[I3(i + 1, j, k)] - v[I3(i - 1, j, k)]) / dx +
                    4 * (u[I3(i, j + 1, k)] - u[I3(i, j - 1, k)]) / dy -
                    u[I3(i, j, k)] * tn / ARAD;
        double cc = 0.5 * SMAG_C / PI_F * (0.5 * dx + 0.5 * dy);
        double c = cc * cc;
        double wk = c * sqrt(dt * dt + ds * ds);
        double am_max = (dx * dx + dy * dy) / (double)(4 * 8 * o->dtbc);
        o->am[I3(i, j, k)] = (am_max < wk) ? am_max : wk;
      }
  swp(o, o->am, NK);
}

/* int_clin, mod_int.f90:298-401 */
void orc_int_clin(orc *o) {
  size_t n2 = o->n2, n3 = o->n3;
  double c_tc = (double)(23.0f / 12.0f), c_tp = (double)(-(16.0f / 12.0f)), c_tpp = (double)(5.0f / 12.0f);
  double bc = (double)o->dtbc;
  double *sch = dalloc(n2), *wk = dalloc(n2), *w2 = dalloc(n2), *pg1 = dalloc(n2), *pg2 = dalloc(n2);
  double *wm1 = dalloc(n2), *wm2 = dalloc(n2), *pu1 = dalloc(n2), *pu2 = dalloc(n2);
  double *u1 = dalloc(n3), *u2 = dalloc(n3);
  dfill(sch, n2, 1.0);
  ter(o, wk, o->mch, 1, 1, 3, 1.0);
#pragma omp parallel for schedule(static)
  for (size_t p = 0; p < n2; ++p) if (o->lev_u[p] > 0) sch[p] = sqrt(wk[p]);
  for (int k = 1; k <= NK; ++k) {
    size_t ko = n2 * (size_t)(k - 1);
#pragma omp parallel for schedule(static)
    for (size_t p = 0; p < n2; ++p) wk[p] = o->phih_t[p] + o->mch[p] * o->adp[ko + p];
    gra(o, pg1, pg2, wk, 1, 1, 2, 4);
    gra(o, wm1, wm2, o->mch, 1, 1, 2, 4);
    ter(o, wk, o->alpha + ko, 1, 1, 2, 0.0);
#pragma omp parallel for schedule(static)
    for (size_t p = 0; p < n2; ++p) pg1[p] = pg1[p] + wk[p] * o->pr2[k - 1] * wm1[p];
    ter(o, wk, o->alpha + ko, 1, 1, 4, 0.0);
#pragma omp parallel for schedule(static)
    for (size_t p = 0; p < n2; ++p) pg2[p] = pg2[p] + wk[p] * o->pr2[k - 1] * wm2[p];
    ter(o, wk, pg1, 1, 2, 3, 0.0);
#pragma omp parallel for schedule(static)
    for (size_t p = 0; p < n2; ++p) pu1[p] = -((wk[p] + A0 * o->pax[p]) * sch[p]);
    ter(o, wk, pg2, 1, 4, 3, 0.0);
#pragma omp parallel for schedule(static)
    for (size_t p = 0; p < n2; ++p) pu2[p] = -((wk[p] + A0 * o->pay[p]) * sch[p]);
#pragma omp parallel for schedule(static)
    for (size_t p = 0; p < n2; ++p) {
      size_t q = ko + p;
      pu1[p] = pu1[p] - (c_tc * o->auc[q] + c_tp * o->aup[q] + c_tpp * o->aupp[q]);
      pu2[p] = pu2[p] - (c_tc * o->avc[q] + c_tp * o->avp[q] + c_tpp * o->avpp[q]);
      pu1[p] = pu1[p] + o->fx[q];
      pu2[p] = pu2[p] + o->fy[q];
      pu1[p] = pu1[p] + o->cor[p] * o->vc[q];
      pu2[p] = pu2[p] - o->cor[p] * o->uc[q];
      double x1 = pu1[p], x2 = pu2[p];
      double c = 0.5 * o->cor[p] * bc;
      pu1[p] = x1 + c * x2;
      pu2[p] = x2 - c * x1;
      double d = 1 + c * c;
      pu1[p] = pu1[p] * (double)o->msk_u[q] / d;
      pu2[p] = pu2[p] * (double)o->msk_u[q] / d;
      o->uc[q] = o->up[q] + pu1[p] * bc;
      o->vc[q] = o->vp[q] + pu2[p] * bc;
    }
  }
  vint(o, wm1, o->uc, o->msk_u, o->prh_u);
  vint(o, wm2, o->vc, o->msk_u, o->prh_u);
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < NK; ++k)
    for (size_t p = 0; p < n2; ++p) {
      size_t q = p + n2 * k;
      if (o->msk_u[q] > 0) {
        o->uc[q] = o->uc[q] - wm1[p] + o->buc[p];
        o->vc[q] = o->vc[q] - wm2[p] + o->bvc[p];
      }
    }
  copy3(o, o->up, o->uc);
  copy3(o, o->vp, o->vc);
  ter(o, wk, o->mch2, 1, 1, 3, 1.0);
#pragma omp parallel for schedule(static)
  for (size_t p = 0; p < n2; ++p) sch[p] = sqrt(wk[p]);
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < NK; ++k)
    for (size_t p = 0; p < n2; ++p) {
      size_t q = p + n2 * k;
      u1[q] = o->uc[q] * sch[p] / o->mch2[p];
      u2[q] = o->vc[q] * sch[p] / o->mch2[p];
    }
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < n3; ++q) { o->acu[q] = o->acu[q] + u1[q]; o->acv[q] = o->acv[q] + u2[q]; }
  o->n_uv += 1;
  smag(o, u1, u2);
  free(sch); free(wk); free(w2); free(pg1); free(pg2); free(wm1); free(wm2); free(pu1); free(pu2);
  free(u1); free(u2);
}

/* int_ssh, mod_int.f90:602-624 */
void orc_int_ssh(orc *o) {
  den_alpha(o, o->alpha, o->tp, o->sp, o->mch, o->msk_t);
#pragma omp parallel for collapse(2) schedule(static)
  for (int j = 1; j <= NJ; ++j)
    for (int i = 1; i <= NI; ++i) {
      double s = 0.0;
      for (int k = 1; k <= NK; ++k)
        s += o->alpha[I3(i, j, k)] * o->dpr2[k - 1] * o->mch[I2(i, j)] * (double)o->msk_t[I3(i, j, k)];
      o->ssh[I2(i, j)] = s / GRAV;
    }
#pragma omp parallel for schedule(static)
  for (size_t p = 0; p < o->n2; ++p) o->ssh[p] = o->phih_t[p] / GRAV + o->ssh[p];
#pragma omp parallel for schedule(static)
  for (size_t p = 0; p < o->n2; ++p) o->acssh[p] = o->acssh[p] + o->ssh[p];
  o->n_ssh += 1;
}

/* mix, mod_int.f90:583-600 (one column; arrays 1-based via pointer shift) */
static void mix_col(int nk, const double *dpr, const double *p1c, const double *p2c,
                    double *t, double *s, double *rhoa, double *rhob, int na, int nb) {
  double st = 0.0, sd = 0.0, ss = 0.0;
  for (int k = na; k <= nb; ++k) st += t[k] * dpr[k];
  for (int k = na; k <= nb; ++k) sd += dpr[k];
  double tv = st / sd;
  for (int k = na; k <= nb; ++k) ss += s[k] * dpr[k];
  double sv = ss / sd;
  for (int k = na; k <= nb; ++k) t[k] = tv;
  for (int k = na; k <= nb; ++k) s[k] = sv;
  for (int k = na; k <= nb; ++k) rhoa[k] = p1c[k] * t[k] + p2c[k] * s[k];
  for (int k = na; k <= nb - 1; ++k) rhob[k] = p1c[k + 1] * t[k] + p2c[k + 1] * s[k];
  if (nb < nk) rhob[nb] = p1c[nb + 1] * t[nb] + p2c[nb + 1] * s[nb];
}

/* convect + adjust, mod_int.f90:509-581, one column with nlev > 1.
 * All arrays are indexed 1..nk (index 0 unused). */
static void convect_col1(int nk, int nlev, const double *dpr, const double *p1c, const double *p2c,
                         double *t, double *s, double *rhoa, double *rhob) {
  for (int k = 1; k <= nk; ++k) { rhoa[k] = 0.0; rhob[k] = 0.0; }
  for (int k = 2; k <= nlev; ++k) rhoa[k] = p1c[k] * t[k] + p2c[k] * s[k];
  for (int k = 1; k <= nlev - 1; ++k) rhob[k] = p1c[k + 1] * t[k] + p2c[k + 1] * s[k];
  int kcnt = 1;
  while (kcnt < nlev) {
    if (rhob[kcnt] > rhoa[kcnt + 1]) {
      int na = kcnt, nb = kcnt + 1;
      mix_col(nk, dpr, p1c, p2c, t, s, rhoa, rhob, na, nb);
      while (nb < nlev) {
        if (rhob[nb] > rhoa[nb + 1]) {
          nb = nb + 1;
          mix_col(nk, dpr, p1c, p2c, t, s, rhoa, rhob, na, nb);
        } else {
          if (na <= 1 || rhob[na - 1] <= rhoa[na]) break;
          na = na - 1;
          mix_col(nk, dpr, p1c, p2c, t, s, rhoa, rhob, na, nb);
        }
      }
      while (na > 1) {
        if (rhob[na - 1] <= rhoa[na]) break;
        na = na - 1;
        mix_col(nk, dpr, p1c, p2c, t, s, rhoa, rhob, na, nb);
      }
      kcnt = nb;
    } else {
      kcnt = kcnt + 1;
    }
  }
}

void orc_convect_column(int nk, int nlev, const double *dpr, const double *prho_t,
                        const double *prho_s, double *t, double *s) {
  if (nlev <= 1) return;
  double *buf = dalloc(7 * (size_t)(nk + 1));
  double *d = buf, *a1 = d + nk + 1, *a2 = a1 + nk + 1, *tt = a2 + nk + 1, *ss = tt + nk + 1;
  double *ra = ss + nk + 1, *rb = ra + nk + 1;
  for (int k = 1; k <= nk; ++k) {
    d[k] = dpr[k - 1]; a1[k] = prho_t[k - 1]; a2[k] = prho_s[k - 1];
    tt[k] = t[k - 1]; ss[k] = s[k - 1];
  }
  convect_col1(nk, nlev, d, a1, a2, tt, ss, ra, rb);
  for (int k = 1; k <= nk; ++k) { t[k - 1] = tt[k]; s[k - 1] = ss[k]; }
  free(buf);
}

/* convect, mod_int.f90:509-542 */
static void convect(orc *o) {
  int nk = NK;
#pragma omp parallel
  {
  /* columns are independent: each thread works in its own column buffers */
  double *buf = dalloc(7 * (size_t)(nk + 1));
  double *d = buf, *a1 = d + nk + 1, *a2 = a1 + nk + 1, *tt = a2 + nk + 1, *ss = tt + nk + 1;
  double *ra = ss + nk + 1, *rb = ra + nk + 1;
  for (int k = 1; k <= nk; ++k) d[k] = o->dpr2[k - 1];
#pragma omp for collapse(2) schedule(dynamic, 64)
  for (int i = 1; i <= NI; ++i)
    for (int j = 1; j <= NJ; ++j) {
      int nlev = o->lev_t[I2(i, j)];
      if (nlev > 1) {
        for (int k = 1; k <= nk; ++k) {
          a1[k] = o->prho1[I3(i, j, k)]; a2[k] = o->prho2[I3(i, j, k)];
          tt[k] = o->tc[I3(i, j, k)]; ss[k] = o->sc[I3(i, j, k)];
        }
        convect_col1(nk, nlev, d, a1, a2, tt, ss, ra, rb);
        for (int k = 1; k <= nk; ++k) { o->tc[I3(i, j, k)] = tt[k]; o->sc[I3(i, j, k)] = ss[k]; }
      }
    }
  free(buf);
  }
}

/* int_ts, mod_int.f90:442-507.  The density accumulator reads a local
 * array that den_alpha leaves unset on land; land points of acr are left
 * unchanged here. */
void orc_int_ts(orc *o) {
  size_t n2 = o->n2, n3 = o->n3;
  double bc = (double)o->dtbc;
  double *pt1 = dalloc(n3), *pt2 = dalloc(n3), *w1 = dalloc(n3), *w2 = dalloc(n3), *wk = dalloc(n3);
  upwelling(o);
  for (int k = 0; k < NK; ++k)
    for (size_t p = 0; p < n2; ++p) {
      size_t q = p + n2 * k;
      o->acw[q] = o->acw[q] + o->wm[q] * o->alpha[q] / (GRAV * o->mch2[p]);
    }
  o->n_w += 1;
  op_adv_ts(o, w1, o->tc, o->wm);
  op_adv_ts(o, w2, o->sc, o->wm);
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < n3; ++q) { pt1[q] = -w1[q]; pt2[q] = -w2[q]; }
  op_dif(o, w1, o->mch, o->tp);
  op_dif(o, w2, o->mch, o->sp);
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < n3; ++q) { pt1[q] = pt1[q] + w1[q]; pt2[q] = pt2[q] + w2[q]; }
#pragma omp parallel for schedule(static)
  for (size_t p = 0; p < n2; ++p) {
    double pb = GAMMA_T * (o->bnd[2][p] - o->tc[p]) * (double)o->msk_t[p] / o->dpr2[0];
    pt1[p] = pt1[p] + pb;
    pb = GAMMA_S * (o->bnd[3][p] - o->sc[p]) * (double)o->msk_t[p] * o->mch[p];
    pt2[p] = pt2[p] + pb;
  }
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < NK; ++k)
    for (size_t p = 0; p < n2; ++p) {
      size_t q = p + n2 * k;
      o->tc[q] = o->tp[q] + pt1[q] * bc * (double)o->msk_t[q] / o->mch[p];
      o->sc[q] = o->sp[q] + pt2[q] * bc * (double)o->msk_t[q] / o->mch[p];
    }
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < n3; ++q)
    if (o->msk_t[q] > 0 && o->tc[q] < TICE) o->tc[q] = TICE;
  convect(o);
  copy3(o, o->tp, o->tc);
  copy3(o, o->sp, o->sc);
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < n3; ++q) { o->act[q] = o->act[q] + o->tc[q]; o->acs[q] = o->acs[q] + o->sc[q]; }
  den_alpha(o, wk, o->tp, o->sp, o->chc, o->msk_t);
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < n3; ++q)
    if (o->msk_t[q] == 1) o->acr[q] = o->acr[q] + 1.0 / wk[q];
  o->n_ts += 1;
  free(pt1); free(pt2); free(w1); free(w2); free(wk);
}

/* ---- calendar, mod_type.f90:403-438, 593-635 --------------------------- */
static int days_year(int y) {
  int d = 365;
  if (y % 100 == 0) { if (y % 400 == 0) d = 366; }
  else if (y % 4 == 0) d = 366;
  return d;
}
static int days_month(int y, int m) {
  static const int dm[12] = {31, 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31};
  int d = dm[m - 1];
  if (m == 2 && days_year(y) == 366) d = 29;
  return d;
}
static otime time_plus(otime t, int dt) {
  t.s += dt;
  if (t.s >= 60) { t.mi += t.s / 60; t.s %= 60; }
  if (t.mi >= 60) { t.h += t.mi / 60; t.mi %= 60; }
  if (t.h >= 24) { t.d += t.h / 24; t.h %= 24; }
  if (t.d > t.mm) { t.m += 1; t.d = 1; }
  if (t.m > 12) { t.y += 1; t.m = 1; }
  t.mm = days_month(t.y, t.m);
  return t;
}

/* int_bnd, mod_int.f90:649-691.  The weight f is evaluated in default
 * REAL (float32) arithmetic, as the source's literals make it. */
void orc_int_bnd(orc *o) {
  int y = o->ct.y, m = o->ct.m, d = o->ct.d, m1, m2, days;
  float f;
  if (d < 15) {
    m2 = m; m1 = m - 1; if (m == 1) m1 = 12;
    days = days_month(y, m1);
    f = ((float)days - 15.0f + (float)d) / ((float)days - 15.0f + 15.0f);
  } else {
    m1 = m; m2 = m + 1; if (m == 12) m2 = 1;
    days = days_month(y, m1);
    f = ((float)d - 15.0f) / ((float)(days - 15) + 15.0f);
  }
  double fd = (double)f;
  for (int v = 0; v < 6; ++v)
#pragma omp parallel for schedule(static)
    for (size_t p = 0; p < o->n2; ++p) {
      double a1 = o->frc[v][p + o->n2 * (size_t)(m1 - 1)], a2 = o->frc[v][p + o->n2 * (size_t)(m2 - 1)];
      o->bnd[v][p] = (a2 - a1) * fd + a1;
    }
}

void orc_advance_time(orc *o) {
  o->pt = o->ct;
  o->ct = time_plus(o->ct, o->dtbc);
}

void orc_set_step_index(orc *o, long long i) { o->step_i = i; }

/* main.f90:164-202 with forcing files: the records are read after int_bnd
 * (daily) or on their own (6-hourly), divided into the halo'd fields as
 * mympi_input does (mod_mympi.f90:638-657, 898-928: interior, then the E/W
 * swap; wall rows keep their values), then den_prho. */
static void put_record(orc *o, double *bnd, const double *rec) {
  for (int j = 2; j <= NJ - 1; ++j)
    for (int i = 2; i <= NI - 1; ++i)
      bnd[I2(i, j)] = rec[(size_t)(i - 2) + (size_t)o->glo_ni * (size_t)(j - 2)];
  swp(o, bnd, 1);
}

void orc_forcing_due(const orc *o, int due[2]) {
  int first = o->step_i + 1 == 1;
  due[0] = first || o->ct.d != o->pt.d;
  due[1] = first || (o->ct.h != o->pt.h && o->ct.h % 6 == 0);
}

int orc_set_forcing(orc *o, const pcom_forcing *f, int interpolate) {
  long long next = o->step_i + 1;
  int daily_due = next == 1 || o->ct.d != o->pt.d;
  if (!interpolate && daily_due && o->forcing_taken != next) return -1;
  if (interpolate) orc_int_bnd(o);
  if (f->taux) put_record(o, o->bnd[0], f->taux);
  if (f->tauy) put_record(o, o->bnd[1], f->tauy);
  if (f->bct) put_record(o, o->bnd[2], f->bct);
  if (f->pa) put_record(o, o->bnd[4], f->pa);
  orc_den_prho(o);
  if (interpolate) o->forcing_taken = next;
  return 0;
}

/* main.f90:128-141: a restart run continues from tctr%ist = step + 1 with
 * tctr%ct and tctr%pt from the restart file. */
void orc_set_clock(orc *o, long long step, const int c[6], const int p[6]) {
  o->nsteps = step;
  o->step_i = step;
  o->forcing_taken = 0;
  otime *t[2] = {&o->ct, &o->pt};
  const int *d[2] = {c, p};
  for (int w = 0; w < 2; ++w) {
    t[w]->y = d[w][0]; t[w]->m = d[w][1]; t[w]->d = d[w][2];
    t[w]->h = d[w][3]; t[w]->mi = d[w][4]; t[w]->s = d[w][5];
    t[w]->mm = days_month(d[w][0], d[w][1]);
  }
}

/* ---- output(): main.f90:726-760 with mod_mympi.f90:178-343 ----------- */
/* interior of a halo'd array into a (glo_ni, glo_nj, nd3) array: the
 * merge_out_* gather of one rank, var(2:ni-1, 2:nj-1, :) */
static void to_glo(const orc *o, const double *a, int nd3, double *out) {
  size_t q = 0;
  for (int k = 1; k <= nd3; ++k)
    for (int j = 2; j <= NJ - 1; ++j)
      for (int i = 2; i <= NI - 1; ++i) out[q++] = a[I3(i, j, k)];
}

int orc_output_means(orc *o, pcom_means *m) {
  const double missing = (double)PCOM_MISSING_FLOAT;
  const size_t n3 = o->n3, n2 = o->n2;
  if (((m->pt || m->sa || m->rho) && o->n_ts == 0) || ((m->u || m->v) && o->n_uv == 0) ||
      (m->w && o->n_w == 0) || (m->ch && o->n_ch == 0) || (m->ssh && o->n_ssh == 0))
    return -1;
  /* merge_out_eqts (178-199) */
  if (o->n_ts) {
    for (size_t p = 0; p < n3; ++p) {
      o->act[p] = o->act[p] / o->n_ts;
      o->acs[p] = o->acs[p] / o->n_ts;
      o->acr[p] = o->acr[p] / o->n_ts;
      if (o->msk_t[p] == 0) o->acr[p] = missing;
    }
    if (m->pt) to_glo(o, o->act, NK, m->pt);
    if (m->sa) to_glo(o, o->acs, NK, m->sa);
    if (m->rho) to_glo(o, o->acr, NK, m->rho);
  }
  dfill(o->act, n3, 0.0); dfill(o->acs, n3, 0.0); dfill(o->acr, n3, 0.0);
  o->n_ts = 0;
  /* merge_out_equvw (201-226): all three masked by the U mask */
  for (size_t p = 0; p < n3; ++p) {
    if (o->n_uv) {
      o->acu[p] = o->acu[p] / o->n_uv;
      o->acv[p] = o->acv[p] / o->n_uv;
    }
    if (o->n_w) o->acw[p] = o->acw[p] / o->n_w;
    if (o->msk_u[p] == 0) {
      o->acu[p] = missing;
      o->acv[p] = missing;
      o->acw[p] = missing;
    }
  }
  if (m->u) to_glo(o, o->acu, NK, m->u);
  if (m->v) to_glo(o, o->acv, NK, m->v);
  if (m->w) to_glo(o, o->acw, NK, m->w);
  dfill(o->acu, n3, 0.0); dfill(o->acv, n3, 0.0); dfill(o->acw, n3, 0.0);
  o->n_uv = 0;
  o->n_w = 0;
  /* eqch%acch (main.f90:743-749) */
  if (o->n_ch) {
    for (size_t p = 0; p < n2; ++p) {
      o->acch[p] = o->acch[p] / o->n_ch;
      if (o->msk_t[p] == 0) o->acch[p] = missing;
    }
    if (m->ch) to_glo(o, o->acch, 1, m->ch);
  }
  dfill(o->acch, n2, 0.0);
  o->n_ch = 0;
  /* merge_out_accu_gr2d_weight + merge_out_r2d_weight (251-306): acssh/n,
   * masked, minus sum(var*wg)/sum(wg) with wg = rh(g1)*msk(:,:,1) */
  if (o->n_ssh) {
    double *wg = dalloc(n2);
    for (size_t p = 0; p < n2; ++p) {
      o->acssh[p] = o->acssh[p] / o->n_ssh;
      if (o->msk_t[p] == 0) o->acssh[p] = missing;
      wg[p] = o->rh[1][p] * (double)o->msk_t[p];
    }
    const size_t g2 = (size_t)o->glo_ni * o->glo_nj;
    double *gv = dalloc(g2), *gw = dalloc(g2);
    to_glo(o, o->acssh, 1, gv);
    to_glo(o, wg, 1, gw);
    double s1 = 0.0, s2 = 0.0;
    for (size_t q = 0; q < g2; ++q) s1 += gv[q] * gw[q];
    for (size_t q = 0; q < g2; ++q) s2 += gw[q];
    const double mean = s1 / s2;
    for (size_t q = 0; q < g2; ++q)
      if (gv[q] != missing) gv[q] = gv[q] - mean;
    if (m->ssh) memcpy(m->ssh, gv, g2 * sizeof(double));
    free(wg); free(gv); free(gw);
  }
  dfill(o->acssh, n2, 0.0);
  o->n_ssh = 0;
  return 0;
}

/* main.f90:158-227 */
void orc_step(orc *o, int n) {
  for (int s = 0; s < n; ++s) {
    o->nsteps += 1;
    o->step_i = o->nsteps;
    if ((o->step_i == 1 || o->ct.d != o->pt.d) && o->forcing_taken != o->step_i) {
      orc_int_bnd(o);
      orc_den_prho(o);
    }
    orc_den_alpha_main(o);
    orc_int_pgra(o);
    orc_int_readyc(o);
    orc_int_trop(o);
    orc_int_clin(o);
    orc_int_ssh(o);
    orc_int_ts(o);
    orc_advance_time(o);
  }
}

/* ---- start-up chain, main.f90:252-704 ---------------------------------- */

/* mympi_div of a global (glo_ni, glo_nj, nd3) array into the interior,
 * then the E/W swap (mod_mympi.f90:863-896) */
static void divide_r(orc *o, double *dst, const double *src, int nd3) {
  for (int k = 1; k <= nd3; ++k)
    for (int j = 2; j <= NJ - 1; ++j)
      for (int i = 2; i <= NI - 1; ++i)
        dst[I3(i, j, k)] = src[(size_t)(i - 2) + (size_t)o->glo_ni * ((size_t)(j - 2) + (size_t)o->glo_nj * (size_t)(k - 1))];
  swp(o, dst, nd3);
}

static void setup(orc *o, const pcom_inputs *in) {
  int ni = NI, nj = NJ, nk = NK, nkp = o->nkp;
  size_t n2 = o->n2;
  /* set_cf, main.f90:312-345; km(:,:,12) is skipped when nk+1 < 12 */
  {
    static const float pf0[12] = {0.1f, 50.0f, 50.0f, 50.0f, 40.0f, 20.0f, 10.0f, 5.0f, 2.0f, 1.0f, 0.5f, 0.2f};
    dfill(o->km, o->n3p, KM_C);
    for (int k = 1; k <= 12 && k <= nkp; ++k) {
      double pf = (double)pf0[k - 1] * (double)1.0e-4f;
      for (size_t p = 0; p < n2; ++p) o->km[p + n2 * (k - 1)] = pf;
    }
    size_t gn = (size_t)o->glo_ni * o->glo_nj;
    double *gk = dalloc(gn * nkp);
    for (int k = 2; k <= nkp; ++k)
      for (size_t p = 0; p < gn; ++p) {
        double v = in->vmix ? in->vmix[p + gn * (k - 2)] : 9.99999e30;
        gk[p + gn * (k - 1)] = v;
      }
    for (size_t q = 0; q < gn * nkp; ++q) {
      if (gk[q] > (double)(9.0e+30f - 1.0f)) gk[q] = (double)0.1f;
      if (gk[q] > 200.0) gk[q] = 200.0;
      gk[q] = gk[q] * (double)1.0e-4f;
    }
    divide_r(o, o->kh, gk, nkp);
    free(gk);
  }
  /* init(): lon, lat (main.f90:292-308) */
  for (int i = 2; i <= ni - 1; ++i) o->lon[i - 1] = in->lon[i - 2];
  o->lon[0] = o->lon[ni - 2];
  o->lon[ni - 1] = o->lon[1];
  for (int j = 2; j <= nj - 1; ++j) o->lat[j - 1] = in->lat[j - 2];
  o->lat[0] = 2 * o->lat[1] - o->lat[2];
  o->lat[nj - 1] = 2 * o->lat[nj - 2] - o->lat[nj - 3];
  /* calc_hpos, main.f90:347-406 */
  double *hx = o->hposx, *hy = o->hposy;
  for (int j = 2; j <= o->njm; ++j)
    for (int i = 2; i <= o->nim; ++i) {
      const double *lon = o->lon, *lat = o->lat;
      double loni = lon[i - 1], latj = (lat[j - 1] + lat[j - 2]) * 0.5;
      hx[I3(i, j, 1)] = loni * TORAD; hy[I3(i, j, 1)] = latj * TORAD;
      loni = (lon[i - 1] + lon[i]) * 0.5;
      if (lon[i] < lon[i - 1]) loni = (lon[i - 1] + lon[i] + 360) * 0.5;
      latj = (lat[j - 1] + lat[j - 2]) * 0.5;
      hx[I3(i, j, 2)] = loni * TORAD; hy[I3(i, j, 2)] = latj * TORAD;
      loni = (lon[i - 1] + lon[i]) * 0.5;
      if (lon[i] < lon[i - 1]) loni = (lon[i - 1] + lon[i] + 360) * 0.5;
      latj = lat[j - 1];
      hx[I3(i, j, 3)] = loni * TORAD; hy[I3(i, j, 3)] = latj * TORAD;
      loni = lon[i - 1]; latj = lat[j - 1];
      hx[I3(i, j, 4)] = loni * TORAD; hy[I3(i, j, 4)] = latj * TORAD;
    }
  swp(o, hx, 4);
  swp(o, hy, 4);
  for (int g = 1; g <= 4; ++g)
    for (int i = 1; i <= ni; ++i) {
      hx[I3(i, 1, g)] = 2 * hx[I3(i, 2, g)] - hx[I3(i, 3, g)];
      hy[I3(i, 1, g)] = 2 * hy[I3(i, 2, g)] - hy[I3(i, 3, g)];
    }
  for (int g = 1; g <= 4; ++g)
    for (int i = 1; i <= ni; ++i) {
      hx[I3(i, nj, g)] = 2 * hx[I3(i, nj - 1, g)] - hx[I3(i, nj - 2, g)];
      hy[I3(i, nj, g)] = 2 * hy[I3(i, nj - 1, g)] - hy[I3(i, nj - 2, g)];
    }
  /* set_hgrid, main.f90:408-497 */
  for (int j = 2; j <= o->njm; ++j)
    for (int i = 2; i <= o->nim; ++i) {
      double d1, d2;
      d1 = hx[I3(i, j, 2)] - hx[I3(i - 1, j, 2)]; if (d1 < 0) d1 = d1 + 360.0 * TORAD;
      d2 = hy[I3(i, j, 4)] - hy[I3(i, j - 1, 4)];
      o->dx1[1][I2(i, j)] = d1; o->dx2[1][I2(i, j)] = d2;
      d1 = hx[I3(i + 1, j, 1)] - hx[I3(i, j, 1)]; if (d1 < 0) d1 = d1 + 360.0 * TORAD;
      d2 = hy[I3(i, j, 3)] - hy[I3(i, j - 1, 3)];
      o->dx1[2][I2(i, j)] = d1; o->dx2[2][I2(i, j)] = d2;
      d1 = hx[I3(i + 1, j, 4)] - hx[I3(i, j, 4)]; if (d1 < 0) d1 = d1 + 360.0 * TORAD;
      d2 = hy[I3(i, j + 1, 2)] - hy[I3(i, j, 2)];
      o->dx1[3][I2(i, j)] = d1; o->dx2[3][I2(i, j)] = d2;
      d1 = hx[I3(i, j, 3)] - hx[I3(i - 1, j, 3)]; if (d1 < 0) d1 = d1 + 360.0 * TORAD;
      d2 = hy[I3(i, j + 1, 1)] - hy[I3(i, j, 1)];
      o->dx1[4][I2(i, j)] = d1; o->dx2[4][I2(i, j)] = d2;
    }
  for (int g = 1; g <= 4; ++g) {
    /* swpbnd_m2d: x(1) then x(2) */
    swp(o, o->dx1[g], 1);
    swp(o, o->dx2[g], 1);
  }
  for (int g = 1; g <= 4; ++g)
    for (int i = 1; i <= ni; ++i) {
      o->dx1[g][I2(i, 1)] = 2 * o->dx1[g][I2(i, 2)] - o->dx1[g][I2(i, 3)];
      o->dx2[g][I2(i, 1)] = 2 * o->dx2[g][I2(i, 2)] - o->dx2[g][I2(i, 3)];
    }
  for (int g = 1; g <= 4; ++g)
    for (int i = 1; i <= ni; ++i) {
      o->dx1[g][I2(i, nj)] = 2 * o->dx1[g][I2(i, nj - 1)] - o->dx1[g][I2(i, nj - 2)];
      o->dx2[g][I2(i, nj)] = 2 * o->dx2[g][I2(i, nj - 1)] - o->dx2[g][I2(i, nj - 2)];
    }
  for (int g = 1; g <= 4; ++g)
    for (int j = 1; j <= nj; ++j)
      for (int i = 1; i <= ni; ++i) {
        o->rh[g][I2(i, j)] = cos(hy[I3(i, j, g)]);
        o->tn[g][I2(i, j)] = tan(hy[I3(i, j, g)]);
      }
  /* set_vgrid, main.f90:499-583 */
  {
    double *dz = dalloc(nk), *rmn = dalloc(nk), *pmn = dalloc(nk);
    const double *z = in->z, *tmn = in->tmn, *smn = in->smn;
    dz[0] = 2 * z[0];
    for (int k = 2; k <= nk; ++k) dz[k - 1] = 2 * (z[k - 1] - z[k - 2]) - dz[k - 2];
    double pmnbot = 1.0;
    for (int k = 0; k < nk; ++k) { pmn[k] = 0.0; rmn[k] = RHO0; }
    while (fabs(pmn[nk - 1] - pmnbot) > (double)1e-6f) {
      pmnbot = pmn[nk - 1];
      rmn[0] = orc_den_rho(tmn[0], smn[0], pmn[0] * 0.5);
      for (int k = 2; k <= nk; ++k)
        rmn[k - 1] = orc_den_rho(tmn[k - 1], smn[k - 1], (pmn[k - 2] + pmn[k - 1]) * 0.5);
      pmn[0] = rmn[0] * GRAV * dz[0];
      for (int k = 2; k <= nk; ++k) pmn[k - 1] = rmn[k - 1] * GRAV * dz[k - 1] + pmn[k - 2];
    }
    o->z1[0] = 0.0;
    for (int k = 2; k <= nkp; ++k) o->z1[k - 1] = o->z1[k - 2] + dz[k - 2];
    for (int k = 0; k < nk; ++k) o->z2[k] = z[k];
    o->pr1[0] = 0.0;
    for (int k = 2; k <= nkp; ++k) o->pr1[k - 1] = pmn[k - 2];
    o->pr2[0] = pmn[0] * 0.5;
    for (int k = 2; k <= nk; ++k) o->pr2[k - 1] = (pmn[k - 1] + pmn[k - 2]) * 0.5;
    for (int k = 0; k < nkp; ++k) o->dz1[k] = 0.0;
    o->dz1[0] = 0.5 * dz[0];
    for (int k = 2; k <= nk; ++k) o->dz1[k - 1] = z[k - 1] - z[k - 2];
    o->dz1[nkp - 1] = o->dz1[nk - 1];
    for (int k = 0; k < nk; ++k) o->dz2[k] = dz[k];
    o->dpr1[0] = 0.5 * pmn[0];
    for (int k = 2; k <= nk; ++k) o->dpr1[k - 1] = o->pr2[k - 1] - o->pr2[k - 2];
    o->dpr1[nkp - 1] = o->dpr1[nk - 1];
    o->dpr2[0] = pmn[0];
    for (int k = 2; k <= nk; ++k) o->dpr2[k - 1] = pmn[k - 1] - pmn[k - 2];
    free(dz); free(rmn); free(pmn);
  }
  /* set_grid, main.f90:585-659 */
  {
    for (int j = 2; j <= nj - 1; ++j)
      for (int i = 2; i <= ni - 1; ++i)
        o->lev_t[I2(i, j)] = in->itn[(size_t)(i - 2) + (size_t)o->glo_ni * (size_t)(j - 2)];
    swpi(o, o->lev_t, 1);
    for (int k = 1; k <= nk; ++k)
      for (int j = 2; j <= o->njm; ++j)
        for (int i = 2; i <= o->nim; ++i)
          if (o->lev_t[I2(i, j)] >= k) o->msk_t[I3(i, j, k)] = 1;
    swpi(o, o->msk_t, nk);
    for (int k = 1; k <= nk; ++k)
      for (int j = 1; j <= o->njm; ++j)
        for (int i = 1; i <= o->nim; ++i)
          o->msk_u[I3(i, j, k)] = o->msk_t[I3(i, j, k)] * o->msk_t[I3(i + 1, j, k)] *
                                  o->msk_t[I3(i, j + 1, k)] * o->msk_t[I3(i + 1, j + 1, k)];
    swpi(o, o->msk_u, nk);
    for (size_t p = 0; p < n2; ++p) {
      int s = 0;
      for (int k = 0; k < nk; ++k) s += o->msk_u[p + n2 * k];
      o->lev_u[p] = s;
    }
    double *phimn = dalloc(nk);
    phimn[0] = -(GRAV * o->dz2[0]);
    for (int k = 2; k <= nk; ++k) phimn[k - 1] = phimn[k - 2] - GRAV * o->dz2[k - 1];
    for (size_t p = 0; p < n2; ++p) {
      int k = o->lev_t[p];
      if (k == 0) o->phih_t[p] = 0.0;
      else { o->phih_t[p] = phimn[k - 1]; o->prh_t[p] = o->pr1[k]; }
      k = o->lev_u[p];
      if (k == 0) o->phih_u[p] = 0.0;
      else { o->phih_u[p] = phimn[k - 1]; o->prh_u[p] = o->pr1[k]; }
    }
    for (size_t p = 0; p < n2; ++p) {
      if (o->lev_t[p] == 0) o->prh_t[p] = (double)0.1f;
      if (o->lev_u[p] == 0) o->prh_u[p] = (double)0.1f;
    }
    free(phimn);
  }
  /* calc_cf, main.f90:661-673 */
  for (size_t p = 0; p < n2; ++p) {
    double c = cos(hy[p + n2 * 2]);
    double t = sqrt(1 - c * c) / cos(hy[p + n2 * 2]);
    o->cv1[p] = (1 - t * t) / (ARAD * ARAD);
    o->cv2[p] = 2 * t / ARAD;
  }
  /* Coriolis, main.f90:119 */
  for (size_t p = 0; p < n2; ++p) o->cor[p] = 2.0 * OMEGA * sin(hy[p + n2 * 2]);
  /* inistat, main.f90:675-704 */
  {
    const double *src[6] = {in->taux, in->tauy, in->bct, in->bcs, in->pa, in->fw};
    for (int v = 0; v < 6; ++v) divide_r(o, o->frc[v], src[v], 12);
    const double missing = (double)(-9.99999e30f);
    for (size_t q = 0; q < n2 * 12; ++q)
      if (o->frc[0][q] == missing)
        for (int v = 0; v < 6; ++v) o->frc[v][q] = 0.0;
    divide_r(o, o->tc, in->pt, nk);
    divide_r(o, o->sc, in->sa, nk);
    copy3(o, o->tp, o->tc);
    copy3(o, o->sp, o->sc);
    double *wkx = dalloc(n2), *wky = dalloc(n2);
    gra(o, wkx, wky, o->phih_t, 1, 1, 2, 4);
    ter(o, o->graphihx, wkx, 1, 2, 3, 0.0);
    ter(o, o->graphihy, wky, 1, 4, 3, 0.0);
    free(wkx); free(wky);
  }
}

orc *orc_create(const pcom_inputs *in, const pcom_opts *opts) {
  if (!in || !opts || in->glo_ni < 3 || in->glo_nj < 3 || in->nk < 2) return NULL;
  orc *o = (orc *)calloc(1, sizeof(orc));
  o->glo_ni = in->glo_ni; o->glo_nj = in->glo_nj; o->nk = in->nk;
  o->ni = in->glo_ni + 2; o->nj = in->glo_nj + 2; o->nkp = in->nk + 1;
  o->nim = o->ni - 1; o->njm = o->nj - 1;
  o->n2 = (size_t)o->ni * o->nj; o->n3 = o->n2 * o->nk; o->n3p = o->n2 * o->nkp;
  o->dtbt = opts->dtbt; o->dtbc = opts->dtbc; o->rst = opts->restart_run;
  size_t n2 = o->n2, n3 = o->n3, n3p = o->n3p;
  for (int g = 1; g <= 4; ++g) {
    o->rh[g] = dalloc(n2); o->tn[g] = dalloc(n2); o->dx1[g] = dalloc(n2); o->dx2[g] = dalloc(n2);
  }
  o->hposx = dalloc(n2 * 4); o->hposy = dalloc(n2 * 4);
  o->lon = dalloc(o->ni); o->lat = dalloc(o->nj);
  o->z1 = dalloc(o->nkp); o->pr1 = dalloc(o->nkp); o->dz1 = dalloc(o->nkp); o->dpr1 = dalloc(o->nkp);
  o->z2 = dalloc(o->nk); o->pr2 = dalloc(o->nk); o->dz2 = dalloc(o->nk); o->dpr2 = dalloc(o->nk);
  o->msk_t = ialloc(n3); o->msk_u = ialloc(n3); o->lev_t = ialloc(n2); o->lev_u = ialloc(n2);
  o->phih_t = dalloc(n2); o->prh_t = dalloc(n2); o->phih_u = dalloc(n2); o->prh_u = dalloc(n2);
  o->cv1 = dalloc(n2); o->cv2 = dalloc(n2); o->cor = dalloc(n2);
  o->graphihx = dalloc(n2); o->graphihy = dalloc(n2);
  o->km = dalloc(n3p); o->kh = dalloc(n3p);
  for (int v = 0; v < 6; ++v) { o->frc[v] = dalloc(n2 * 12); o->bnd[v] = dalloc(n2); }
  o->tc = dalloc(n3); o->sc = dalloc(n3); o->tp = dalloc(n3); o->sp = dalloc(n3);
  o->uc = dalloc(n3); o->vc = dalloc(n3); o->up = dalloc(n3); o->vp = dalloc(n3);
  o->auc = dalloc(n3); o->avc = dalloc(n3); o->aup = dalloc(n3); o->avp = dalloc(n3);
  o->aupp = dalloc(n3); o->avpp = dalloc(n3);
  o->fx = dalloc(n3); o->fy = dalloc(n3); o->am = dalloc(n3);
  o->alpha = dalloc(n3); o->adp = dalloc(n3); o->wm = dalloc(n3p);
  o->prho1 = dalloc(n3); o->prho2 = dalloc(n3);
  o->act = dalloc(n3); o->acs = dalloc(n3); o->acr = dalloc(n3);
  o->acu = dalloc(n3); o->acv = dalloc(n3); o->acw = dalloc(n3);
  o->chc = dalloc(n2); o->chp = dalloc(n2); o->mch = dalloc(n2); o->mch2 = dalloc(n2);
  o->acch = dalloc(n2); o->acssh = dalloc(n2); o->ssh = dalloc(n2);
  o->buc = dalloc(n2); o->bvc = dalloc(n2); o->bup = dalloc(n2); o->bvp = dalloc(n2);
  o->but = dalloc(n2); o->bvt = dalloc(n2); o->pax = dalloc(n2); o->pay = dalloc(n2);
  for (int q = 0; q < 4; ++q) { o->bphi[q] = dalloc(n2); o->bgra[q] = dalloc(n2); }
  /* initial values, mod_arrays.f90:257-313 */
  dfill(o->alpha, n3, (double)(1.0f / 1000.0f));
  dfill(o->am, n3, AM_C);
  dfill(o->chc, n2, 1.0); dfill(o->chp, n2, 1.0); dfill(o->mch, n2, 1.0); dfill(o->mch2, n2, 1.0);
  setup(o, in);
  o->ct.y = opts->year; o->ct.m = opts->month; o->ct.d = opts->day; o->ct.h = opts->hour;
  o->ct.mi = 0; o->ct.s = 0; o->ct.mm = days_month(o->ct.y, o->ct.m);
  o->pt = o->ct;
  return o;
}

void orc_destroy(orc *o) {
  if (!o) return;
  for (int g = 1; g <= 4; ++g) { free(o->rh[g]); free(o->tn[g]); free(o->dx1[g]); free(o->dx2[g]); }
  double *ptrs[] = {o->hposx, o->hposy, o->lon, o->lat, o->z1, o->pr1, o->dz1, o->dpr1, o->z2, o->pr2,
                    o->dz2, o->dpr2, o->phih_t, o->prh_t, o->phih_u, o->prh_u, o->cv1, o->cv2, o->cor,
                    o->graphihx, o->graphihy, o->km, o->kh, o->tc, o->sc, o->tp, o->sp, o->uc, o->vc,
                    o->up, o->vp, o->auc, o->avc, o->aup, o->avp, o->aupp, o->avpp, o->fx, o->fy, o->am,
                    o->alpha, o->adp, o->wm, o->prho1, o->prho2, o->act, o->acs, o->acr, o->acu, o->acv,
                    o->acw, o->chc, o->chp, o->mch, o->mch2, o->acch, o->acssh, o->ssh, o->buc, o->bvc,
                    o->bup, o->bvp, o->but, o->bvt, o->pax, o->pay};
  for (size_t q = 0; q < sizeof(ptrs) / sizeof(ptrs[0]); ++q) free(ptrs[q]);
  for (int v = 0; v < 6; ++v) { free(o->frc[v]); free(o->bnd[v]); }
  for (int q = 0; q < 4; ++q) { free(o->bphi[q]); free(o->bgra[q]); }
  free(o->msk_t); free(o->msk_u); free(o->lev_t); free(o->lev_u);
  free(o);
}

/* field table: pointer and level count (1 = 2-D) */
static double *field_ptr(const orc *o, int id, int *nd3) {
  int nk = o->nk, nkp = o->nkp;
  *nd3 = 1;
  switch (id) {
    case PCOM_F_TC: *nd3 = nk; return o->tc;
    case PCOM_F_SC: *nd3 = nk; return o->sc;
    case PCOM_F_TP: *nd3 = nk; return o->tp;
    case PCOM_F_SP: *nd3 = nk; return o->sp;
    case PCOM_F_UC: *nd3 = nk; return o->uc;
    case PCOM_F_VC: *nd3 = nk; return o->vc;
    case PCOM_F_UP: *nd3 = nk; return o->up;
    case PCOM_F_VP: *nd3 = nk; return o->vp;
    case PCOM_F_AUC: *nd3 = nk; return o->auc;
    case PCOM_F_AVC: *nd3 = nk; return o->avc;
    case PCOM_F_AUP: *nd3 = nk; return o->aup;
    case PCOM_F_AVP: *nd3 = nk; return o->avp;
    case PCOM_F_AUPP: *nd3 = nk; return o->aupp;
    case PCOM_F_AVPP: *nd3 = nk; return o->avpp;
    case PCOM_F_FX: *nd3 = nk; return o->fx;
    case PCOM_F_FY: *nd3 = nk; return o->fy;
    case PCOM_F_AM: *nd3 = nk; return o->am;
    case PCOM_F_ALPHA: *nd3 = nk; return o->alpha;
    case PCOM_F_ADP: *nd3 = nk; return o->adp;
    case PCOM_F_WM: *nd3 = nkp; return o->wm;
    case PCOM_F_PRHO_T: *nd3 = nk; return o->prho1;
    case PCOM_F_PRHO_S: *nd3 = nk; return o->prho2;
    case PCOM_F_ACT: *nd3 = nk; return o->act;
    case PCOM_F_ACS: *nd3 = nk; return o->acs;
    case PCOM_F_ACR: *nd3 = nk; return o->acr;
    case PCOM_F_ACU: *nd3 = nk; return o->acu;
    case PCOM_F_ACV: *nd3 = nk; return o->acv;
    case PCOM_F_ACW: *nd3 = nk; return o->acw;
    case PCOM_F_CHC: return o->chc;
    case PCOM_F_CHP: return o->chp;
    case PCOM_F_MCH: return o->mch;
    case PCOM_F_MCH2: return o->mch2;
    case PCOM_F_ACCH: return o->acch;
    case PCOM_F_ACSSH: return o->acssh;
    case PCOM_F_BUC: return o->buc;
    case PCOM_F_BVC: return o->bvc;
    case PCOM_F_BUP: return o->bup;
    case PCOM_F_BVP: return o->bvp;
    case PCOM_F_BUT: return o->but;
    case PCOM_F_BVT: return o->bvt;
    case PCOM_F_PAX: return o->pax;
    case PCOM_F_PAY: return o->pay;
    case PCOM_F_BPHI_XN: return o->bphi[0];
    case PCOM_F_BPHI_XS: return o->bphi[1];
    case PCOM_F_BPHI_YE: return o->bphi[2];
    case PCOM_F_BPHI_YW: return o->bphi[3];
    case PCOM_F_BGRA_XN: return o->bgra[0];
    case PCOM_F_BGRA_XS: return o->bgra[1];
    case PCOM_F_BGRA_YE: return o->bgra[2];
    case PCOM_F_BGRA_YW: return o->bgra[3];
    case PCOM_F_BND_TAUX: return o->bnd[0];
    case PCOM_F_BND_TAUY: return o->bnd[1];
    case PCOM_F_BND_T: return o->bnd[2];
    case PCOM_F_BND_S: return o->bnd[3];
    case PCOM_F_BND_PA: return o->bnd[4];
    case PCOM_F_BND_FW: return o->bnd[5];
    case PCOM_F_SSH: return o->ssh;
    case PCOM_F_GRAPHIHX: return o->graphihx;
    case PCOM_F_GRAPHIHY: return o->graphihy;
    case PCOM_F_COR: return o->cor;
    case PCOM_F_PHIH_T: return o->phih_t;
    case PCOM_F_PRH_T: return o->prh_t;
    case PCOM_F_PHIH_U: return o->phih_u;
    case PCOM_F_PRH_U: return o->prh_u;
    case PCOM_F_KM: *nd3 = nkp; return o->km;
    case PCOM_F_KH: *nd3 = nkp; return o->kh;
    default: return NULL;
  }
}

size_t orc_field_size(const orc *o, int id) {
  int nd3;
  double *p = field_ptr(o, id, &nd3);
  return p ? o->n2 * (size_t)nd3 : 0;
}

int orc_get(const orc *o, int id, double *out, size_t n) {
  int nd3;
  double *p = field_ptr(o, id, &nd3);
  if (!p || n != o->n2 * (size_t)nd3) return -1;
  memcpy(out, p, n * sizeof(double));
  return 0;
}

int orc_set(orc *o, int id, const double *in, size_t n) {
  int nd3;
  double *p = field_ptr(o, id, &nd3);
  if (!p || n != o->n2 * (size_t)nd3) return -1;
  memcpy(p, in, n * sizeof(double));
  return 0;
}

void orc_dims(const orc *o, int *ni, int *nj, int *nk) { *ni = o->ni; *nj = o->nj; *nk = o->nk; }

void orc_counters(const orc *o, pcom_counters *c) {
  c->step = o->nsteps;
  c->n_ts = o->n_ts; c->n_uv = o->n_uv; c->n_w = o->n_w; c->n_ch = o->n_ch; c->n_ssh = o->n_ssh;
  c->year = o->ct.y; c->month = o->ct.m; c->day = o->ct.d; c->hour = o->ct.h;
  c->minute = o->ct.mi; c->second = o->ct.s;
}

void orc_op_ter(orc *o, int ga, int gb, int nd3, double dft, const double *in, double *out) {
  ter(o, out, in, nd3, ga, gb, dft);
}
void orc_op_px1(orc *o, int ga, int gb, int nd3, const double *in, double *out) {
  p1(o, out, in, nd3, ga, gb, NULL);
}
void orc_op_px2(orc *o, int ga, int gb, int nd3, const double *in, double *out) {
  p2(o, out, in, nd3, ga, gb, NULL);
}
void orc_op_div(orc *o, int ga, int gb, int gc, int nd3, const double *va, const double *vb, double *out) {
  divg(o, out, va, vb, nd3, ga, gb, gc);
}
void orc_op_lap(orc *o, int ga, int gb, const double *in, double *out) { lap(o, out, in, ga, gb); }
void orc_op_adv(orc *o, const double *var, const double *sch, double *out) { op_adv(o, out, var, sch); }
void orc_op_adv_ts(orc *o, const double *var, const double *wm, double *out) { op_adv_ts(o, out, var, wm); }
void orc_op_dif(orc *o, const double *ch, const double *var, double *out) { op_dif(o, out, ch, var); }
void orc_op_fri(orc *o, const double *sch) { op_fri(o, sch); }
void orc_upwelling(orc *o) { upwelling(o); }
void orc_smag(orc *o, const double *u, const double *v) { smag(o, u, v); }

int orc_metric(const orc *o, int g, int which, double *out, size_t n) {
  if (g < 1 || g > 4 || n != o->n2) return -1;
  const double *src = which == 0 ? o->rh[g] : which == 1 ? o->tn[g] : which == 2 ? o->dx1[g]
                    : which == 3 ? o->dx2[g] : which == 4 ? o->hposx + n * (size_t)(g - 1)
                                                          : o->hposy + n * (size_t)(g - 1);
  memcpy(out, src, n * sizeof(double));
  return 0;
}

int orc_vgrid(const orc *o, int gi, int which, double *out, size_t n) {
  const double *src;
  size_t len = gi == 1 ? (size_t)o->nkp : (size_t)o->nk;
  if (n != len) return -1;
  if (gi == 1) src = which == 0 ? o->z1 : which == 1 ? o->pr1 : which == 2 ? o->dz1 : o->dpr1;
  else src = which == 0 ? o->z2 : which == 1 ? o->pr2 : which == 2 ? o->dz2 : o->dpr2;
  memcpy(out, src, n * sizeof(double));
  return 0;
}

int orc_igrid(const orc *o, int which, int *out, size_t n) {
  const int *src;
  size_t len = which < 2 ? o->n3 : o->n2;
  if (n != len) return -1;
  src = which == 0 ? o->msk_t : which == 1 ? o->msk_u : which == 2 ? o->lev_t : o->lev_u;
  memcpy(out, src, n * sizeof(int));
  return 0;
}
