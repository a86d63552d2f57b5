/*
 * pcom_oracle.h - CPU FP64 restatement of the PCOM time-stepping path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the MI355X library is
 * compared against; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product never links it.
 *
 * The restatement follows the reference routine by routine on full halo'd
 * arrays (ni = glo_ni+2, nj = glo_nj+2), reproducing its write ranges,
 * prefills, one-rank halo swap, loop orders and the float32 rounding of its
 * default-REAL literals.  Citations are to ouyuyuan/pcom v2.2 files.
 */
#ifndef PCOM_ORACLE_H
#define PCOM_ORACLE_H

#include <stddef.h>
#include "../include/pcom_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc orc;

orc *orc_create(const pcom_inputs *in, const pcom_opts *opts);
void orc_destroy(orc *o);
/* main.f90:158-227 for n steps (no output). */
void orc_step(orc *o, int n);
/* Same field ids and halo'd layout as pcom_get_field. Returns 0 or -1. */
int orc_get(const orc *o, int id, double *out, size_t n);
int orc_set(orc *o, int id, const double *in, size_t n);
size_t orc_field_size(const orc *o, int id);
void orc_dims(const orc *o, int *ni, int *nj, int *nk);
void orc_counters(const orc *o, pcom_counters *c);

/* Individual routines (state in o), for stage-wise parity tests. */
void orc_den_prho(orc *o);
void orc_den_alpha_main(orc *o);       /* main.f90:205                  */
void orc_int_pgra(orc *o);
void orc_int_readyc(orc *o);
void orc_int_trop(orc *o);
void orc_int_clin(orc *o);
void orc_int_ssh(orc *o);
void orc_int_ts(orc *o);
void orc_int_bnd(orc *o);
/* Advance the calendar as main.f90:225-226 does after int_ts. */
void orc_advance_time(orc *o);
/* Set tctr%i (the step index the AB3 bootstrap reads, mod_int.f90:122). */
void orc_set_step_index(orc *o, long long i);
/* Restart clock, as pcom_set_clock (main.f90:128-141). */
void orc_set_clock(orc *o, long long step, const int cdate[6], const int pdate[6]);
/* Caller forcing records for the next step, as pcom_set_forcing;
 * -1 when interpolate = 0 comes before a due daily refresh. */
int orc_set_forcing(orc *o, const pcom_forcing *f, int interpolate);
/* The refreshes the next step performs, as pcom_forcing_due. */
void orc_forcing_due(const orc *o, int due[2]);
/* output() time means, as pcom_output_means; -1 when a count is zero. */
int orc_output_means(orc *o, pcom_means *m);

/* Single operators on the context's grid (halo'd arrays, nd3 levels). */
void orc_op_ter(orc *o, int ga, int gb, int nd3, double dft,
                const double *in, double *out);
void orc_op_px1(orc *o, int ga, int gb, int nd3, const double *in, double *out);
void orc_op_px2(orc *o, int ga, int gb, int nd3, const double *in, double *out);
void orc_op_div(orc *o, int ga, int gb, int gc, int nd3,
                const double *va, const double *vb, double *out);
void orc_op_lap(orc *o, int ga, int gb, const double *in, double *out);
/* Composite operators on the context's fields (for invariant tests):
 * op_adv of var (g3, nk) with the context's uc, vc, wm, chc and the given
 * sch (n2); op_adv_ts of var (g1, nk) with the given wm (nk+1); op_dif of
 * var (g1, nk) with ch (n2); op_fri into the context's fx, fy; upwelling
 * into the context's wm. */
void orc_op_adv(orc *o, const double *var, const double *sch, double *out);
void orc_op_adv_ts(orc *o, const double *var, const double *wm, double *out);
void orc_op_dif(orc *o, const double *ch, const double *var, double *out);
void orc_op_fri(orc *o, const double *sch);
void orc_upwelling(orc *o);
/* smag (mod_int.f90:403-440) of unweighted velocities u, v (g3, nk) into
 * the context's am. */
void orc_smag(orc *o, const double *u, const double *v);
/* Horizontal metric arrays: which = 0 rh, 1 tn, 2 dx1, 3 dx2, 4 longitude,
 * 5 latitude (radians) of stagger g. */
int orc_metric(const orc *o, int g, int which, double *out, size_t n);
/* Vertical grid: which = 0 z, 1 pr, 2 dz, 3 dpr of gi1 (nk+1) or gi2 (nk). */
int orc_vgrid(const orc *o, int gi, int which, double *out, size_t n);
/* Integer grid: which = 0 g12 msk (3D), 1 g32 msk (3D), 2 g12 lev, 3 g32 lev. */
int orc_igrid(const orc *o, int which, int *out, size_t n);

/* Host threads of the OpenMP loops (1 in a serial build); set n > 0. */
int orc_threads(void);
void orc_set_threads(int n);

/* mod_den.f90:57-118 */
double orc_den_rho(double t, double s, double p0);
void orc_den_rho_n(size_t n, const double *t, const double *s,
                   const double *p, double *rho);
/* mod_con.f90:16-41 constants as the active build evaluates them:
 * names[i] and values[i] for i < count. */
int orc_constants(const char ***names, const double **values);
/* mod_int.f90:509-600 on one column (for brute-force tests). */
void orc_convect_column(int nk, int nlev, const double *dpr,
                        const double *prho_t, const double *prho_s,
                        double *t, double *s);

#ifdef __cplusplus
}
#endif
#endif
