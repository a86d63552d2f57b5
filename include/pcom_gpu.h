/*
 * pcom_gpu.h - C ABI of the MI355X time-stepping library for the PCOM
 * pressure-coordinate B-grid ocean model (ouyuyuan/pcom v2.2).
 *
 * The library replaces the Fortran module procedures that main.f90 calls
 * once per baroclinic step (main.f90:158-238):
 *
 *   den_prho  (mod_den.f90:120-153)   int_bnd   (mod_int.f90:649-691)
 *   den_alpha (mod_den.f90:33-55)     int_pgra  (mod_int.f90:64-88)
 *   int_readyc(mod_int.f90:90-151)    int_trop  (mod_int.f90:153-296)
 *   int_clin  (mod_int.f90:298-401)   int_ssh   (mod_int.f90:602-624)
 *   int_ts    (mod_int.f90:442-507)
 *
 * and the start-up chain that fixes every static field those routines read
 * (main.f90:252-704: set_cf, calc_hpos, set_hgrid, set_vgrid, set_grid,
 * calc_cf, Coriolis, inistat).  The reference passes non-interoperable
 * derived types and module globals (mod_type.f90:138-205, mod_op.f90:13-19);
 * here the caller hands over plain arrays once, the device owns the state,
 * and fields are copied in and out on request.
 *
 * Layout at the boundary: every 2-D/3-D field is FP64, Fortran
 * column-major, WITH the reference's one-cell halo: (ni, nj[, nk|nkp]) with
 * ni = glo_ni + 2, nj = glo_nj + 2 (mod_param.f90:313-342 with one rank).
 * Raw inputs (pcom_inputs) are the global arrays of initial.nc/forcing.nc
 * (glo_ni, glo_nj[, nk|12]) without halo, as mod_io reads them.
 *
 * Errors: every call returns 0 or a negative PCOM_ERR_* code and never
 * aborts; pcom_last_error() returns the message of the last failure on the
 * calling thread.  The reference stops the program instead (mod_op.f90:337,
 * mod_io.f90:540-550).
 *
 * Threading: one host thread per context; calls are stream-ordered on the
 * context's own HIP stream; pcom_get_field() synchronises.
 */
#ifndef PCOM_GPU_H
#define PCOM_GPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  PCOM_OK = 0,
  PCOM_ERR_ARG = -1,      /* bad argument, size or field id               */
  PCOM_ERR_HIP = -2,      /* HIP runtime error (message says which call)  */
  PCOM_ERR_NAN = -3,      /* non-finite value found by pcom_check_finite  */
  PCOM_ERR_STATE = -4,    /* call not valid in the context's state        */
  PCOM_ERR_NOGPU = -5     /* no HIP device visible                        */
};

/* Arithmetic of the 3-D baroclinic fields.  In every form the prognostic
 * state (T, S, u, v) is stored and stepped in FP64, and the barotropic
 * subcycle, its time filter and the global sums are FP64. */
enum {
  PCOM_FP64 = 0,        /* everything FP64, bit-identical to the oracle    */
  PCOM_MIXED = 1,       /* 3-D tendency stencils (advection, friction,
                           diffusion, upwelling, Smagorinsky) in FP32; the
                           equation of state and the pressure integral and
                           gradient in FP64                                */
  PCOM_MIXED_EOS32 = 2  /* as PCOM_MIXED, with the equation of state and
                           the pressure integral in FP32 as well           */
};

/* Raw inputs, as read from initial.nc, forcing.nc and vmix.nc
 * (main.f90:263-345, mod_mympi.f90:493-560).  Column-major, no halo. */
typedef struct pcom_inputs {
  int glo_ni, glo_nj, nk;
  const double *lon;   /* [glo_ni]  degrees east, grid g4 (initial.nc lon) */
  const double *lat;   /* [glo_nj]  degrees north, grid g4                 */
  const double *z;     /* [nk]      depth of layer centres, m              */
  const int    *itn;   /* [glo_ni*glo_nj] number of wet levels (0 = land)  */
  const double *tmn;   /* [nk]      layer-mean T for the vertical grid     */
  const double *smn;   /* [nk]      layer-mean S                           */
  const double *pt;    /* [glo_ni*glo_nj*nk] initial potential temperature */
  const double *sa;    /* [glo_ni*glo_nj*nk] initial salinity              */
  const double *vmix;  /* [glo_ni*glo_nj*nk] kh in cm^2/s at interfaces
                          2..nk+1, or NULL = fill value everywhere         */
  const double *taux, *tauy, *bct, *bcs, *pa, *fw; /* [glo_ni*glo_nj*12]   */
} pcom_inputs;

typedef struct pcom_opts {
  int dtbt;        /* barotropic step, s (namelist time_ctrl dtbt)         */
  int dtbc;        /* baroclinic step, s (namelist time_ctrl dtbc)         */
  int year, month, day, hour;   /* bdate                                   */
  int precision;   /* PCOM_FP64 or PCOM_MIXED                              */
  int restart_run; /* namelist restart_run: 1 skips the AB3 bootstrap      */
  int device;      /* HIP device ordinal                                   */
  int trop_tile_i, trop_tile_j; /* barotropic subcycle tile (0, 0 and 0
                                   substeps = the library's choice)        */
  int trop_substeps;            /* substeps between tile exchanges         */
  int trop_mode;   /* PCOM_TROP_AUTO, _PERSISTENT or _LAUNCHES             */
} pcom_opts;

/* How the barotropic subcycle runs.  PERSISTENT: one launch per baroclinic
 * step whose tiles stay resident and exchange halos through device memory
 * (needs every tile co-resident: grids up to about 1 degree).  LAUNCHES:
 * one launch per `substeps` substeps, the kernel boundary being the
 * exchange (any grid size).  AUTO: PERSISTENT when the grid fits, else
 * LAUNCHES.  The results are the same bits in every mode.               */
enum { PCOM_TROP_AUTO = 0, PCOM_TROP_PERSISTENT = 1, PCOM_TROP_LAUNCHES = 2 };

typedef struct pcom_ctx pcom_ctx;

/* Field ids for pcom_get_field / pcom_set_field.  Shapes are the halo'd
 * reference shapes; (3D) = (ni,nj,nk), (3Dp) = (ni,nj,nk+1), (2D) = (ni,nj). */
enum {
  PCOM_F_TC = 0, PCOM_F_SC, PCOM_F_TP, PCOM_F_SP,            /* 3D, g1 */
  PCOM_F_UC, PCOM_F_VC, PCOM_F_UP, PCOM_F_VP,                /* 3D, g3 */
  PCOM_F_AUC, PCOM_F_AVC, PCOM_F_AUP, PCOM_F_AVP,
  PCOM_F_AUPP, PCOM_F_AVPP,                                  /* 3D, g3 */
  PCOM_F_FX, PCOM_F_FY, PCOM_F_AM,                           /* 3D, g3 */
  PCOM_F_ALPHA, PCOM_F_ADP,                                  /* 3D, g1 */
  PCOM_F_WM,                                                 /* 3Dp, g1 */
  PCOM_F_PRHO_T, PCOM_F_PRHO_S,                              /* 3D, g1 */
  PCOM_F_ACT, PCOM_F_ACS, PCOM_F_ACR, PCOM_F_ACU, PCOM_F_ACV,
  PCOM_F_ACW,                                                /* 3D     */
  PCOM_F_CHC, PCOM_F_CHP, PCOM_F_MCH, PCOM_F_MCH2,           /* 2D, g1 */
  PCOM_F_ACCH, PCOM_F_ACSSH,                                 /* 2D, g1 */
  PCOM_F_BUC, PCOM_F_BVC, PCOM_F_BUP, PCOM_F_BVP,
  PCOM_F_BUT, PCOM_F_BVT,                                    /* 2D, g3 */
  PCOM_F_PAX, PCOM_F_PAY,                                    /* 2D, g3 */
  PCOM_F_BPHI_XN, PCOM_F_BPHI_XS, PCOM_F_BPHI_YE, PCOM_F_BPHI_YW,
  PCOM_F_BGRA_XN, PCOM_F_BGRA_XS, PCOM_F_BGRA_YE, PCOM_F_BGRA_YW,
  PCOM_F_BND_TAUX, PCOM_F_BND_TAUY, PCOM_F_BND_T, PCOM_F_BND_S,
  PCOM_F_BND_PA, PCOM_F_BND_FW,                              /* 2D     */
  PCOM_F_SSH,        /* 2D: sea surface height of the last int_ssh call */
  PCOM_F_GRAPHIHX, PCOM_F_GRAPHIHY, PCOM_F_COR,              /* 2D, g3 */
  PCOM_F_PHIH_T, PCOM_F_PRH_T, PCOM_F_PHIH_U, PCOM_F_PRH_U,  /* 2D     */
  PCOM_F_KM, PCOM_F_KH,                                      /* 3Dp    */
  PCOM_F_COUNT
};

/* Counters of the time-mean accumulators (eqts%n, equv%n, eqw%n, eqch%n,
 * acssh%n in mod_type.f90) and of the time control (tctr%i). */
typedef struct pcom_counters {
  long long step;          /* baroclinic steps taken since create (tctr%i) */
  int n_ts, n_uv, n_w, n_ch, n_ssh;
  int year, month, day, hour, minute, second;   /* tctr%ct */
} pcom_counters;

/* Create a context: runs the reference's start-up chain on the host
 * (main.f90:104-122, 312-704), uploads the state and builds the step.   */
int pcom_create(const pcom_inputs *in, const pcom_opts *opts, pcom_ctx **out);

/* Advance nsteps baroclinic steps: main.f90:158-227 without output,
 * including the daily int_bnd + den_prho on day change.  Returns with the
 * work queued on the context's two streams; every entry point that reads
 * or writes the state waits for it first.                              */
int pcom_step(pcom_ctx *ctx, int nsteps);

/* Dimensions of the halo'd arrays. */
int pcom_get_dims(const pcom_ctx *ctx, int *ni, int *nj, int *nk);

/* Number of doubles of field id (0 for an unknown id). */
size_t pcom_field_size(const pcom_ctx *ctx, int id);

/* Copy a field to / from host memory (FP64, halo'd, column-major).  In
 * the mixed forms, fields stored in FP32 on the device are converted.   */
int pcom_get_field(pcom_ctx *ctx, int id, double *host, size_t n);
int pcom_set_field(pcom_ctx *ctx, int id, const double *host, size_t n);

int pcom_get_counters(pcom_ctx *ctx, pcom_counters *c);

/* Restart clock (main.f90:128-141): after loading a restart file, set the
 * index of the last completed step (tctr%ist - 1), the current date
 * (tctr%ct = cdate) and the previous date (tctr%pt = pdate), each date as
 * {year, month, day, hour, minute, second}.  The next pcom_step is step
 * index+1: it refreshes the daily forcing only on a day change, as the
 * reference does for i /= 1, and skips the AB3 bootstrap when the context
 * was created with restart_run = 1.                                     */
int pcom_set_clock(pcom_ctx *ctx, long long step, const int cdate[6], const int pdate[6]);

/* Forcing records the caller reads from its daily or 6-hourly files
 * (main.f90:164-202: namelist ff_ws -> taux, tauy; ff_pa -> pa; ff_bct ->
 * bct).  Each is a global interior array (glo_ni, glo_nj), column-major,
 * as mympi_input reads one record (mod_mympi.f90:638-657); NULL keeps the
 * value int_bnd interpolated.                                            */
typedef struct pcom_forcing {
  const double *taux, *tauy;  /* wind stress, N m-2          (bnd%taux, bnd%tauy) */
  const double *pa;           /* atmospheric pressure         (bnd%pa)            */
  const double *bct;          /* restoring temperature, degC  (bnd%t)             */
} pcom_forcing;

/* Which forcing refreshes the next pcom_step performs, by the reference's
 * conditions for its step i (main.f90:164, 184-185): due[0] = 1 when i == 1
 * or the day changed (int_bnd, then the daily records, then den_prho);
 * due[1] = 1 when i == 1 or the hour changed to a multiple of 6 (the
 * 6-hourly records of out_per = '6hour', then den_prho).                 */
int pcom_forcing_due(pcom_ctx *ctx, int due[2]);

/* Apply forcing records for the next step, before it runs, in the order of
 * main.f90:164-202.  interpolate = 1 (the daily refresh): int_bnd at the
 * current date, then the non-NULL records of f replace its fields, then
 * den_prho; the next step does not repeat its own daily refresh.
 * interpolate = 0 (the 6-hourly refresh): the records replace the current
 * fields, then den_prho; it fails with PCOM_ERR_STATE while the next
 * step's daily refresh is due and not yet applied, which must come first.
 * The pressure gradient of int_readyc follows every change of bnd pa
 * (also through pcom_set_field).  Not calling this leaves pcom_step to
 * refresh the daily forcing from the monthly fields, as the reference does
 * without forcing files.                                                 */
int pcom_set_forcing(pcom_ctx *ctx, const pcom_forcing *f, int interpolate);

/* missing_float of mod_param.f90:62 (a float32 constant) */
#define PCOM_MISSING_FLOAT (-9.99999e30f)

/* Time means of one output interval, as main.f90:726-760 (output) with
 * mod_mympi.f90:178-343 computes them: each accumulator divided by its
 * record count; acr masked by the T mask; acu, acv and acw by the U mask;
 * acch and SSH by the surface T mask (land = PCOM_MISSING_FLOAT); SSH
 * minus its global mean weighted by rh*msk.  Arrays are the global
 * interior (glo_ni, glo_nj[, nk]), FP64, column-major, as mod_io writes
 * them; a NULL pointer skips that copy.                                  */
typedef struct pcom_means {
  double *pt, *sa, *rho;   /* eqts%act, acs, acr / eqts%n        (3-D) */
  double *u, *v, *w;       /* equv%acu, acv / equv%n, eqw%acw / eqw%n  */
  double *ch;              /* eqch%acch / eqch%n                  (2-D) */
  double *ssh;             /* acssh / acssh%n, minus the mean     (2-D) */
} pcom_means;

/* Compute the means, copy them out, then reset every accumulator and its
 * count to zero (the reference resets after each output).  Fails with
 * PCOM_ERR_STATE when a count is zero (no step since the last output).  */
int pcom_output_means(pcom_ctx *ctx, pcom_means *out);

/* pcom_output_means that allocates nothing: the means are staged in a
 * static device buffer of the library (16 MiB, held by each GPU the library
 * runs on until the process ends) instead of one allocated at the first
 * output.  No counterpart in the reference: for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Limit: glo_ni*glo_nj*nk + 2 <=
 * 2097152 doubles (1 degree x 30 levels fits); a larger grid fails with
 * PCOM_ERR_ARG.  All calls on one GPU share the buffer, which keeps the
 * last call's contents: calls that overlap on one GPU (from several host
 * threads) race.  Same results and accumulator reset as pcom_output_means. */
int pcom_output_means_noalloc(pcom_ctx *ctx, pcom_means *out);

/* Horizontal metric of stagger g (1..4) from the start-up chain
 * (main.f90:408-497): which = 0 rh = cos(lat), 1 tn = tan(lat),
 * 2 dx%x(1), 3 dx%x(2) (radians); halo'd 2-D, column-major.             */
int pcom_get_metric(const pcom_ctx *ctx, int g, int which, double *out, size_t n);

/* Area/volume-weighted global sums over sea points, accumulated in FP64 on
 * the device: out[0] = sum(rh*dx1*dx2*dpr*msk*T), out[1] = same for S,
 * out[2] = sum(rh*dx1*dx2*msk*chc) (column mass).  New: the reference has
 * no global reduction besides the SSH mean (mod_mympi.f90:308-343).      */
int pcom_global_sums(pcom_ctx *ctx, double out[3]);

/* Returns PCOM_ERR_NAN if any prognostic field holds a non-finite value. */
int pcom_check_finite(pcom_ctx *ctx);

/* Wait for all work queued on the context's streams. */
int pcom_sync(pcom_ctx *ctx);

/* Average device duration (ms) of each phase over the steps since the last
 * reset, measured with HIP events on the context's stream; phases are
 * 0 = whole step, 1 = everything but the subcycle kernel (3-D baroclinic
 * kernels and the 2-D set-up of the subcycle), 2 = the barotropic
 * subcycle kernel alone (k_trop_tiles, one launch per step).            */
int pcom_phase_times(pcom_ctx *ctx, int enable, double out_ms[3], int *nsteps);

int pcom_destroy(pcom_ctx *ctx);

const char *pcom_last_error(void);

/* Single-operator entry points on the context's grid, for parity tests.
 * Stagger ids 1..4 are the reference's g1j..g4j (mod_arrays.f90:148-166);
 * in/out are halo'd host arrays of nd3 levels.                          */

/* op_ter for 2-D/3-D reals (mod_op.f90:361-427); dft is the prefill.   */
int pcom_op_ter(pcom_ctx *ctx, int ga, int gb, int nd3, double dft,
                const double *in, double *out);
/* op_px1 / op_px2 (mod_op.f90:496-527, 597-652, 689-713).              */
int pcom_op_px1(pcom_ctx *ctx, int ga, int gb, int nd3,
                const double *in, double *out);
int pcom_op_px2(pcom_ctx *ctx, int ga, int gb, int nd3,
                const double *in, double *out);
/* op_div (mod_op.f90:813-848): (va on ga, vb on gb) -> gc.             */
int pcom_op_div(pcom_ctx *ctx, int ga, int gb, int gc, int nd3,
                const double *va, const double *vb, double *out);
/* op_lap (mod_op.f90:904-925) on 2-D fields.                           */
int pcom_op_lap(pcom_ctx *ctx, int ga, int gb, const double *in, double *out);
/* den_rho (mod_den.f90:57-118) on n points: t degC, s psu, p Pa.       */
int pcom_den_rho(pcom_ctx *ctx, size_t n, const double *t, const double *s,
                 const double *p, double *rho);

/* ---- Ensembles and the subcycle autotuner.  New: the reference runs one
 * ocean per (MPI) job.
 *
 * pcom_ensemble_step advances every member by nsteps steps with the same
 * results as pcom_step on each.  Each member's 3-D kernels run on the
 * member's own stream; the members' barotropic subcycles run in packed
 * launches, as many members per launch as the device holds co-resident
 * (first fit by tile count), so small members fill the device together.
 * Members may differ in grid size and form and must outlive the ensemble;
 * a member with a surrogate in use runs its own barotropic part.        */
typedef struct pcom_ensemble pcom_ensemble;
int pcom_ensemble_create(pcom_ctx *const *members, int count, pcom_ensemble **out);
int pcom_ensemble_step(pcom_ensemble *ens, int nsteps);
/* Packing of the last pcom_ensemble_step: number of launches and the
 * launch of each member (-1: the member runs alone).                    */
int pcom_ensemble_packs(pcom_ensemble *ens, int *npacks, int *pack_of_member);
int pcom_ensemble_destroy(pcom_ensemble *ens);
/* Autotuner: time the subcycle on this context's state under the `trials`
 * tilings (tile i, tile j, substeps per exchange) a cost model ranks best,
 * restoring the state after each, and keep the fastest.  Returns the chosen
 * tiling and its time per launch.  The state is unchanged.               */
int pcom_tune_subcycle(pcom_ctx *ctx, int trials, int tile[3], double *ms);
/* Set the subcycle tiling {tile i, tile j, substeps per exchange, mode}
 * (as pcom_opts.trop_tile_*, trop_substeps and trop_mode at create; zero
 * tile sizes and substeps = the cost model's choice for the mode) or read
 * the tiling in use, its mode resolved (set = 0).  The results do not
 * depend on the tiling or the mode.                                      */
int pcom_subcycle_tiling(pcom_ctx *ctx, int set, int tile[4]);
/* Bound on each neighbour wait of the persistent subcycle: a wait that has
 * polled `polls` times without success ends in an error (0: the default,
 * 2^22 polls, seconds, where a normal wait takes microseconds).  The tile
 * stops, the launch still ends, the state is marked invalid and every
 * later synchronising call reports PCOM_ERR_STATE.  No counterpart in the
 * reference; a bound of 1 forces that path, for tests of failure handling. */
int pcom_subcycle_wait_limit(pcom_ctx *ctx, unsigned long long polls);

/* ---- Learned barotropic surrogate.  New: the reference has no such part.
 * A BF16 convolutional network (six 3x3 layers, 64 hidden channels, FP32
 * accumulation on the matrix cores) predicts what int_trop returns
 * (mod_int.f90:153-296): the step-mean free surface mch2 and mch and the
 * step-mean barotropic velocity, from chc, buc, bvc, but, bvt, the face
 * integrals bphi and bgraphi of int_pgra, sea masks and Coriolis.  It is
 * trained on the device from snapshots of the plain subcycle.  When in
 * use, every check_every-th step still runs the plain subcycle, which is
 * the fallback and measures the network's error there; an error above tol
 * turns the surrogate off.                                               */
typedef struct pcom_nn_stats {
  long long recorded;        /* snapshots held in the store                */
  long long surrogate_steps; /* steps whose subcycle the network replaced  */
  long long plain_steps;     /* steps that ran the plain subcycle          */
  long long checks;          /* plain steps that also checked a prediction */
  long long fallbacks;       /* checks that turned the surrogate off       */
  int in_use;
  double last_check_rmse[4]; /* sea RMSE of mch2, mch (dimensionless ch),  */
  double max_check_rmse[4];  /* ubt, vbt (m/s) at the last / worst check   */
  double last_persist_rmse[4]; /* same for persistence (no change over the
                                  step) at the last check: the method
                                  without the network                      */
} pcom_nn_stats;

/* Attach a network with seeded random weights and a device store of up to
 * `capacity` snapshots (0: no store).  Needs glo_ni*glo_nj % 16 == 0.   */
int pcom_nn_create(pcom_ctx *ctx, int capacity, unsigned long long seed);
/* Record (inputs, outputs) of every plain subcycle from now on (1) or not.
 * While the network is in use (pcom_nn_use), a step it takes is recorded
 * too: its inputs with the plain subcycle's outputs from the same state,
 * run beside it and then undone, so that training sees the states the
 * network itself leads to.                                              */
int pcom_nn_record(pcom_ctx *ctx, int on);
/* While recording, also record, before a `fraction` of the steps, two
 * extra snapshots: the state perturbed by seeded noise of random smoothness
 * and sea RMS up to amp_ch (ch) and amp_u (m/s, u and v) with its plain
 * subcycle, and the state that subcycle leaves with its own; then the
 * state is restored and the step itself runs unperturbed.
 * Such snapshots teach the network how the plain cycle damps unbalanced
 * states, which its own errors create when it is in use (fraction 0: off). */
int pcom_nn_augment(pcom_ctx *ctx, double fraction, double amp_ch, double amp_u, unsigned long long seed);
/* Adam training on the recorded snapshots: `iters` minibatches of `batch`
 * snapshots drawn with `seed`, learning rate lr decayed to 0 by a cosine
 * over the call; loss[iters] (mean squared error in normalised units) may
 * be NULL.  A non-finite gradient element leaves its weight unchanged;
 * weights that still end non-finite fail the call (PCOM_ERR_HIP,
 * "training diverged").                                                 */
int pcom_nn_train(pcom_ctx *ctx, int iters, int batch, double lr, unsigned long long seed, double *loss);
/* Use the network in pcom_step with a plain, checking subcycle every
 * check_every steps (0 turns the surrogate off).  At each check, tol > 0
 * bounds the sea RMSE of every output, tol < 0 bounds it by -tol times the
 * RMSE of persistence (no change over the step), both at this check and at
 * the network's first check, whichever is smaller (a run drifting into
 * noise raises the first with the error); tol = 0 never falls back; a
 * failed check turns the surrogate off, or holds it for a number of plain
 * steps set by pcom_nn_resume.  The first check also sets the pointwise
 * limit of the network's increments (50 times its persistence RMSE); the
 * network's first step after this call is a check until one has done so,
 * and the limit then holds across later calls.                          */
int pcom_nn_use(pcom_ctx *ctx, int check_every, double tol);
/* After a failed check, run plain_steps plain steps and then resume the
 * network (counted in pcom_nn_stats.fallbacks), whose first step after them
 * is a check, so that a network that failed never runs unchecked; 0 (the
 * default) keeps the plain subcycle for good.  Extends the check of the surrogate config
 * (plain subcycle every 10th step as check and fallback): the reference
 * has no surrogate.                                                      */
int pcom_nn_resume(pcom_ctx *ctx, int plain_steps);
/* Root-mean-square error over sea points of the network on recorded
 * snapshots [first, first+count) (in recording order), in physical units:
 * rmse[0..3] for the mch2, mch, ubt, vbt increments and rmse[4..7] for
 * persistence (no change over the step).  Held-out snapshots, recorded
 * after training, measure what the network gains.                       */
int pcom_nn_eval(pcom_ctx *ctx, int first, int count, double rmse[8]);
int pcom_nn_get_stats(pcom_ctx *ctx, pcom_nn_stats *s);
/* Parameters: per layer the FP32 master weights [cout][3][3][cin] (channels
 * padded: input 32 (16 used), hidden 64, output 32 (4 used)) and biases
 * [cout], then 31 normalisation values (offsets and scales of the 12
 * recorded inputs, output scales 4, Coriolis and bottom-pressure scales,
 * the floor of the per-snapshot scale).  Setting them marks the network as
 * trained.                                                               */
size_t pcom_nn_param_count(const pcom_ctx *ctx);
int pcom_nn_get_params(pcom_ctx *ctx, float *host, size_t n);
int pcom_nn_set_params(pcom_ctx *ctx, const float *host, size_t n);
/* One forward and backward pass on normalised host tensors, for gradient
 * checks: x [batch][glo_nj][glo_ni][16], target [batch][glo_nj][glo_ni][4];
 * out gets the output (same shape as target), grads the loss gradient in
 * parameter order (zero for the normalisation values).                  */
int pcom_nn_forward_backward(pcom_ctx *ctx, int batch, const float *x, const float *target, float *out,
                             float *grads, double *loss);

#ifdef __cplusplus
}
#endif
#endif /* PCOM_GPU_H */
