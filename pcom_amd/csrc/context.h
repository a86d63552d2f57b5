// context.h - the device-resident state of one PCOM time stepper.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "pcom_dev.h"
#include "setup.h"

namespace pcom {

struct Nn;        // surrogate.hip: the learned barotropic surrogate
struct Ensemble;  // trop.hip: contexts stepped together

// 3-D fields are stored in R (double, or float in the mixed form); 2-D
// fields (free surface, barotropic mode, forcing) are always double.
struct Fields3 {
  void *T[2], *S[2];        // potential temperature / salinity, ping-pong:
                            // T[cur] is both tc and tp (int_ts ends with
                            // tp = tc, mod_int.f90:496-497)
  void *U, *V;              // uc == up, vc == vp (int_clin ends with up = uc)
  void *AU[3], *AV[3];      // AB3 advection ring: auc = AU[ab], aup =
                            // AU[ab-1], aupp = AU[ab-2]
  void *fx, *fy, *am, *alpha, *adp, *wm, *prho1, *prho2;
  void *alphn, *adpn;  // the next step's den_alpha and int_pgra adp, formed by convect
  void *act, *acs, *acr, *acu, *acv, *acw;
  void *km, *kh;
};

struct Fields2 {
  double *chc, *chp, *mch, *mch2, *acch, *acssh, *ssh;
  double *buc, *bvc, *bup, *bvp, *but, *bvt, *pax, *pay;
  double *bphi[4], *bgra[4];   // xn xs ye yw
  double *bnd[6];              // taux tauy t s pa fw
  double *frc[6];              // (ni,nj,12)
  double *graphihx, *graphihy;
  double *vz1, *vz2, *am1;                      // subcycle statics
  double *schr, *sqp, *schs, *schm;             // sqrt(ter(ch)) factors on g3
  double *xc[2], *xu[2], *xv[2];                // subcycle tile exchange
  double *xacc[4];  // subcycle accumulators between launches (PCOM_TROP_LAUNCHES)
  double *red;                                  // reduction scratch
  int *flag;
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // a second stream for the kernels of a step that can run beside others
  // (den_alpha + int_pgra beside int_readyc, int_ssh beside the mode
  // coupling and smag), forked from and joined back into `stream` by events
  hipStream_t side = nullptr;
  hipEvent_t fork[3] = {nullptr, nullptr, nullptr}, join[3] = {nullptr, nullptr, nullptr};
  // convect (the end of a step) runs on `side` and overlaps the next step's
  // int_readyc; set while it has not been joined back into `stream`
  bool side_pending = false;
  // a step also forms what the next step's start would compute again from
  // unchanged inputs (schr; the upwelling, which int_ts's already is;
  // convect's den_alpha and adp, in alphn and adpn) and sets next_ready; the
  // next step starts from them unless an entry point changed the state
  bool next_ready = false;
  // int_readyc's pressure gradient (pax, pay) must be formed again from
  // bnd pa before the next step: set at create and by every entry point
  // that changes the state; the daily refresh forms it too
  bool pa_stale = true;
  // the step index whose daily forcing refresh the caller applied
  // (pcom_set_forcing with interpolate = 1), which that step then skips
  long long forcing_taken = 0;
  double* fstage = nullptr;  // (glo_ni, glo_nj) staging of caller forcing records
  int precision = PCOM_FP64;
  int rst = 0;
  int dtbt = 60, dtbc = 3600;
  int ni = 0, nj = 0, nk = 0, nkp = 0;
  size_t n2 = 0, n3 = 0, n3p = 0;
  HostSetup hs;
  Calendar ct, pt;
  long long step_i = 0;
  int n_ts = 0, n_uv = 0, n_w = 0, n_ch = 0, n_ssh = 0;
  int cur = 0, ab = 0;
  Grid<double> gd{};
  Grid<float> gf{};
  Fields3 f3{};
  Fields2 f2{};
  std::vector<void*> allocs;
  int trop_tw = 0, trop_th = 0, trop_s = 0;   // barotropic tile and substeps/exchange
  int trop_mode = 0;                          // requested PCOM_TROP_* mode
  bool trop_launches = false;                 // mode in use: one launch per exchange
  int trop_tiles_x = 0, trop_blocks = 0;
  size_t trop_lds = 0;
  int trop_block = 0;  // subcycle workgroup size of the tiling
  unsigned long long* trop_flags = nullptr;  // per-tile exchange flags (zeroed at create)
  long trop_flag_cap = 0;                    // tiles the flag array holds
  int* trop_error = nullptr;                 // set by a timed-out neighbour wait
  unsigned long long trop_spin_limit = 0;    // polls per neighbour wait (0: kSpinLimit)
  unsigned long long trop_epoch = 0;
  // phase timing (HIP events on `stream`)
  bool timing = false;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  double t_ms[3] = {0, 0, 0};
  int t_steps = 0;
  double* stage = nullptr;  // output staging, allocated on first use
  Nn* nn = nullptr;         // learned subcycle surrogate, if attached
  size_t rsize() const { return precision == PCOM_FP64 ? sizeof(double) : sizeof(float); }
};

// kernels.hip
std::string launch_step(Ctx& c);                 // one baroclinic step
std::string launch_step_pre(Ctx& c);             // up to int_trop
std::string launch_step_post(Ctx& c);            // after int_trop
std::string settle_side(Ctx& c);                 // c.stream waits for pending side-stream work
std::string launch_global_sums(Ctx& c, double out[3]);
// int_bnd at the current date (when interpolate), caller records over it
// (rec[v]: device (glo_ni, glo_nj) arrays for bnd v, or null), den_prho
std::string launch_set_forcing(Ctx& c, const double* const rec[6], bool interpolate);
std::string launch_check_finite(Ctx& c, int* bad);
// time means into `stage` (>= glo_ni*glo_nj*nk + 2 doubles), copied to the
// host pointers of o, then accumulators and counts reset
std::string launch_output_means(Ctx& c, pcom_means* o, double* stage);
std::string trop_configure(Ctx& c);              // trop.hip: size the subcycle tiles
std::string launch_trop(Ctx& c);                 // trop.hip: int_trop (k_vz, then the subcycle)
std::string launch_trop_kernel(Ctx& c);          // trop.hip: the subcycle alone
void trop_chain_forget(int device, hipStream_t st);  // trop.hip: before a stream is destroyed
std::string launch_vz(Ctx& c);                   // trop.hip: the subcycle's set-up kernel
std::string ensemble_create(const std::vector<Ctx*>& members, Ensemble** out);
std::string ensemble_step(Ensemble& en, int nsteps);
void ensemble_packs(const Ensemble& en, int* npacks, int* pack_of);
void ensemble_destroy(Ensemble* en);
std::string tune_subcycle(Ctx& c, int trials, int out[3], double* best_ms);
// surrogate.hip
std::string nn_create(Ctx& c, int capacity, unsigned long long seed);
void nn_destroy(Ctx& c);
void nn_set_record(Ctx& c, bool on);
std::string nn_set_augment(Ctx& c, double fraction, double amp_ch, double amp_u, unsigned long long seed);
std::string nn_set_use(Ctx& c, int every, double tol);
std::string nn_set_resume(Ctx& c, int plain_steps);
std::string nn_trop_step(Ctx& c, std::string (*plain)(Ctx&));
std::string nn_train(Ctx& c, int iters, int batch, double lr, unsigned long long seed, double* loss);
size_t nn_param_count(const Ctx& c);
std::string nn_eval(Ctx& c, int first, int count, double* res);
std::string nn_params(Ctx& c, float* host, bool set);
std::string nn_forward_backward(Ctx& c, int batch, const float* x, const float* target, float* out, float* grads,
                                double* loss);
void nn_stats(const Ctx& c, pcom_nn_stats* s);
// single-operator entry points on device buffers (FP64)
std::string launch_op_ter(Ctx& c, int ga, int gb, int nd3, double dft, const double* in, double* out);
std::string launch_op_px(Ctx& c, int dir, int ga, int gb, int nd3, const double* in, double* out);
std::string launch_op_div(Ctx& c, int ga, int gb, int gc, int nd3, const double* va, const double* vb, double* out);
std::string launch_op_lap(Ctx& c, int ga, int gb, const double* in, double* out);
std::string launch_den_rho(hipStream_t st, size_t n, const double* t, const double* s, const double* p, double* rho);

}  // namespace pcom
