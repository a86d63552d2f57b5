// api.hip - the extern "C" boundary declared in include/pcom_gpu.h.
//
// A context owns a HIP stream and the device copy of the model state.  The
// start-up chain runs on the host once (setup.cpp); afterwards fields cross
// the boundary only on explicit get/set calls.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pcom_gpu.h"
#include "context.h"

using namespace pcom;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

#define HIPCHK(call, what)                                                              \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess) return fail(PCOM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e_)); \
  } while (0)

// allocate and optionally fill from host doubles, converting to float when
// `as_float`
int dev_alloc(Ctx& c, void** out, size_t count, size_t elem) {
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, count * elem);
  if (e != hipSuccess) return fail(PCOM_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
  c.allocs.push_back(p);
  *out = p;
  return PCOM_OK;
}

// host -> device on the null stream: the context's stream is non-blocking,
// so its queued work is drained first
int upload(Ctx& c, void* dst, const double* src, size_t n, bool as_float) {
  if (c.stream) HIPCHK(hipStreamSynchronize(c.stream), "sync before upload");
  if (!as_float) {
    HIPCHK(hipMemcpy(dst, src, n * sizeof(double), hipMemcpyHostToDevice), "upload");
  } else {
    std::vector<float> tmp(n);
    for (size_t q = 0; q < n; ++q) tmp[q] = (float)src[q];
    HIPCHK(hipMemcpy(dst, tmp.data(), n * sizeof(float), hipMemcpyHostToDevice), "upload");
  }
  return PCOM_OK;
}

// a barotropic tile that timed out waiting for its neighbours leaves the
// state undefined: report it at the next synchronising call
int check_async_error(Ctx& c) {
  if (!c.trop_error) return PCOM_OK;
  int err = 0;
  HIPCHK(hipStreamSynchronize(c.stream), "sync");
  HIPCHK(hipMemcpy(&err, c.trop_error, sizeof(int), hipMemcpyDeviceToHost), "read subcycle status");
  if (err) return fail(PCOM_ERR_STATE, "barotropic subcycle: a tile exchange wait exceeded its bound; state invalid");
  return PCOM_OK;
}

// A step leaves int_ssh, int_ts, convect and the next step's int_pgra
// queued on the side stream.  An entry point that reads or writes the state
// first joins that work into c.stream; one that changes the state also drops
// what the last step formed for the next (the next step then forms it again).
int quiesce(Ctx& c, bool changes_state) {
  std::string e = settle_side(c);
  if (!e.empty()) return fail(PCOM_ERR_HIP, e);
  if (changes_state) {
    c.next_ready = false;
    c.pa_stale = true;
  }
  return PCOM_OK;
}

int download(Ctx& c, double* dst, const void* src, size_t n, bool as_float) {
  HIPCHK(hipStreamSynchronize(c.stream), "sync before download");
  if (int rc = check_async_error(c)) return rc;
  if (!as_float) {
    HIPCHK(hipMemcpy(dst, src, n * sizeof(double), hipMemcpyDeviceToHost), "download");
  } else {
    std::vector<float> tmp(n);
    HIPCHK(hipMemcpy(tmp.data(), src, n * sizeof(float), hipMemcpyDeviceToHost), "download");
    for (size_t q = 0; q < n; ++q) dst[q] = (double)tmp[q];
  }
  return PCOM_OK;
}

int fill_const(Ctx& c, void* dst, size_t n, double v, bool as_float) {
  std::vector<double> tmp(n, v);
  return upload(c, dst, tmp.data(), n, as_float);
}

// new device array of n values initialised from host doubles (or a constant)
int make(Ctx& c, void** out, size_t n, bool as_float, const double* src, double v = 0.0) {
  int rc = dev_alloc(c, out, n, as_float ? sizeof(float) : sizeof(double));
  if (rc) return rc;
  return src ? upload(c, *out, src, n, as_float) : fill_const(c, *out, n, v, as_float);
}

int make2(Ctx& c, double** out, size_t n, const double* src, double v = 0.0) {
  void* p;
  int rc = make(c, &p, n, false, src, v);
  *out = static_cast<double*>(p);
  return rc;
}

template <typename R>
int build_grid(Ctx& c, Grid<R>& g, bool as_float, const double* prh_t_d, const double* prh_u_d,
               const double* phih_t_d, const int* lev_t_d, const int* lev_u_d) {
  HostSetup& hs = c.hs;
  g.ni = c.ni; g.nj = c.nj; g.nk = c.nk; g.nkp = c.nkp; g.nim = c.ni - 1; g.njm = c.nj - 1;
  g.n2 = (long)c.n2;
  void* p;
  int rc;
  g.rh[0] = g.tn[0] = g.dx1[0] = g.dx2[0] = nullptr;
  for (int s = 1; s <= 4; ++s) {
    if ((rc = make(c, &p, c.n2, as_float, hs.rh[s].data()))) return rc;
    g.rh[s] = (const R*)p;
    if ((rc = make(c, &p, c.n2, as_float, hs.tn[s].data()))) return rc;
    g.tn[s] = (const R*)p;
    if ((rc = make(c, &p, c.n2, as_float, hs.dx1[s].data()))) return rc;
    g.dx1[s] = (const R*)p;
    if ((rc = make(c, &p, c.n2, as_float, hs.dx2[s].data()))) return rc;
    g.dx2[s] = (const R*)p;
  }
  g.lev_t = lev_t_d; g.lev_u = lev_u_d;
  g.phih_t = phih_t_d; g.prh_t = prh_t_d; g.prh_u = prh_u_d;
  if ((rc = make(c, &p, c.n2, as_float, hs.prh_t.data()))) return rc;
  g.prh_t_r = (const R*)p;
  if ((rc = make(c, &p, c.n2, as_float, hs.prh_u.data()))) return rc;
  g.prh_u_r = (const R*)p;
  if ((rc = make(c, &p, c.n2, as_float, hs.cv1.data()))) return rc;
  g.cv1 = (const R*)p;
  if ((rc = make(c, &p, c.n2, as_float, hs.cv2.data()))) return rc;
  g.cv2 = (const R*)p;
  if ((rc = make(c, &p, c.n2, as_float, hs.cor.data()))) return rc;
  g.cor = (const R*)p;
  if ((rc = make(c, &p, (size_t)c.nk, as_float, hs.pr2.data()))) return rc;
  g.pr2 = (const R*)p;
  if ((rc = make(c, &p, (size_t)c.nk, as_float, hs.dpr2.data()))) return rc;
  g.dpr2 = (const R*)p;
  if ((rc = make(c, &p, (size_t)c.nkp, as_float, hs.dz1.data()))) return rc;
  g.dz1 = (const R*)p;
  if ((rc = make(c, &p, c.n3p, as_float, hs.km.data()))) return rc;
  g.km = (const R*)p;
  if ((rc = make(c, &p, c.n3p, as_float, hs.kh.data()))) return rc;
  g.kh = (const R*)p;
  return PCOM_OK;
}

struct FieldRef {
  void* ptr = nullptr;
  size_t n = 0;
  bool is_float = false;  // stored in FP32 on the device
  const std::vector<double>* host = nullptr;  // host-only static fields
};

FieldRef field_ref(Ctx& c, int id) {
  FieldRef r;
  Fields3& f = c.f3;
  Fields2& h = c.f2;
  // 3-D fields: the FP64 state, E (EOS and pressure) or R (the rest)
  const bool rf = c.precision != PCOM_FP64, ef = c.precision == PCOM_MIXED_EOS32;
  auto state = [&](void* p, size_t n) {
    r.ptr = p;
    r.n = n;
  };
  auto eos = [&](void* p, size_t n) {
    r.ptr = p;
    r.n = n;
    r.is_float = ef;
  };
  auto three = [&](void* p, size_t n) {
    r.ptr = p;
    r.n = n;
    r.is_float = rf;
  };
  auto two = [&](double* p) {
    r.ptr = p;
    r.n = c.n2;
  };
  switch (id) {
    case PCOM_F_TC: case PCOM_F_TP: state(f.T[c.cur], c.n3); break;
    case PCOM_F_SC: case PCOM_F_SP: state(f.S[c.cur], c.n3); break;
    case PCOM_F_UC: case PCOM_F_UP: state(f.U, c.n3); break;
    case PCOM_F_VC: case PCOM_F_VP: state(f.V, c.n3); break;
    case PCOM_F_AUC: three(f.AU[c.ab], c.n3); break;
    case PCOM_F_AVC: three(f.AV[c.ab], c.n3); break;
    case PCOM_F_AUP: three(f.AU[(c.ab + 2) % 3], c.n3); break;
    case PCOM_F_AVP: three(f.AV[(c.ab + 2) % 3], c.n3); break;
    case PCOM_F_AUPP: three(f.AU[(c.ab + 1) % 3], c.n3); break;
    case PCOM_F_AVPP: three(f.AV[(c.ab + 1) % 3], c.n3); break;
    case PCOM_F_FX: three(f.fx, c.n3); break;
    case PCOM_F_FY: three(f.fy, c.n3); break;
    case PCOM_F_AM: three(f.am, c.n3); break;
    case PCOM_F_ALPHA: eos(f.alpha, c.n3); break;
    case PCOM_F_ADP: eos(f.adp, c.n3); break;
    case PCOM_F_WM: three(f.wm, c.n3p); break;
    case PCOM_F_PRHO_T: eos(f.prho1, c.n3); break;
    case PCOM_F_PRHO_S: eos(f.prho2, c.n3); break;
    case PCOM_F_ACT: three(f.act, c.n3); break;
    case PCOM_F_ACS: three(f.acs, c.n3); break;
    case PCOM_F_ACR: three(f.acr, c.n3); break;
    case PCOM_F_ACU: three(f.acu, c.n3); break;
    case PCOM_F_ACV: three(f.acv, c.n3); break;
    case PCOM_F_ACW: three(f.acw, c.n3); break;
    case PCOM_F_KM: three(f.km, c.n3p); break;
    case PCOM_F_KH: three(f.kh, c.n3p); break;
    case PCOM_F_CHC: two(h.chc); break;
    case PCOM_F_CHP: two(h.chp); break;
    case PCOM_F_MCH: two(h.mch); break;
    case PCOM_F_MCH2: two(h.mch2); break;
    case PCOM_F_ACCH: two(h.acch); break;
    case PCOM_F_ACSSH: two(h.acssh); break;
    case PCOM_F_BUC: two(h.buc); break;
    case PCOM_F_BVC: two(h.bvc); break;
    case PCOM_F_BUP: two(h.bup); break;
    case PCOM_F_BVP: two(h.bvp); break;
    case PCOM_F_BUT: two(h.but); break;
    case PCOM_F_BVT: two(h.bvt); break;
    case PCOM_F_PAX: two(h.pax); break;
    case PCOM_F_PAY: two(h.pay); break;
    case PCOM_F_BPHI_XN: two(h.bphi[0]); break;
    case PCOM_F_BPHI_XS: two(h.bphi[1]); break;
    case PCOM_F_BPHI_YE: two(h.bphi[2]); break;
    case PCOM_F_BPHI_YW: two(h.bphi[3]); break;
    case PCOM_F_BGRA_XN: two(h.bgra[0]); break;
    case PCOM_F_BGRA_XS: two(h.bgra[1]); break;
    case PCOM_F_BGRA_YE: two(h.bgra[2]); break;
    case PCOM_F_BGRA_YW: two(h.bgra[3]); break;
    case PCOM_F_BND_TAUX: two(h.bnd[0]); break;
    case PCOM_F_BND_TAUY: two(h.bnd[1]); break;
    case PCOM_F_BND_T: two(h.bnd[2]); break;
    case PCOM_F_BND_S: two(h.bnd[3]); break;
    case PCOM_F_BND_PA: two(h.bnd[4]); break;
    case PCOM_F_BND_FW: two(h.bnd[5]); break;
    case PCOM_F_SSH: two(h.ssh); break;
    case PCOM_F_GRAPHIHX: two(h.graphihx); break;
    case PCOM_F_GRAPHIHY: two(h.graphihy); break;
    case PCOM_F_COR: r.host = &c.hs.cor; r.n = c.n2; break;
    case PCOM_F_PHIH_T: r.host = &c.hs.phih_t; r.n = c.n2; break;
    case PCOM_F_PRH_T: r.host = &c.hs.prh_t; r.n = c.n2; break;
    case PCOM_F_PHIH_U: r.host = &c.hs.phih_u; r.n = c.n2; break;
    case PCOM_F_PRH_U: r.host = &c.hs.prh_u; r.n = c.n2; break;
    default: break;
  }
  return r;
}

int create_impl(const pcom_inputs* in, const pcom_opts* opts, pcom_ctx** out) {
  if (!in || !opts || !out) return fail(PCOM_ERR_ARG, "pcom_create: NULL argument");
  *out = nullptr;
  if (opts->dtbt <= 0 || opts->dtbc <= 0 || opts->dtbc % opts->dtbt != 0 || 86400 % opts->dtbt != 0 ||
      86400 % opts->dtbc != 0)
    return fail(PCOM_ERR_ARG, "dtbt and dtbc must divide 86400 and dtbt must divide dtbc (mod_param.f90:228-235)");
  if (opts->trop_tile_i < 0 || opts->trop_tile_j < 0 || opts->trop_substeps < 0 || opts->trop_tile_i > 256 ||
      opts->trop_tile_j > 256 || opts->trop_substeps > 16)
    return fail(PCOM_ERR_ARG, "barotropic tiling out of range (tile <= 256, substeps per exchange <= 16)");
  if (opts->trop_mode < PCOM_TROP_AUTO || opts->trop_mode > PCOM_TROP_LAUNCHES)
    return fail(PCOM_ERR_ARG, "trop_mode must be PCOM_TROP_AUTO, PCOM_TROP_PERSISTENT or PCOM_TROP_LAUNCHES");
  if (opts->precision != PCOM_FP64 && opts->precision != PCOM_MIXED && opts->precision != PCOM_MIXED_EOS32)
    return fail(PCOM_ERR_ARG, "precision must be PCOM_FP64, PCOM_MIXED or PCOM_MIXED_EOS32");
  if (opts->month < 1 || opts->month > 12 || opts->day < 1 || opts->day > 31 || opts->year < 1)
    return fail(PCOM_ERR_ARG, "bad start date");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(PCOM_ERR_NOGPU, "no HIP device visible");
  if (opts->device < 0 || opts->device >= ndev) return fail(PCOM_ERR_ARG, "device ordinal out of range");
  Ctx* c = new Ctx();
  c->device = opts->device;
  c->precision = opts->precision;
  c->rst = opts->restart_run;
  c->dtbt = opts->dtbt;
  c->dtbc = opts->dtbc;
  std::string err = build_setup(*in, c->hs);
  if (!err.empty()) {
    delete c;
    return fail(PCOM_ERR_ARG, "pcom_create: " + err);
  }
  HostSetup& hs = c->hs;
  if (hs.nk > 64) return fail(PCOM_ERR_ARG, "pcom_create: at most 64 levels are supported");
  c->ni = hs.ni; c->nj = hs.nj; c->nk = hs.nk; c->nkp = hs.nkp;
  c->n2 = hs.n2; c->n3 = hs.n3; c->n3p = hs.n3p;
  c->ct.y = opts->year; c->ct.m = opts->month; c->ct.d = opts->day; c->ct.h = opts->hour;
  c->ct.mi = 0; c->ct.s = 0; c->ct.mm = days_in_month(c->ct.y, c->ct.m);
  c->pt = c->ct;
  *out = reinterpret_cast<pcom_ctx*>(c);
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return fail(PCOM_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
  if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess ||
      (e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking)) != hipSuccess)
    return fail(PCOM_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  for (hipEvent_t* ev : {&c->fork[0], &c->fork[1], &c->fork[2], &c->join[0], &c->join[1], &c->join[2]})
    if ((e = hipEventCreateWithFlags(ev, hipEventDisableTiming)) != hipSuccess)
      return fail(PCOM_ERR_HIP, std::string("hipEventCreate: ") + hipGetErrorString(e));

  const bool fl = c->precision != PCOM_FP64;             // R = float
  const bool el = c->precision == PCOM_MIXED_EOS32;      // E = float
  const size_t n2 = c->n2, n3 = c->n3, n3p = c->n3p;
  int rc;
  // 2-D static (always FP64) and integer grids
  double *prh_t, *prh_u, *phih_t;
  if ((rc = make2(*c, &prh_t, n2, hs.prh_t.data()))) return rc;
  if ((rc = make2(*c, &prh_u, n2, hs.prh_u.data()))) return rc;
  if ((rc = make2(*c, &phih_t, n2, hs.phih_t.data()))) return rc;
  int *lev_t, *lev_u;
  {
    void* p;
    if ((rc = dev_alloc(*c, &p, n2, sizeof(int)))) return rc;
    lev_t = (int*)p;
    HIPCHK(hipMemcpy(lev_t, hs.lev_t.data(), n2 * sizeof(int), hipMemcpyHostToDevice), "upload lev");
    if ((rc = dev_alloc(*c, &p, n2, sizeof(int)))) return rc;
    lev_u = (int*)p;
    HIPCHK(hipMemcpy(lev_u, hs.lev_u.data(), n2 * sizeof(int), hipMemcpyHostToDevice), "upload lev");
  }
  if ((rc = build_grid<double>(*c, c->gd, false, prh_t, prh_u, phih_t, lev_t, lev_u))) return rc;
  if (fl && (rc = build_grid<float>(*c, c->gf, true, prh_t, prh_u, phih_t, lev_t, lev_u))) return rc;

  // 3-D prognostic state: FP64 in every form
  Fields3& f = c->f3;
  for (int q = 0; q < 2; ++q) {
    if ((rc = make(*c, &f.T[q], n3, false, hs.tc.data()))) return rc;
    if ((rc = make(*c, &f.S[q], n3, false, hs.sc.data()))) return rc;
  }
  if ((rc = make(*c, &f.U, n3, false, nullptr))) return rc;
  if ((rc = make(*c, &f.V, n3, false, nullptr))) return rc;
  for (int q = 0; q < 3; ++q) {
    if ((rc = make(*c, &f.AU[q], n3, fl, nullptr))) return rc;
    if ((rc = make(*c, &f.AV[q], n3, fl, nullptr))) return rc;
  }
  if ((rc = make(*c, &f.fx, n3, fl, nullptr))) return rc;
  if ((rc = make(*c, &f.fy, n3, fl, nullptr))) return rc;
  if ((rc = make(*c, &f.am, n3, fl, nullptr, AM_C))) return rc;
  if ((rc = make(*c, &f.alpha, n3, el, nullptr, (double)(1.0f / 1000.0f)))) return rc;
  if ((rc = make(*c, &f.adp, n3, el, nullptr))) return rc;
  if ((rc = make(*c, &f.alphn, n3, el, nullptr))) return rc;
  if ((rc = make(*c, &f.adpn, n3, el, nullptr))) return rc;
  if ((rc = make(*c, &f.wm, n3p, fl, nullptr))) return rc;
  if ((rc = make(*c, &f.prho1, n3, el, nullptr))) return rc;
  if ((rc = make(*c, &f.prho2, n3, el, nullptr))) return rc;
  void** acc[] = {&f.act, &f.acs, &f.acr, &f.acu, &f.acv, &f.acw};
  for (void** a : acc)
    if ((rc = make(*c, a, n3, fl, nullptr))) return rc;
  f.km = const_cast<void*>(fl ? (const void*)c->gf.km : (const void*)c->gd.km);
  f.kh = const_cast<void*>(fl ? (const void*)c->gf.kh : (const void*)c->gd.kh);

  // 2-D state (FP64)
  Fields2& h = c->f2;
  double** ones[] = {&h.chc, &h.chp, &h.mch, &h.mch2};
  for (double** a : ones)
    if ((rc = make2(*c, a, n2, nullptr, 1.0))) return rc;
  double** zeros[] = {&h.acch, &h.acssh, &h.ssh, &h.buc, &h.bvc, &h.bup, &h.bvp, &h.but, &h.bvt, &h.pax, &h.pay,
                      &h.vz1, &h.vz2, &h.am1, &h.schr, &h.sqp, &h.schs, &h.schm, &h.xc[0], &h.xc[1], &h.xu[0], &h.xu[1], &h.xv[0], &h.xv[1],
                      &h.xacc[0], &h.xacc[1], &h.xacc[2], &h.xacc[3]};
  for (double** a : zeros)
    if ((rc = make2(*c, a, n2, nullptr))) return rc;
  for (int q = 0; q < 4; ++q) {
    if ((rc = make2(*c, &h.bphi[q], n2, nullptr))) return rc;
    if ((rc = make2(*c, &h.bgra[q], n2, nullptr))) return rc;
  }
  for (int v = 0; v < 6; ++v) {
    if ((rc = make2(*c, &h.bnd[v], n2, nullptr))) return rc;
    if ((rc = make2(*c, &h.frc[v], n2 * 12, hs.frc[v].data()))) return rc;
  }
  if ((rc = make2(*c, &h.graphihx, n2, hs.graphihx.data()))) return rc;
  if ((rc = make2(*c, &h.graphihy, n2, hs.graphihy.data()))) return rc;
  {
    void* p;
    if ((rc = dev_alloc(*c, &p, 4 * (size_t)(c->ni - 2) * (c->nj - 2), sizeof(double)))) return rc;
    c->fstage = static_cast<double*>(p);
  }
  {
    size_t nb = (n2 + 255) / 256;
    if ((rc = make2(*c, &h.red, n2 + 3 * nb + 8, nullptr))) return rc;
    void* p;
    if ((rc = dev_alloc(*c, &p, 1, sizeof(int)))) return rc;
    h.flag = (int*)p;
  }
  c->trop_tw = opts->trop_tile_i;
  c->trop_th = opts->trop_tile_j;
  c->trop_s = opts->trop_substeps;
  c->trop_mode = opts->trop_mode;
  {
    // exchange flags for as many tiles as the device can hold co-resident
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, c->device), "hipGetDeviceProperties");
    c->trop_flag_cap = (long)prop.multiProcessorCount * 8;
  }
  std::string te = trop_configure(*c);
  if (!te.empty()) return fail(PCOM_ERR_HIP, te);
  {
    void* p;
    size_t nflag = (size_t)c->trop_flag_cap * 32 + 32;
    if ((rc = dev_alloc(*c, &p, nflag, sizeof(unsigned long long)))) return rc;
    HIPCHK(hipMemset(p, 0, nflag * sizeof(unsigned long long)), "zero tile flags");
    c->trop_flags = (unsigned long long*)p;
    if ((rc = dev_alloc(*c, &p, 1, sizeof(int)))) return rc;
    HIPCHK(hipMemset(p, 0, sizeof(int)), "zero trop error");
    c->trop_error = (int*)p;
  }
  HIPCHK(hipDeviceSynchronize(), "create sync");
  return PCOM_OK;
}

Ctx* C(pcom_ctx* p) { return reinterpret_cast<Ctx*>(p); }
const Ctx* C(const pcom_ctx* p) { return reinterpret_cast<const Ctx*>(p); }

// device scratch for the single-operator entry points
struct Scratch {
  std::vector<void*> ptrs;
  ~Scratch() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  double* get(size_t n) {
    void* p = nullptr;
    if (hipMalloc(&p, n * sizeof(double) + 8) != hipSuccess) return nullptr;
    ptrs.push_back(p);
    return (double*)p;
  }
};

}  // namespace

extern "C" {

const char* pcom_last_error(void) { return g_last_error.c_str(); }

int pcom_create(const pcom_inputs* in, const pcom_opts* opts, pcom_ctx** out) {
  int rc = create_impl(in, opts, out);
  if (rc != PCOM_OK && out && *out) {
    std::string keep = g_last_error;
    pcom_destroy(*out);
    *out = nullptr;
    g_last_error = keep;
  }
  return rc;
}

int pcom_destroy(pcom_ctx* ctx) {
  if (!ctx) return PCOM_OK;
  Ctx* c = C(ctx);
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->side) (void)hipStreamSynchronize(c->side);  // work left by a failed step
  if (c->stream) trop_chain_forget(c->device, c->stream);
  nn_destroy(*c);
  for (void* p : c->allocs) (void)hipFree(p);
  for (hipEvent_t& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : {c->fork[0], c->fork[1], c->fork[2], c->join[0], c->join[1], c->join[2]})
    if (e) (void)hipEventDestroy(e);
  if (c->side) (void)hipStreamDestroy(c->side);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return PCOM_OK;
}

int pcom_step(pcom_ctx* ctx, int nsteps) {
  if (!ctx || nsteps < 0) return fail(PCOM_ERR_ARG, "pcom_step: bad argument");
  Ctx* c = C(ctx);
  HIPCHK(hipSetDevice(c->device), "hipSetDevice");
  // the last step's side-stream work stays queued: the next step overlaps
  // it, and every entry point that touches the state joins it first
  for (int s = 0; s < nsteps; ++s) {
    std::string e = launch_step(*c);
    if (!e.empty()) return fail(PCOM_ERR_HIP, "pcom_step: " + e);
  }
  return PCOM_OK;
}

int pcom_sync(pcom_ctx* ctx) {
  if (!ctx) return fail(PCOM_ERR_ARG, "pcom_sync: NULL context");
  HIPCHK(hipStreamSynchronize(C(ctx)->stream), "hipStreamSynchronize");
  HIPCHK(hipStreamSynchronize(C(ctx)->side), "hipStreamSynchronize");
  return check_async_error(*C(ctx));
}

int pcom_get_dims(const pcom_ctx* ctx, int* ni, int* nj, int* nk) {
  if (!ctx || !ni || !nj || !nk) return fail(PCOM_ERR_ARG, "pcom_get_dims: NULL argument");
  const Ctx* c = C(ctx);
  *ni = c->ni;
  *nj = c->nj;
  *nk = c->nk;
  return PCOM_OK;
}

size_t pcom_field_size(const pcom_ctx* ctx, int id) {
  if (!ctx) return 0;
  FieldRef r = field_ref(*const_cast<Ctx*>(C(ctx)), id);
  return (r.ptr || r.host) ? r.n : 0;
}

int pcom_get_field(pcom_ctx* ctx, int id, double* host, size_t n) {
  if (!ctx || !host) return fail(PCOM_ERR_ARG, "pcom_get_field: NULL argument");
  Ctx* c = C(ctx);
  if (int rc = quiesce(*c, false)) return rc;
  FieldRef r = field_ref(*c, id);
  if (!r.ptr && !r.host) return fail(PCOM_ERR_ARG, "pcom_get_field: unknown field id " + std::to_string(id));
  if (n != r.n) return fail(PCOM_ERR_ARG, "pcom_get_field: size " + std::to_string(n) + " != " + std::to_string(r.n));
  if (r.host) {
    std::memcpy(host, r.host->data(), n * sizeof(double));
    return PCOM_OK;
  }
  HIPCHK(hipSetDevice(c->device), "hipSetDevice");
  return download(*c, host, r.ptr, n, r.is_float);
}

int pcom_set_field(pcom_ctx* ctx, int id, const double* host, size_t n) {
  if (!ctx || !host) return fail(PCOM_ERR_ARG, "pcom_set_field: NULL argument");
  Ctx* c = C(ctx);
  if (int rc = quiesce(*c, true)) return rc;
  FieldRef r = field_ref(*c, id);
  if (!r.ptr) return fail(PCOM_ERR_ARG, "pcom_set_field: field id " + std::to_string(id) + " is not settable");
  if (n != r.n) return fail(PCOM_ERR_ARG, "pcom_set_field: size " + std::to_string(n) + " != " + std::to_string(r.n));
  HIPCHK(hipSetDevice(c->device), "hipSetDevice");
  HIPCHK(hipStreamSynchronize(c->stream), "sync before upload");
  return upload(*c, r.ptr, host, n, r.is_float);
}

int pcom_get_counters(pcom_ctx* ctx, pcom_counters* k) {
  if (!ctx || !k) return fail(PCOM_ERR_ARG, "pcom_get_counters: NULL argument");
  Ctx* c = C(ctx);
  k->step = c->step_i;
  k->n_ts = c->n_ts; k->n_uv = c->n_uv; k->n_w = c->n_w; k->n_ch = c->n_ch; k->n_ssh = c->n_ssh;
  k->year = c->ct.y; k->month = c->ct.m; k->day = c->ct.d; k->hour = c->ct.h;
  k->minute = c->ct.mi; k->second = c->ct.s;
  return PCOM_OK;
}

int pcom_set_clock(pcom_ctx* ctx, long long step, const int cdate[6], const int pdate[6]) {
  if (!ctx || !cdate || !pdate || step < 0) return fail(PCOM_ERR_ARG, "pcom_set_clock: bad argument");
  Ctx* c = C(ctx);
  if (int rc = quiesce(*c, true)) return rc;
  Calendar t[2];
  const int* d[2] = {cdate, pdate};
  for (int w = 0; w < 2; ++w) {
    const int* x = d[w];
    if (x[0] < 1 || x[1] < 1 || x[1] > 12 || x[2] < 1 || x[2] > days_in_month(x[0], x[1]) || x[3] < 0 ||
        x[3] > 23 || x[4] < 0 || x[4] > 59 || x[5] < 0 || x[5] > 59)
      return fail(PCOM_ERR_ARG, "pcom_set_clock: invalid date");
    t[w].y = x[0]; t[w].m = x[1]; t[w].d = x[2]; t[w].h = x[3]; t[w].mi = x[4]; t[w].s = x[5];
    t[w].mm = days_in_month(x[0], x[1]);
  }
  HIPCHK(hipSetDevice(c->device), "hipSetDevice");
  HIPCHK(hipStreamSynchronize(c->stream), "sync");
  c->step_i = step;
  c->forcing_taken = 0;
  c->ct = t[0];
  c->pt = t[1];
  return PCOM_OK;
}

// the reference's forcing-refresh conditions for step tctr%i = step_i + 1
// (main.f90:164, 184-185)
static void forcing_due(const Ctx& c, int due[2]) {
  const bool first = c.step_i + 1 == 1;
  due[0] = first || c.ct.d != c.pt.d;
  due[1] = first || (c.ct.h != c.pt.h && c.ct.h % 6 == 0);
}

int pcom_forcing_due(pcom_ctx* ctx, int due[2]) {
  if (!ctx || !due) return fail(PCOM_ERR_ARG, "pcom_forcing_due: NULL argument");
  forcing_due(*C(ctx), due);
  return PCOM_OK;
}

int pcom_set_forcing(pcom_ctx* ctx, const pcom_forcing* f, int interpolate) {
  if (!ctx || !f || (interpolate != 0 && interpolate != 1))
    return fail(PCOM_ERR_ARG, "pcom_set_forcing: bad argument");
  Ctx* c = C(ctx);
  int due[2];
  forcing_due(*c, due);
  if (!interpolate && due[0] && c->forcing_taken != c->step_i + 1)
    return fail(PCOM_ERR_STATE, "pcom_set_forcing: the next step's daily refresh (interpolate = 1) comes first");
  if (int rc = quiesce(*c, true)) return rc;
  HIPCHK(hipSetDevice(c->device), "hipSetDevice");
  const size_t g2 = (size_t)(c->ni - 2) * (c->nj - 2);
  // record q of the staging area -> bnd field (taux 0, tauy 1, t 2, pa 4)
  const double* src[4] = {f->taux, f->tauy, f->bct, f->pa};
  const int field[4] = {0, 1, 2, 4};
  const double* rec[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  for (int q = 0; q < 4; ++q)
    if (src[q]) {
      if (int rc = upload(*c, c->fstage + q * g2, src[q], g2, false)) return rc;
      rec[field[q]] = c->fstage + q * g2;
    }
  std::string e = launch_set_forcing(*c, rec, interpolate == 1);
  if (!e.empty()) return fail(PCOM_ERR_HIP, "pcom_set_forcing: " + e);
  if (interpolate) c->forcing_taken = c->step_i + 1;
  return PCOM_OK;
}

// the record counts an output needs (main.f90:726-760)
static int means_ready(const Ctx& c, const pcom_means* out, const char* who) {
  const std::string w(who);
  if ((out->pt || out->sa || out->rho) && c.n_ts == 0) return fail(PCOM_ERR_STATE, w + ": no T/S records");
  if ((out->u || out->v) && c.n_uv == 0) return fail(PCOM_ERR_STATE, w + ": no u/v records");
  if (out->w && c.n_w == 0) return fail(PCOM_ERR_STATE, w + ": no w records");
  if (out->ch && c.n_ch == 0) return fail(PCOM_ERR_STATE, w + ": no ch records");
  if (out->ssh && c.n_ssh == 0) return fail(PCOM_ERR_STATE, w + ": no ssh records");
  return PCOM_OK;
}

int pcom_output_means(pcom_ctx* ctx, pcom_means* out) {
  if (!ctx || !out) return fail(PCOM_ERR_ARG, "pcom_output_means: NULL argument");
  Ctx* c = C(ctx);
  if (int rc = quiesce(*c, false)) return rc;
  if (int rc = means_ready(*c, out, "pcom_output_means")) return rc;
  HIPCHK(hipSetDevice(c->device), "hipSetDevice");
  if (int rc = check_async_error(*c)) return rc;
  if (!c->stage) {
    void* p = nullptr;
    size_t n = (size_t)(c->ni - 2) * (c->nj - 2) * c->nk + 2;
    if (int rc = dev_alloc(*c, &p, n, sizeof(double))) return rc;
    c->stage = static_cast<double*>(p);
  }
  std::string e = launch_output_means(*c, out, c->stage);
  if (!e.empty()) return fail(PCOM_ERR_HIP, "pcom_output_means: " + e);
  return PCOM_OK;
}

}  // extern "C"

// Staging of pcom_output_means_noalloc: one static device array per GPU
// (each device holds its own copy of it for the life of the process), so
// that entry point allocates nothing.  Host code never uses the array's
// name as an address: the device address comes from hipGetSymbolAddress,
// taken once per device with that device current, and kept here.
constexpr size_t kScratchDoubles = (size_t)1 << 21;  // 16 MiB
static __device__ __align__(256) double scratch[kScratchDoubles];
namespace {
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[kScratchDevices] = {};

int scratch_of_device(int device, double** out) {
  if (device < 0 || device >= kScratchDevices) return fail(PCOM_ERR_ARG, "scratch: device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    HIPCHK(hipSetDevice(device), "hipSetDevice");
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!p) return fail(PCOM_ERR_HIP, "hipGetSymbolAddress(scratch): null address");
    g_scratch_ptr[device] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[device];
  return PCOM_OK;
}
}  // namespace

extern "C" {

int pcom_output_means_noalloc(pcom_ctx* ctx, pcom_means* out) {
  if (!ctx || !out) return fail(PCOM_ERR_ARG, "pcom_output_means_noalloc: NULL argument");
  Ctx* c = C(ctx);
  const size_t need = (size_t)(c->ni - 2) * (c->nj - 2) * c->nk + 2;
  if (need > kScratchDoubles)
    return fail(PCOM_ERR_ARG, "pcom_output_means_noalloc: the grid needs " + std::to_string(need) +
                                  " doubles of staging, above the static buffer's " + std::to_string(kScratchDoubles) +
                                  " (glo_ni*glo_nj*nk + 2); use pcom_output_means");
  if (int rc = quiesce(*c, false)) return rc;
  if (int rc = means_ready(*c, out, "pcom_output_means_noalloc")) return rc;
  HIPCHK(hipSetDevice(c->device), "hipSetDevice");
  if (int rc = check_async_error(*c)) return rc;
  double* stage = nullptr;
  if (int rc = scratch_of_device(c->device, &stage)) return rc;
  std::string e = launch_output_means(*c, out, stage);
  if (!e.empty()) return fail(PCOM_ERR_HIP, "pcom_output_means_noalloc: " + e);
  return PCOM_OK;
}

int pcom_ensemble_create(pcom_ctx* const* members, int count, pcom_ensemble** out) {
  if (!members || count < 1 || !out) return fail(PCOM_ERR_ARG, "pcom_ensemble_create: bad argument");
  std::vector<Ctx*> m;
  for (int q = 0; q < count; ++q) {
    if (!members[q]) return fail(PCOM_ERR_ARG, "pcom_ensemble_create: NULL member");
    for (Ctx* x : m)
      if (x == C(members[q])) return fail(PCOM_ERR_ARG, "pcom_ensemble_create: a member is listed twice");
    m.push_back(C(members[q]));
  }
  for (Ctx* x : m)
    if (int rc = quiesce(*x, false)) return rc;
  Ensemble* en = nullptr;
  std::string e = ensemble_create(m, &en);
  if (!e.empty()) return fail(PCOM_ERR_ARG, "pcom_ensemble_create: " + e);
  *out = reinterpret_cast<pcom_ensemble*>(en);
  return PCOM_OK;
}

int pcom_ensemble_step(pcom_ensemble* ens, int nsteps) {
  if (!ens || nsteps < 0) return fail(PCOM_ERR_ARG, "pcom_ensemble_step: bad argument");
  std::string e = ensemble_step(*reinterpret_cast<Ensemble*>(ens), nsteps);
  if (!e.empty()) return fail(PCOM_ERR_HIP, "pcom_ensemble_step: " + e);
  return PCOM_OK;
}

int pcom_ensemble_packs(pcom_ensemble* ens, int* npacks, int* pack_of_member) {
  if (!ens || !npacks || !pack_of_member) return fail(PCOM_ERR_ARG, "pcom_ensemble_packs: NULL argument");
  ensemble_packs(*reinterpret_cast<Ensemble*>(ens), npacks, pack_of_member);
  return PCOM_OK;
}

int pcom_ensemble_destroy(pcom_ensemble* ens) {
  ensemble_destroy(reinterpret_cast<Ensemble*>(ens));
  return PCOM_OK;
}

int pcom_tune_subcycle(pcom_ctx* ctx, int trials, int tile[3], double* ms) {
  if (!ctx || trials < 1 || !tile || !ms) return fail(PCOM_ERR_ARG, "pcom_tune_subcycle: bad argument");
  Ctx* c = C(ctx);
  if (int rc = quiesce(*c, true)) return rc;
  HIPCHK(hipSetDevice(c->device), "hipSetDevice");
  if (int rc = check_async_error(*c)) return rc;
  std::string e = tune_subcycle(*c, trials, tile, ms);
  if (!e.empty()) return fail(PCOM_ERR_HIP, "pcom_tune_subcycle: " + e);
  return PCOM_OK;
}

int pcom_subcycle_tiling(pcom_ctx* ctx, int set, int tile[4]) {
  if (!ctx || !tile) return fail(PCOM_ERR_ARG, "pcom_subcycle_tiling: NULL argument");
  Ctx* c = C(ctx);
  if (int rc = quiesce(*c, false)) return rc;
  if (set) {
    if (tile[0] < 0 || tile[1] < 0 || tile[2] < 0 || tile[0] > 256 || tile[1] > 256 || tile[2] > 16 ||
        tile[3] < PCOM_TROP_AUTO || tile[3] > PCOM_TROP_LAUNCHES)
      return fail(PCOM_ERR_ARG, "pcom_subcycle_tiling: tiling out of range");
    HIPCHK(hipSetDevice(c->device), "hipSetDevice");
    HIPCHK(hipStreamSynchronize(c->stream), "sync");
    const int keep[4] = {c->trop_tw, c->trop_th, c->trop_s, c->trop_mode};
    c->trop_tw = tile[0];
    c->trop_th = tile[1];
    c->trop_s = tile[2];
    c->trop_mode = tile[3];
    std::string e = trop_configure(*c);
    if (!e.empty()) {
      c->trop_tw = keep[0];
      c->trop_th = keep[1];
      c->trop_s = keep[2];
      c->trop_mode = keep[3];
      (void)trop_configure(*c);
      return fail(PCOM_ERR_ARG, "pcom_subcycle_tiling: " + e);
    }
  }
  tile[0] = c->trop_tw;
  tile[1] = c->trop_th;
  tile[2] = c->trop_s;
  tile[3] = c->trop_launches ? PCOM_TROP_LAUNCHES : PCOM_TROP_PERSISTENT;
  return PCOM_OK;
}

int pcom_subcycle_wait_limit(pcom_ctx* ctx, unsigned long long polls) {
  if (!ctx) return fail(PCOM_ERR_ARG, "pcom_subcycle_wait_limit: NULL context");
  Ctx* c = C(ctx);
  if (int rc = quiesce(*c, false)) return rc;
  c->trop_spin_limit = polls;
  return PCOM_OK;
}

int pcom_nn_create(pcom_ctx* ctx, int capacity, unsigned long long seed) {
  if (!ctx || capacity < 0) return fail(PCOM_ERR_ARG, "pcom_nn_create: bad argument");
  Ctx* c = C(ctx);
  if (int rc = quiesce(*c, false)) return rc;
  HIPCHK(hipSetDevice(c->device), "hipSetDevice");
  std::string e = nn_create(*c, capacity, seed);
  if (!e.empty()) {
    nn_destroy(*c);
    return fail(e.find("hip") != std::string::npos ? PCOM_ERR_HIP : PCOM_ERR_ARG, "pcom_nn_create: " + e);
  }
  return PCOM_OK;
}

int pcom_nn_record(pcom_ctx* ctx, int on) {
  if (!ctx) return fail(PCOM_ERR_ARG, "pcom_nn_record: NULL context");
  Ctx* c = C(ctx);
  if (int rc = quiesce(*c, false)) return rc;
  if (!c->nn) return fail(PCOM_ERR_STATE, "pcom_nn_record: no surrogate attached");
  nn_set_record(*c, on != 0);
  return PCOM_OK;
}

int pcom_nn_augment(pcom_ctx* ctx, double fraction, double amp_ch, double amp_u, unsigned long long seed) {
  if (!ctx) return fail(PCOM_ERR_ARG, "pcom_nn_augment: NULL context");
  Ctx* c = C(ctx);
  if (int rc = quiesce(*c, false)) return rc;
  if (!c->nn) return fail(PCOM_ERR_STATE, "pcom_nn_augment: no surrogate attached");
  std::string e = nn_set_augment(*c, fraction, amp_ch, amp_u, seed);
  if (!e.empty()) return fail(PCOM_ERR_ARG, "pcom_nn_augment: " + e);
  return PCOM_OK;
}

int pcom_nn_train(pcom_ctx* ctx, int iters, int batch, double lr, unsigned long long seed, double* loss) {
  if (!ctx || iters < 0 || batch < 1 || !(lr > 0)) return fail(PCOM_ERR_ARG, "pcom_nn_train: bad argument");
  Ctx* c = C(ctx);
  if (int rc = quiesce(*c, false)) return rc;
  if (!c->nn) return fail(PCOM_ERR_STATE, "pcom_nn_train: no surrogate attached");
  HIPCHK(hipSetDevice(c->device), "hipSetDevice");
  std::string e = nn_train(*c, iters, batch, lr, seed, loss);
  if (!e.empty()) return fail(PCOM_ERR_STATE, "pcom_nn_train: " + e);
  return PCOM_OK;
}

int pcom_nn_use(pcom_ctx* ctx, int check_every, double tol) {
  if (!ctx || check_every < 0 || !(tol == tol)) return fail(PCOM_ERR_ARG, "pcom_nn_use: bad argument");
  Ctx* c = C(ctx);
  if (int rc = quiesce(*c, false)) return rc;
  if (!c->nn) return fail(PCOM_ERR_STATE, "pcom_nn_use: no surrogate attached");
  std::string e = nn_set_use(*c, check_every, tol);
  if (!e.empty()) return fail(PCOM_ERR_STATE, "pcom_nn_use: " + e);
  return PCOM_OK;
}

int pcom_nn_resume(pcom_ctx* ctx, int plain_steps) {
  if (!ctx || plain_steps < 0) return fail(PCOM_ERR_ARG, "pcom_nn_resume: bad argument");
  Ctx* c = C(ctx);
  if (int rc = quiesce(*c, false)) return rc;
  if (!c->nn) return fail(PCOM_ERR_STATE, "pcom_nn_resume: no surrogate attached");
  std::string e = nn_set_resume(*c, plain_steps);
  if (!e.empty()) return fail(PCOM_ERR_STATE, "pcom_nn_resume: " + e);
  return PCOM_OK;
}

int pcom_nn_eval(pcom_ctx* ctx, int first, int count, double rmse[8]) {
  if (!ctx || !rmse) return fail(PCOM_ERR_ARG, "pcom_nn_eval: NULL argument");
  Ctx* c = C(ctx);
  if (int rc = quiesce(*c, false)) return rc;
  if (!c->nn) return fail(PCOM_ERR_STATE, "pcom_nn_eval: no surrogate attached");
  HIPCHK(hipSetDevice(c->device), "hipSetDevice");
  std::string e = nn_eval(*c, first, count, rmse);
  if (!e.empty()) return fail(PCOM_ERR_STATE, "pcom_nn_eval: " + e);
  return PCOM_OK;
}

int pcom_nn_get_stats(pcom_ctx* ctx, pcom_nn_stats* s) {
  if (!ctx || !s) return fail(PCOM_ERR_ARG, "pcom_nn_get_stats: NULL argument");
  nn_stats(*C(ctx), s);
  return PCOM_OK;
}

size_t pcom_nn_param_count(const pcom_ctx* ctx) {
  if (!ctx) return 0;
  const Ctx* c = C(ctx);
  return c->nn ? nn_param_count(*c) : 0;
}

int pcom_nn_get_params(pcom_ctx* ctx, float* host, size_t n) {
  if (!ctx || !host) return fail(PCOM_ERR_ARG, "pcom_nn_get_params: NULL argument");
  Ctx* c = C(ctx);
  if (int rc = quiesce(*c, false)) return rc;
  if (!c->nn) return fail(PCOM_ERR_STATE, "pcom_nn_get_params: no surrogate attached");
  if (n != pcom_nn_param_count(ctx)) return fail(PCOM_ERR_ARG, "pcom_nn_get_params: size mismatch");
  HIPCHK(hipSetDevice(c->device), "hipSetDevice");
  std::string e = nn_params(*c, host, false);
  if (!e.empty()) return fail(PCOM_ERR_HIP, "pcom_nn_get_params: " + e);
  return PCOM_OK;
}

int pcom_nn_set_params(pcom_ctx* ctx, const float* host, size_t n) {
  if (!ctx || !host) return fail(PCOM_ERR_ARG, "pcom_nn_set_params: NULL argument");
  Ctx* c = C(ctx);
  if (int rc = quiesce(*c, false)) return rc;
  if (!c->nn) return fail(PCOM_ERR_STATE, "pcom_nn_set_params: no surrogate attached");
  if (n != pcom_nn_param_count(ctx)) return fail(PCOM_ERR_ARG, "pcom_nn_set_params: size mismatch");
  HIPCHK(hipSetDevice(c->device), "hipSetDevice");
  std::string e = nn_params(*c, const_cast<float*>(host), true);
  if (!e.empty()) return fail(PCOM_ERR_HIP, "pcom_nn_set_params: " + e);
  return PCOM_OK;
}

int pcom_nn_forward_backward(pcom_ctx* ctx, int batch, const float* x, const float* target, float* out, float* grads,
                             double* loss) {
  if (!ctx || batch < 1 || !x || !target || !out || !grads || !loss)
    return fail(PCOM_ERR_ARG, "pcom_nn_forward_backward: bad argument");
  Ctx* c = C(ctx);
  if (int rc = quiesce(*c, false)) return rc;
  if (!c->nn) return fail(PCOM_ERR_STATE, "pcom_nn_forward_backward: no surrogate attached");
  HIPCHK(hipSetDevice(c->device), "hipSetDevice");
  std::string e = nn_forward_backward(*c, batch, x, target, out, grads, loss);
  if (!e.empty()) return fail(PCOM_ERR_HIP, "pcom_nn_forward_backward: " + e);
  return PCOM_OK;
}

int pcom_get_metric(const pcom_ctx* ctx, int g, int which, double* out, size_t n) {
  if (!ctx || !out) return fail(PCOM_ERR_ARG, "pcom_get_metric: NULL argument");
  const Ctx* c = C(ctx);
  if (g < 1 || g > 4 || which < 0 || which > 3) return fail(PCOM_ERR_ARG, "pcom_get_metric: bad stagger or metric id");
  if (n != c->n2) return fail(PCOM_ERR_ARG, "pcom_get_metric: size mismatch");
  const std::vector<double>& v = which == 0 ? c->hs.rh[g] : which == 1 ? c->hs.tn[g] : which == 2 ? c->hs.dx1[g] : c->hs.dx2[g];
  std::memcpy(out, v.data(), n * sizeof(double));
  return PCOM_OK;
}

int pcom_global_sums(pcom_ctx* ctx, double out[3]) {
  if (!ctx || !out) return fail(PCOM_ERR_ARG, "pcom_global_sums: NULL argument");
  Ctx* c = C(ctx);
  if (int rc = quiesce(*c, false)) return rc;
  HIPCHK(hipSetDevice(c->device), "hipSetDevice");
  std::string e = launch_global_sums(*c, out);
  if (!e.empty()) return fail(PCOM_ERR_HIP, e);
  return PCOM_OK;
}

int pcom_check_finite(pcom_ctx* ctx) {
  if (!ctx) return fail(PCOM_ERR_ARG, "pcom_check_finite: NULL context");
  Ctx* c = C(ctx);
  if (int rc = quiesce(*c, false)) return rc;
  HIPCHK(hipSetDevice(c->device), "hipSetDevice");
  if (int rc = check_async_error(*c)) return rc;
  int bad = 0;
  std::string e = launch_check_finite(*c, &bad);
  if (!e.empty()) return fail(PCOM_ERR_HIP, e);
  if (bad) return fail(PCOM_ERR_NAN, "non-finite value in T, S, u, v, ch or the barotropic velocity");
  return PCOM_OK;
}

int pcom_phase_times(pcom_ctx* ctx, int enable, double out_ms[3], int* nsteps) {
  if (!ctx) return fail(PCOM_ERR_ARG, "pcom_phase_times: NULL context");
  Ctx* c = C(ctx);
  if (int rc = quiesce(*c, false)) return rc;
  if (out_ms) {
    for (int q = 0; q < 3; ++q) out_ms[q] = c->t_steps ? c->t_ms[q] / c->t_steps : 0.0;
  }
  if (nsteps) *nsteps = c->t_steps;
  if (enable && !c->ev[0]) {
    for (hipEvent_t& e : c->ev) HIPCHK(hipEventCreate(&e), "hipEventCreate");
  }
  c->timing = enable != 0;
  for (double& t : c->t_ms) t = 0;
  c->t_steps = 0;
  return PCOM_OK;
}

// ---- single operators ------------------------------------------------------
static int op_common(pcom_ctx* ctx, Ctx** c) {
  if (!ctx) return fail(PCOM_ERR_ARG, "operator entry point: NULL context");
  *c = C(ctx);
  HIPCHK(hipSetDevice((*c)->device), "hipSetDevice");
  return PCOM_OK;
}

int pcom_op_ter(pcom_ctx* ctx, int ga, int gb, int nd3, double dft, const double* in, double* out) {
  Ctx* c;
  int rc = op_common(ctx, &c);
  if (rc) return rc;
  if (!in || !out || nd3 < 1) return fail(PCOM_ERR_ARG, "pcom_op_ter: bad argument");
  size_t n = c->n2 * (size_t)nd3;
  Scratch s;
  double *di = s.get(n), *dout = s.get(n);
  if (!di || !dout) return fail(PCOM_ERR_HIP, "pcom_op_ter: hipMalloc failed");
  HIPCHK(hipMemcpy(di, in, n * sizeof(double), hipMemcpyHostToDevice), "upload");
  std::string e = launch_op_ter(*c, ga, gb, nd3, dft, di, dout);
  if (!e.empty()) return fail(PCOM_ERR_ARG, e);
  HIPCHK(hipStreamSynchronize(c->stream), "sync");
  HIPCHK(hipMemcpy(out, dout, n * sizeof(double), hipMemcpyDeviceToHost), "download");
  return PCOM_OK;
}

static int op_px(pcom_ctx* ctx, int dir, int ga, int gb, int nd3, const double* in, double* out) {
  Ctx* c;
  int rc = op_common(ctx, &c);
  if (rc) return rc;
  if (!in || !out || nd3 < 1) return fail(PCOM_ERR_ARG, "pcom_op_px: bad argument");
  size_t n = c->n2 * (size_t)nd3;
  Scratch s;
  double *di = s.get(n), *dout = s.get(n);
  if (!di || !dout) return fail(PCOM_ERR_HIP, "pcom_op_px: hipMalloc failed");
  HIPCHK(hipMemcpy(di, in, n * sizeof(double), hipMemcpyHostToDevice), "upload");
  std::string e = launch_op_px(*c, dir, ga, gb, nd3, di, dout);
  if (!e.empty()) return fail(PCOM_ERR_ARG, e);
  HIPCHK(hipStreamSynchronize(c->stream), "sync");
  HIPCHK(hipMemcpy(out, dout, n * sizeof(double), hipMemcpyDeviceToHost), "download");
  return PCOM_OK;
}

int pcom_op_px1(pcom_ctx* ctx, int ga, int gb, int nd3, const double* in, double* out) {
  return op_px(ctx, 1, ga, gb, nd3, in, out);
}
int pcom_op_px2(pcom_ctx* ctx, int ga, int gb, int nd3, const double* in, double* out) {
  return op_px(ctx, 2, ga, gb, nd3, in, out);
}

int pcom_op_div(pcom_ctx* ctx, int ga, int gb, int gc, int nd3, const double* va, const double* vb, double* out) {
  Ctx* c;
  int rc = op_common(ctx, &c);
  if (rc) return rc;
  if (!va || !vb || !out || nd3 < 1) return fail(PCOM_ERR_ARG, "pcom_op_div: bad argument");
  size_t n = c->n2 * (size_t)nd3;
  Scratch s;
  double *da = s.get(n), *db = s.get(n), *dout = s.get(n);
  if (!da || !db || !dout) return fail(PCOM_ERR_HIP, "pcom_op_div: hipMalloc failed");
  HIPCHK(hipMemcpy(da, va, n * sizeof(double), hipMemcpyHostToDevice), "upload");
  HIPCHK(hipMemcpy(db, vb, n * sizeof(double), hipMemcpyHostToDevice), "upload");
  std::string e = launch_op_div(*c, ga, gb, gc, nd3, da, db, dout);
  if (!e.empty()) return fail(PCOM_ERR_ARG, e);
  HIPCHK(hipStreamSynchronize(c->stream), "sync");
  HIPCHK(hipMemcpy(out, dout, n * sizeof(double), hipMemcpyDeviceToHost), "download");
  return PCOM_OK;
}

int pcom_op_lap(pcom_ctx* ctx, int ga, int gb, const double* in, double* out) {
  Ctx* c;
  int rc = op_common(ctx, &c);
  if (rc) return rc;
  if (!in || !out) return fail(PCOM_ERR_ARG, "pcom_op_lap: bad argument");
  size_t n = c->n2;
  Scratch s;
  double *di = s.get(n), *dout = s.get(n);
  if (!di || !dout) return fail(PCOM_ERR_HIP, "pcom_op_lap: hipMalloc failed");
  HIPCHK(hipMemcpy(di, in, n * sizeof(double), hipMemcpyHostToDevice), "upload");
  std::string e = launch_op_lap(*c, ga, gb, di, dout);
  if (!e.empty()) return fail(PCOM_ERR_ARG, e);
  HIPCHK(hipStreamSynchronize(c->stream), "sync");
  HIPCHK(hipMemcpy(out, dout, n * sizeof(double), hipMemcpyDeviceToHost), "download");
  return PCOM_OK;
}

int pcom_den_rho(pcom_ctx* ctx, size_t n, const double* t, const double* s, const double* p, double* rho) {
  if (!t || !s || !p || !rho) return fail(PCOM_ERR_ARG, "pcom_den_rho: NULL argument");
  if (n == 0) return PCOM_OK;
  hipStream_t st = nullptr;
  if (ctx) {
    HIPCHK(hipSetDevice(C(ctx)->device), "hipSetDevice");
    st = C(ctx)->stream;
  } else {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(PCOM_ERR_NOGPU, "no HIP device visible");
  }
  Scratch sc;
  double *dt = sc.get(n), *ds = sc.get(n), *dp = sc.get(n), *dr = sc.get(n);
  if (!dt || !ds || !dp || !dr) return fail(PCOM_ERR_HIP, "pcom_den_rho: hipMalloc failed");
  HIPCHK(hipMemcpy(dt, t, n * sizeof(double), hipMemcpyHostToDevice), "upload");
  HIPCHK(hipMemcpy(ds, s, n * sizeof(double), hipMemcpyHostToDevice), "upload");
  HIPCHK(hipMemcpy(dp, p, n * sizeof(double), hipMemcpyHostToDevice), "upload");
  std::string e = launch_den_rho(st, n, dt, ds, dp, dr);
  if (!e.empty()) return fail(PCOM_ERR_HIP, e);
  HIPCHK(hipStreamSynchronize(st), "sync");
  HIPCHK(hipMemcpy(rho, dr, n * sizeof(double), hipMemcpyDeviceToHost), "download");
  return PCOM_OK;
}

}  // extern "C"
