// kernels.hip - CDNA4 kernels of one PCOM baroclinic step (main.f90:158-227).
//
// Each kernel evaluates one or more reference routines per grid point with
// the pointwise operators of pcom_dev.h, so intermediates the reference
// materialises as whole arrays (and halo-swaps) stay in registers / cache.
// Precision forms: the prognostic state T, S, u, v is FP64 in every form;
// 3-D tendency stencils compute in R and store their fields in R; the
// equation of state and the pressure integral (alpha, adp, prho, the
// face integrals of int_pgra, the pressure gradient of int_clin, int_ssh)
// compute in E.  FP64: R = E = double.  PCOM_MIXED: R = float, E = double.
// PCOM_MIXED_EOS32: R = E = float.  The barotropic subcycle is FP64 always.
//
// Thread mapping: point kernels use one thread per (i, j, k) with i fastest
// (coalesced along the Fortran-contiguous dimension).  The stencil kernels
// whose operators are divergences of face fluxes (int_ts, op_fri, op_adv,
// smag) run on 64 x 4 tiles of canonical columns per level and stage each
// face value in LDS once, instead of once per cell beside it.  Column
// kernels give each column several k-lanes that evaluate the per-level
// terms in parallel; one lane then walks k in the reference's summation
// order, so every column sum is accumulated exactly as Fortran's sum()
// does.  A second stream runs int_ssh, int_ts, convect and int_pgra beside
// smag and the next step's int_readyc (step_pre_t, step_post_t).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>
#include <type_traits>

#include "context.h"

namespace pcom {

namespace {

constexpr int kBlock = 256;
// column kernels: workgroups of kColCW columns x kColKL k-lanes (the
// column kernels are bound by load latency: more lanes per column keep
// more loads in flight), with [nk][kColCW] LDS arrays
constexpr int kColCW = 32, kColKL = 8;

inline std::string hip_err(hipError_t e, const char* what) {
  if (e == hipSuccess) return "";
  return std::string(what) + ": " + hipGetErrorString(e);
}

inline unsigned nblocks(size_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// decode a point index into 1-based (i, j, k)
template <typename G>
__device__ __forceinline__ void ijk(const G& g, long p, int& i, int& j, int& k) {
  i = (int)(p % g.ni) + 1;
  long q = p / g.ni;
  j = (int)(q % g.nj) + 1;
  k = (int)(q / g.nj) + 1;
}

// ---------------------------------------------------------------------------
// daily forcing: int_bnd (mod_int.f90:678-689)
struct Six {
  const double* a[6];
  double* b[6];
};

__global__ void k_bnd(long n2, Six s, int m1, int m2, double f) {
  long p = (long)xcd_block() * blockDim.x + threadIdx.x;
  if (p >= n2) return;
#pragma unroll
  for (int v = 0; v < 6; ++v) {
    double a1 = s.a[v][p + n2 * (m1 - 1)], a2 = s.a[v][p + n2 * (m2 - 1)];
    s.b[v][p] = (a2 - a1) * f + a1;
  }
}

// one global record of a forcing file into a halo'd bnd field, as
// mympi_input does (mod_mympi.f90:638-657, 898-928): interior points from
// the record, the wall rows 1 and nj kept, then the E/W swap (halo columns
// of the wall rows copy interior columns of those rows, which no thread
// writes)
__global__ void k_put_forcing(Grid<double> g, const double* __restrict__ rec, double* bnd) {
  long p = (long)xcd_block() * blockDim.x + threadIdx.x;
  if (p >= g.n2) return;
  int i = (int)(p % g.ni) + 1, j = (int)(p / g.ni) + 1;
  const int ic = i == 1 ? g.ni - 1 : (i == g.ni ? 2 : i);
  if (j >= 2 && j <= g.nj - 1)
    bnd[p] = rec[(long)(ic - 2) + (long)(g.ni - 2) * (j - 2)];
  else if (ic != i)
    bnd[p] = bnd[g.ix(ic, j)];
}

// atmospheric pressure gradient (mod_int.f90:100-103): gra_r2d_b(pa, g1)
// then ter to g3.  The reference forms it every step from bnd%pa, which
// changes only with the daily forcing or a caller's set: it runs then.
__global__ void k_pa(Grid<double> g, const double* __restrict__ pa, double* pax, double* pay) {
  long p = (long)xcd_block() * blockDim.x + threadIdx.x;
  if (p >= g.n2) return;
  int i = (int)(p % g.ni) + 1, j = (int)(p / g.ni) + 1;
  auto PA = [&](int m, int jj) -> double { return pa[g.ix(m, jj)]; };
  auto WX = [&](int m, int jj) -> double { return grax<1, 2>(g, PA, m, jj); };
  auto WY = [&](int m, int jj) -> double { return gray<1, 4>(g, PA, m, jj); };
  pax[p] = ter_ns<2>(g, WX, i, j, 0.0);
  pay[p] = ter_ew<4>(g, WY, i, j);
}

// sqrt of ch interpolated to the U points, the factor every 3-D operator of
// the step divides or multiplies by.  Computed once per use in FP64:
//  before int_readyc: schr = sqrt(ter(chc,g1->g3,dft=1)), 1 on land
//                     (mod_int.f90:107-109); sqp = sqrt(ter(chp)) (upwelling,
//                     mod_int.f90:636-637);
//  after int_trop:    schs = sqrt(ter(chc, dft=1)) (op_adv_ts, mod_op.f90:
//                     294-295; equal to int_clin's sqrt(ter(mch2, dft=1)) since
//                     chc = mch2 there, mod_int.f90:284, 385-386); sqp again;
//                     schm = sqrt(ter(mch, dft=1)), 1 on land (mod_int.f90:318-320).
__global__ void k_sch(Grid<double> g, const double* __restrict__ chc, const double* __restrict__ chp,
                      const double* __restrict__ mch, double* schr, double* sqp, double* schs, double* schm) {
  long p = (long)xcd_block() * blockDim.x + threadIdx.x;
  if (p >= g.n2) return;
  int i = (int)(p % g.ni) + 1, j = (int)(p / g.ni) + 1;
  auto CHC = [&](int m, int jj) -> double { return chc[g.ix(m, jj)]; };
  auto CHP = [&](int m, int jj) -> double { return chp[g.ix(m, jj)]; };
  const bool land = g.lev_u[p] == 0;
  double c1 = sqrt(ter_di<1>(g, CHC, i, j, 1.0));
  if (schr) schr[p] = land ? 1.0 : c1;
  if (schs) schs[p] = c1;
  sqp[p] = sqrt(ter_di<1>(g, CHP, i, j, 0.0));
  if (schm) {
    auto MCH = [&](int m, int jj) -> double { return mch[g.ix(m, jj)]; };
    schm[p] = land ? 1.0 : sqrt(ter_di<1>(g, MCH, i, j, 1.0));
  }
}

// den_prho (mod_den.f90:120-153)
template <typename R>
__global__ void k_prho(Grid<R> g, const double* __restrict__ tc, const double* __restrict__ sc,
                       const double* __restrict__ chc, R* prho1, R* prho2) {
  long p = (long)xcd_block() * blockDim.x + threadIdx.x;
  if (p >= g.n2 * g.nk) return;
  int i, j, k;
  ijk(g, p, i, j, k);
  if (g.mskt(i, j, k) > 0) {
    R t = (R)tc[p], s = (R)sc[p];
    R pr = (R)chc[g.ix(i, j)] * g.pr2[k - 1];
    R ra = den_rho<R>(t + (R)DELTAT, s, pr);
    R rb = den_rho<R>(t - (R)DELTAT, s, pr);
    prho1[p] = (ra - rb) / ((R)2 * (R)DELTAT);
    ra = den_rho<R>(t, s + (R)DELTAS, pr);
    rb = den_rho<R>(t, s - (R)DELTAS, pr);
    prho2[p] = (ra - rb) / ((R)2 * (R)DELTAS);
  }
}

// den_alpha (mod_den.f90:33-55): alpha = 1/rho where mask == 1; land is
// left as it is (1/1000.0 from mod_arrays.f90:298).
template <typename R>
__global__ void k_alpha(Grid<R> g, const double* __restrict__ t, const double* __restrict__ s,
                        const double* __restrict__ ch, R* alpha) {
  long p = (long)xcd_block() * blockDim.x + threadIdx.x;
  if (p >= g.n2 * g.nk) return;
  int i, j, k;
  ijk(g, p, i, j, k);
  if (g.mskt(i, j, k) == 1) alpha[p] = (R)1.0 / den_rho<R>((R)t[p], (R)s[p], (R)ch[g.ix(i, j)] * g.pr2[k - 1]);
}

// int_tobot (mod_op.f90:927-944): adp(k) = t(k)/2 + sum(t(k+1:nk)),
// t = alpha*dpr*msk; the suffix sum is accumulated in increasing k as
// Fortran's sum() does, so FP64 results are the reference's bits.  t holds
// the column's terms at a stride of `st` between levels.
template <typename R>
__device__ __forceinline__ R tobot(const R* t, int st, int k, int nk) {
  if constexpr (sizeof(R) == sizeof(double)) {
    R s = 0;
    for (int kk = k + 1; kk <= nk; ++kk) s += t[st * (kk - 1)];
    return t[st * (k - 1)] * (R)0.5 + s;
  } else {
    double s = 0;
    for (int kk = k + 1; kk <= nk; ++kk) s += (double)t[st * (kk - 1)];
    return (R)((double)t[st * (k - 1)] * 0.5 + s);
  }
}

template <typename R>
__global__ void __launch_bounds__(kColCW * kColKL) k_adp(Grid<R> g, const R* __restrict__ alpha, R* adp) {
  extern __shared__ unsigned char adp_lds[];
  R(*t)[kColCW] = reinterpret_cast<R(*)[kColCW]>(adp_lds);  // [nk][kColCW]
  const int tx = threadIdx.x, kl = threadIdx.y, nk = g.nk;
  const long p = (long)xcd_block() * kColCW + tx;
  if (p < g.n2) {
    const int lev = g.lev_t[p];
    for (int k = 1 + kl; k <= nk; k += kColKL)
      t[k - 1][tx] = alpha[p + g.n2 * (k - 1)] * g.dpr2[k - 1] * (R)(lev >= k ? 1 : 0);
  }
  __syncthreads();
  if (p >= g.n2) return;
  for (int k = 1 + kl; k <= nk; k += kColKL) adp[p + g.n2 * (k - 1)] = tobot(&t[0][tx], kColCW, k, nk);
}

// int_pgra's face integrals (mod_int.f90:72-86 with op_vint_ns / op_vint_ew,
// mod_op.f90:946-981): bphi = vint(ter(adp) + pr*ter(alpha)),
// bgraphi = vint(gra(adp)), each with the local mask and the mask cshift'ed
// by -1 (row 1 wraps to row nj; the E/W wrap is overwritten by the swap).
// The lanes evaluate the per-level terms of the columns in parallel into
// LDS, one lane per column then sums them in k order (the reference's
// order).

template <typename R>
__global__ void __launch_bounds__(kColCW * kColKL)
    k_bphi(Grid<R> g, const R* __restrict__ adp, const R* __restrict__ alpha, double* b0, double* b1, double* b2,
           double* b3, double* c0, double* c1, double* c2, double* c3) {
  extern __shared__ unsigned char bphi_lds[];
  const int tx = threadIdx.x, kl = threadIdx.y, nk = g.nk;
  R* tv2 = reinterpret_cast<R*>(bphi_lds);  // [nk][kColCW] each
  R* tv4 = tv2 + nk * kColCW;
  R* tvx = tv4 + nk * kColCW;
  R* tvy = tvx + nk * kColCW;
  const long p = (long)xcd_block() * kColCW + tx;
  const bool ok = p < g.n2;
  const int i = ok ? (int)(p % g.ni) + 1 : 2, j = ok ? (int)(p / g.ni) + 1 : 1;
  const int ic = g.canon(i);
  const int js = (j == 1) ? g.nj : j - 1;
  if (ok) {
    for (int k = 1 + kl; k <= nk; k += kColKL) {
      const R* ad = adp + g.n2 * (k - 1);
      const R* al = alpha + g.n2 * (k - 1);
      auto AD = [&](int m, int jj) -> R { return ad[g.ix(m, jj)]; };
      auto AL = [&](int m, int jj) -> R { return al[g.ix(m, jj)]; };
      R pr = g.pr2[k - 1], dpr = g.dpr2[k - 1];
      R wk2 = ter_ew<1>(g, AD, ic, j) + pr * ter_ew<1>(g, AL, ic, j);
      R wk4 = ter_ns<1>(g, AD, ic, j, (R)0) + pr * ter_ns<1>(g, AL, ic, j, (R)0);
      R wx = grax<1, 2>(g, AD, ic, j);
      R wy = gray<1, 4>(g, AD, ic, j);
      const int o = (k - 1) * kColCW + tx;
      tv2[o] = wk2 * dpr;
      tv4[o] = wk4 * dpr;
      tvx[o] = wx * dpr;
      tvy[o] = wy * dpr;
    }
  }
  __syncthreads();
  if (!ok || kl != 0) return;
  const int lu = g.lev_u[g.ix(ic, j)], lus = g.lev_u[g.ix(ic, js)], luw = g.lev_u[g.ix(ic - 1, j)];
  double sxn = 0, sxs = 0, sye = 0, syw = 0, gxn = 0, gxs = 0, gye = 0, gyw = 0;
  for (int k = 1; k <= nk; ++k) {
    const int o = (k - 1) * kColCW + tx;
    const R v2 = tv2[o], v4 = tv4[o], vx = tvx[o], vy = tvy[o];
    R mu = (R)(lu >= k ? 1 : 0), mus = (R)(lus >= k ? 1 : 0), muw = (R)(luw >= k ? 1 : 0);
    sxn += (double)(v2 * mu);
    sxs += (double)(v2 * mus);
    sye += (double)(v4 * mu);
    syw += (double)(v4 * muw);
    gxn += (double)(vx * mu);
    gxs += (double)(vx * mus);
    gye += (double)(vy * mu);
    gyw += (double)(vy * muw);
  }
  double ph = g.prh_u[g.ix(ic, j)], phs = g.prh_u[g.ix(ic, js)], phw = g.prh_u[g.ix(ic - 1, j)];
  b0[p] = sxn / ph;
  b1[p] = sxs / phs;
  b2[p] = sye / ph;
  b3[p] = syw / phw;
  c0[p] = gxn / ph;
  c1[p] = gxs / phs;
  c2[p] = gye / ph;
  c3[p] = gyw / phw;
}

// upwelling (mod_int.f90:626-647): wm (levels 2..nk) from the continuity
// equation
template <typename R>
__global__ void __launch_bounds__(kColCW * kColKL) k_upw(Grid<R> g, const double* __restrict__ uc,
                                                         const double* __restrict__ vc,
                                                         const double* __restrict__ sqp, R* wm) {
  // one LDS column array [nk][kColCW]: the divergence per level, then in
  // place the upwelling
  extern __shared__ unsigned char upw_lds[];
  R(*w3)[kColCW] = reinterpret_cast<R(*)[kColCW]>(upw_lds);
  const int tx = threadIdx.x, kl = threadIdx.y, nk = g.nk;
  const long p = (long)xcd_block() * kColCW + tx;
  const bool ok = p < g.n2;
  const int i = ok ? (int)(p % g.ni) + 1 : 1, j = ok ? (int)(p / g.ni) + 1 : 1;
  auto SQ = [&](int m, int jj) -> R { return (R)sqp[g.ix(m, jj)]; };
  if (ok) {
    for (int k = 1 + kl; k <= nk; k += kColKL) {
      const double* u = uc + g.n2 * (k - 1);
      const double* v = vc + g.n2 * (k - 1);
      auto VA = [&](int m, int jj) -> R { return SQ(m, jj) * (R)u[g.ix(m, jj)]; };
      auto VB = [&](int m, int jj) -> R { return SQ(m, jj) * (R)v[g.ix(m, jj)]; };
      w3[k - 1][tx] = divg<3, 3, 1>(g, VA, VB, i, j);
    }
  }
  __syncthreads();
  if (ok && kl == 0) {
    const int lev = g.lev_t[p];
    R s = 0;
    for (int k = 1; k <= nk; ++k) s += w3[k - 1][tx] * g.dpr2[k - 1] * (R)(lev >= k ? 1 : 0);
    R wk2 = -dv(s, g.prh_t_r[p]);
    R w = wm[p];  // wm(1) is never written by upwelling
    R d = w3[0][tx];
    w3[0][tx] = w;
    for (int k = 2; k <= nk; ++k) {
      w = w + g.dpr2[k - 2] * (wk2 + d);  // d = the divergence of level k-1
      w = w * (R)(lev >= k ? 1 : 0);
      d = w3[k - 1][tx];
      w3[k - 1][tx] = w;
    }
  }
  __syncthreads();
  if (!ok) return;
  for (int k = 2 + kl; k <= nk; k += kColKL) wm[p + g.n2 * (k - 1)] = w3[k - 1][tx];
}

// int_readyc's sch (mod_int.f90:107-109), precomputed by k_sch
template <typename R>
__device__ __forceinline__ R sch_ready(const Grid<R>& g, const double* schr, int m, int jj) {
  return (R)schr[g.ix(m, jj)];
}

// op_adv (mod_op.f90:224-278) for uc and vc, with the face quantities
// staged in LDS.  On the east-west faces of a U cell (g4 points) live the
// transport UU = ter(u/sch) and the fluxes M1 = ter(var)*UU, on the
// north-south faces (g2 points) VV = ter(v/sch*rh) and M2 = ter(var)*VV;
// each face is shared by the two cells beside it.  A workgroup owns a
// kAdvTX x kAdvTY tile of canonical columns at one level, evaluates every
// face once, then finishes each cell (flux divergence, vertical flux,
// -var*div(v)/2) with the same operations in the same order as the
// per-cell composition.  Halo columns take their faces from their
// canonical columns and everything else from themselves.
constexpr int kAdvTX = 64, kAdvTY = 4;

template <typename R>
struct AdvCellIn {
  R wk[2][4], ch[4];  // wm at k, k+1 and chc at T(ic, j), (ic, j+1), (ic+1, j), (ic+1, j+1)
  double var[2][3];   // uc, vc at k-1, k, k+1 (0 outside 1..nk)
  int lev;
  R rh3, dx13, dx23;
};

template <typename R>
__global__ void __launch_bounds__(kAdvTX * kAdvTY)
    k_adv_tiles(Grid<R> g, const double* __restrict__ uc, const double* __restrict__ vc, const R* __restrict__ wm,
                const double* __restrict__ chc, const double* __restrict__ schr, R* auc, R* avc) {
  const int G = g.ni - 2;
  const int ntx = (G + kAdvTX - 1) / kAdvTX;
  const int b = (int)xcd_block();
  // levels fastest: a tile's levels run together on one XCD, whose L2 then
  // serves the vertical neighbours each level reads
  const int k = b % g.nk + 1;
  const int r = b / g.nk;
  const int c0 = 2 + (r % ntx) * kAdvTX, j0 = 1 + (r / ntx) * kAdvTY;
  const int tx = threadIdx.x % kAdvTX, ty = threadIdx.x / kAdvTX;
  const int ic = c0 + tx, j = j0 + ty;
  const long ko = g.n2 * (k - 1);
  const double* u = uc + ko;
  const double* v = vc + ko;
  // east-west faces [row][f] at g4 points c0 + f (f is cell f's west
  // face); north-south faces [row][tx] at g2 points of row j0 + row (row
  // ty is cell ty's south face)
  __shared__ R fuu[kAdvTY][kAdvTX + 1], fm1[2][kAdvTY][kAdvTX + 1];
  __shared__ R fvv[kAdvTY + 1][kAdvTX], fm2[2][kAdvTY + 1][kAdvTX];
  auto SCH = [&](int m, int jj) -> R { return (R)schr[g.ix(m, jj)]; };
  auto X = [&](int m, int jj) -> R { return dv((R)u[g.ix(m, jj)], SCH(m, jj)); };
  auto Y = [&](int m, int jj) -> R {
    R y = dv((R)v[g.ix(m, jj)], SCH(m, jj));
    return y * g.RH(3, m, jj);
  };
  auto VK = [&](int c, int m, int jj) -> R { return (R)(c ? v : u)[g.ix(m, jj)]; };
  // ter_ew<3> (g3 -> g4) at g4 point (m, jj), m canonical
  auto face_ew = [&](int m, int jj, int row, int f) {
    const R uu = (R)0.5 * (X(m - 1, jj) + X(m, jj));
    fuu[row][f] = uu;
    for (int c = 0; c < 2; ++c) fm1[c][row][f] = (R)0.5 * (VK(c, m - 1, jj) + VK(c, m, jj)) * uu;
  };
  // ter_ns<3> (g3 -> g2, rows 2..nj) at g2 point (m, jj)
  auto face_ns = [&](int m, int jj, int row, int f) {
    const bool w = jj >= 2 && jj <= g.nj;
    const R vv = w ? (R)0.5 * (Y(m, jj - 1) + Y(m, jj)) : (R)0;
    fvv[row][f] = vv;
    for (int c = 0; c < 2; ++c) fm2[c][row][f] = (w ? (R)0.5 * (VK(c, m, jj - 1) + VK(c, m, jj)) : (R)0) * vv;
  };
  // the inputs of cell (i, j) besides the LDS faces (the thread's own cell
  // loads them before the face phase, so both rounds of loads overlap)
  const bool rows = j >= 1 && j <= g.njm;  // ter_di's and p2's rows
  auto cell_in = [&](int i) -> AdvCellIn<R> {
    AdvCellIn<R> c;
    const int m0 = g.canon(i);
    for (int o = 0; o < 2; ++o) {
      const R* w = wm + g.n2 * (k - 1 + o);
      for (int e = 0; e < 4; ++e) {
        const long qq = g.ix(m0 + (e >> 1), j + (e & 1));
        c.wk[o][e] = rows ? w[qq] : (R)0;
        if (o == 0) c.ch[e] = rows ? (R)chc[qq] : (R)1;
      }
    }
    const long q = g.ix(i, j);
    for (int o = 0; o < 3; ++o) {
      const int kk = k - 1 + o;
      const bool ok = kk >= 1 && kk <= g.nk;
      c.var[0][o] = ok ? uc[q + g.n2 * (kk - 1)] : 0.0;
      c.var[1][o] = ok ? vc[q + g.n2 * (kk - 1)] : 0.0;
    }
    c.lev = g.lev_u[q];
    c.rh3 = g.RH(3, i, j);
    c.dx13 = g.DX1(3, ic, j);
    c.dx23 = g.DX2(3, ic, j);
    return c;
  };
  const bool in = ic <= g.nim && j <= g.nj;
  AdvCellIn<R> own{};
  if (in) {
    own = cell_in(ic);
    face_ew(ic, j, ty, tx);
    face_ns(ic, j, ty, tx);
    if (tx == kAdvTX - 1 || ic == g.nim) face_ew(g.canon(ic + 1), j, ty, tx + 1);
    if ((ty == kAdvTY - 1 || j == g.nj) && j < g.nj) face_ns(ic, j + 1, ty + 1, tx);
  }
  __syncthreads();
  if (!in) return;

  auto cell = [&](int i, const AdvCellIn<R>& in) {
    const long p = ko + g.ix(i, j);
    // w = ter(wm/chc, g1 -> g3) on level kk of nkp (ter_di: rows 1..njm)
    auto W = [&](int o) -> R {
      if (!rows) return (R)0;
      return (dv(in.wk[o][0], in.ch[0]) + dv(in.wk[o][1], in.ch[1]) + dv(in.wk[o][2], in.ch[2]) +
              dv(in.wk[o][3], in.ch[3])) *
             (R)0.25;
    };
    const R w_k = W(0), w_k1 = W(1);
    const R rh3 = in.rh3;
    const R dpr = g.dpr2[k - 1];
    const R d1 = dv(fuu[ty][tx + 1] - fuu[ty][tx], in.dx13);
    const R d2 = dv(rows ? fvv[ty + 1][tx] - fvv[ty][tx] : (R)0, in.dx23);
    const R d3 = dv(w_k - w_k1, dpr);
    const int lev = in.lev;
    const R mu = (R)(lev >= k ? 1 : 0);
    R out[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const R p1m = dv(fm1[c][ty][tx + 1] - fm1[c][ty][tx], in.dx13);
      const R p2m = dv(rows ? fm2[c][ty + 1][tx] - fm2[c][ty][tx] : (R)0, in.dx23);
      R ans = dv(p1m + p2m, (R)ARAD * rh3);
      // vertical flux through the U-cell faces, zero unless both layers wet
      auto WB = [&](int kk) -> R {
        if (kk < 2 || kk > g.nk) return (R)0;
        const int o = kk - k;
        R a = ((R)in.var[c][o] + (R)in.var[c][o + 1]) * (R)0.5;
        a = a * (kk == k ? w_k : w_k1);
        return a * (R)(lev >= kk - 1 ? 1 : 0) * (R)(lev >= kk ? 1 : 0);
      };
      ans = ans + dv(WB(k) - WB(k + 1), dpr);
      ans = ans - (R)0.5 * (R)in.var[c][1] * (dv(d1 + d2, (R)ARAD * rh3) + d3);
      out[c] = ans * mu;
    }
    auc[p] = out[0];
    avc[p] = out[1];
  };
  cell(ic, own);
  if (ic == g.nim) cell(1, cell_in(1));
  if (ic == 2) cell(g.ni, cell_in(g.ni));
}

// op_fri for both components (mod_op.f90:103-184) with the horizontal
// viscosity's face fluxes staged in LDS.  fri_r3d's first term is
// div(ch*gra(v)) on U points: VA = ter(ch)*grax(v) on the east-west faces
// (g4 points) and VB*rh = ter(ch)*gray(v)*rh on the north-south faces (g2
// points), each shared by the two U cells beside it.  A workgroup owns a
// kFriTX x kFriTY tile of canonical columns at one level, evaluates every
// face once for u and v, then finishes each cell (metric terms, vertical
// viscosity, wind stress, bottom drag) as fri_r3d orders it, with the
// same operations in the same order.  Halo columns take their faces from their
// canonical columns and everything else from themselves.
constexpr int kFriTX = 64, kFriTY = 4;

template <typename R>
struct FriCellIn {
  R sch, a;          // schr and am at the point
  int lev;           // lev_u
  double tau[2];     // taux, tauy
  R cv1, cv2;
  double vp[2][3];   // uc, vc at k-1, k, k+1 (0 outside 1..nk)
  R km0, km1;        // km at k, k+1
  double ue, uw, ve, vw;  // uc, vc at the east and west neighbours of level k
  R sche, schw;
  R dx13, dx23, rh3, rh4, d14, d14e;
  double ub, vb;     // uc, vc at the bottom level (when the drag enters)
};

template <typename R>
__global__ void __launch_bounds__(kFriTX * kFriTY)
    k_fri_tiles(Grid<R> g, const double* __restrict__ uc, const double* __restrict__ vc, const R* __restrict__ am,
                const double* __restrict__ chc, const double* __restrict__ schr, const double* __restrict__ taux,
                const double* __restrict__ tauy, R* fx, R* fy) {
  const int G = g.ni - 2;
  const int ntx = (G + kFriTX - 1) / kFriTX;
  const int b = (int)xcd_block();
  // levels fastest: a tile's levels run together on one XCD, whose L2 then
  // serves the vertical neighbours each level reads
  const int k = b % g.nk + 1;
  const int r = b / g.nk;
  const int c0 = 2 + (r % ntx) * kFriTX, j0 = 1 + (r / ntx) * kFriTY;
  const int tx = threadIdx.x % kFriTX, ty = threadIdx.x / kFriTX;
  const int ic = c0 + tx, j = j0 + ty;
  const long ko = g.n2 * (k - 1);
  // [component][row][face]: fa east-west faces at g4 points c0 + f (face f
  // is cell f's west face, f + 1 its east face), fb north-south faces at
  // g2 points of rows j0 + row (row ty is cell ty's south face)
  __shared__ R fa[2][kFriTY][kFriTX + 1];
  __shared__ R fb[2][kFriTY + 1][kFriTX];
  // the horizontal views at level k: v = uc/sch (u) or vc/sch (v)
  auto VV = [&](int c, int m, int jj) -> R { return dv((R)(c ? vc : uc)[ko + g.ix(m, jj)], (R)schr[g.ix(m, jj)]); };
  auto CH = [&](int m, int jj) -> R { return (R)chc[g.ix(m, jj)]; };
  // VA(m, jj) = ter_ns<1>(ch) * grax<3,4>(v), m canonical
  auto face_ew = [&](int m, int jj, int row, int f) {
    const R chy = (jj < 1 || jj > g.njm) ? (R)0 : (R)0.5 * (CH(m, jj) + CH(m, jj + 1));
    for (int c = 0; c < 2; ++c) {
      const R gx = dv(dv(VV(c, m, jj) - VV(c, m - 1, jj), g.DX1(4, m, jj)), (R)ARAD * g.RH(4, m, jj));
      fa[c][row][f] = chy * gx;
    }
  };
  // VB(m, jj) * rh_g2 = ter_ew<1>(ch) * gray<3,2>(v) * rh_g2
  auto face_ns = [&](int m, int jj, int row, int f) {
    const R chx = (R)0.5 * (CH(m, jj) + CH(m + 1, jj));
    for (int c = 0; c < 2; ++c) {
      const R d = (jj < 2 || jj > g.nj) ? (R)0 : (VV(c, m, jj) - VV(c, m, jj - 1));
      const R gy = dv(dv(d, g.DX2(2, m, jj)), (R)ARAD);
      fb[c][row][f] = chx * gy * g.RH(2, m, jj);
    }
  };
  // the inputs of cell (i, j) besides the LDS faces (the thread's own cell
  // loads them before the face phase, so both rounds of loads overlap)
  auto cell_in = [&](int i) -> FriCellIn<R> {
    FriCellIn<R> c;
    const long q = g.ix(i, j);
    c.sch = (R)schr[q];
    c.a = am[ko + q];
    c.lev = g.lev_u[q];
    c.tau[0] = taux[q];
    c.tau[1] = tauy[q];
    c.cv1 = g.cv1[q];
    c.cv2 = g.cv2[q];
    for (int o = 0; o < 3; ++o) {
      const int kk = k - 1 + o;
      const bool ok = kk >= 1 && kk <= g.nk;
      c.vp[0][o] = ok ? uc[q + g.n2 * (kk - 1)] : 0.0;
      c.vp[1][o] = ok ? vc[q + g.n2 * (kk - 1)] : 0.0;
    }
    c.km0 = g.km[g.ix(i, j, k)];
    c.km1 = g.km[g.ix(i, j, k + 1)];
    const long qe = g.ix(ic + 1, j), qw = g.ix(ic - 1, j);
    c.ue = uc[ko + qe];
    c.uw = uc[ko + qw];
    c.ve = vc[ko + qe];
    c.vw = vc[ko + qw];
    c.sche = (R)schr[qe];
    c.schw = (R)schr[qw];
    c.dx13 = g.DX1(3, ic, j);
    c.dx23 = g.DX2(3, ic, j);
    c.rh3 = g.RH(3, ic, j);
    c.rh4 = g.RH(4, ic, j);
    c.d14 = g.DX1(4, ic, j);
    c.d14e = g.DX1(4, ic + 1, j);
    // the bottom velocity, for the drag in the layer above the bottom face
    const bool drag = c.lev > 0 && (c.lev == k - 1 || c.lev == k);
    c.ub = drag ? uc[g.ix(i, j, c.lev)] : 0.0;
    c.vb = drag ? vc[g.ix(i, j, c.lev)] : 0.0;
    return c;
  };
  const bool in = ic <= g.nim && j <= g.nj;
  FriCellIn<R> own{};
  if (in) {
    own = cell_in(ic);
    face_ew(ic, j, ty, tx);
    face_ns(ic, j, ty, tx);
    if (tx == kFriTX - 1 || ic == g.nim) face_ew(g.canon(ic + 1), j, ty, tx + 1);
    if ((ty == kFriTY - 1 || j == g.nj) && j < g.nj) face_ns(ic, j + 1, ty + 1, tx);
  }
  __syncthreads();
  if (!in) return;

  auto cell = [&](int i, const FriCellIn<R>& in) {
    const long p = ko + g.ix(i, j);
    const R sch = in.sch;
    const R a = in.a;
    const int lev = in.lev;
    R out[2];
    for (int c = 0; c < 2; ++c) {
      // fx from (u, uc, -v, taux), fy from (v, vc, u, tauy) (mod_op.f90:169-184)
      const R tau = (R)in.tau[c];
      // u or v / sch at the bottom level of this column
      auto v3b = [&]() -> R { return dv((R)(c ? in.vb : in.ub), sch); };
      // -v for fx, u for fy, at the bottom level and at the east and west
      // neighbours of level k
      auto vccb = [&]() -> R { return c ? dv((R)in.ub, sch) : -dv((R)in.vb, sch); };
      auto vcce = [&]() -> R { return c ? dv((R)in.ue, in.sche) : -dv((R)in.ve, in.sche); };
      auto vccw = [&]() -> R { return c ? dv((R)in.uw, in.schw) : -dv((R)in.vw, in.schw); };
      // divg<4,2,3>(VA, VB) (mod_op.f90:813-848)
      const R a1 = dv(fa[c][ty][tx + 1] - fa[c][ty][tx], in.dx13);
      const R d = (j < 1 || j > g.njm) ? (R)0 : (fb[c][ty + 1][tx] - fb[c][ty][tx]);
      const R b1 = dv(d, in.dx23);
      const R wka = dv(a1 + b1, (R)ARAD * in.rh3);
      R ans = dv(a, sch) * wka;
      ans = ans + a * in.cv1 * (R)in.vp[c][1];
      const R dx = (R)ARAD * in.rh4 * (in.d14 + in.d14e);
      const R pxc = dv(vcce() - vccw(), dx);
      ans = ans + a * in.cv2 * sch * pxc;
      auto WKP = [&](int kk) -> R {
        if (kk == 1) return dv((R)GRAV * tau, sch);
        if (lev > 0 && kk == lev + 1) {
          R vv = v3b(), cc = vccb();
          R mag = sqrt(vv * vv + cc * cc);
          return dv(vv * (R)GRAV * (R)CDBOT * mag, sch);
        }
        if (kk > g.nk) return (R)0;
        const int o = kk - k;
        R pz = dv(((R)in.vp[c][o] - (R)in.vp[c][o + 1]) * (R)(lev >= kk - 1 ? 1 : 0) * (R)(lev >= kk ? 1 : 0),
                  g.dz1[kk - 1]);
        return dv((R)GRAV * (R)RHO0 * (kk == k ? in.km0 : in.km1) * pz, sch * sch);
      };
      const R vert = dv(WKP(k) - WKP(k + 1), g.dpr2[k - 1]);
      out[c] = ans * (R)(lev >= k ? 1 : 0) + vert;
    }
    fx[p] = out[0];
    fy[p] = out[1];
  };
  cell(ic, own);
  if (ic == g.nim) cell(1, cell_in(1));
  if (ic == 2) cell(g.ni, cell_in(g.ni));
}

// column integral of the baroclinic tendency (mod_int.f90:136-149)
template <typename R>
__global__ void __launch_bounds__(kColCW * kColKL)
    k_ut(Grid<R> g, const R* __restrict__ auc, const R* __restrict__ avc, const R* __restrict__ fx,
         const R* __restrict__ fy, const double* __restrict__ schr, const double* __restrict__ pax,
         const double* __restrict__ pay, double* ut, double* vt) {
  extern __shared__ unsigned char ut_lds[];
  const int tx = threadIdx.x, kl = threadIdx.y, nk = g.nk;
  R* tx1 = reinterpret_cast<R*>(ut_lds);  // [nk][kColCW]
  R* ty1 = tx1 + nk * kColCW;
  const long p = (long)xcd_block() * kColCW + tx;
  const bool ok = p < g.n2;
  const int i = ok ? (int)(p % g.ni) + 1 : 1, j = ok ? (int)(p / g.ni) + 1 : 1;
  if (ok) {
    const R sch = sch_ready<R>(g, schr, i, j);
    const R px = (R)pax[p], py = (R)pay[p];
    for (int k = 1 + kl; k <= nk; k += kColKL) {
      long q = p + g.n2 * (k - 1);
      R x = -auc[q] - sch * (R)A0 * px + fx[q];
      R y = -avc[q] - sch * (R)A0 * py + fy[q];
      R mu = (R)g.msku(i, j, k);
      const int o = (k - 1) * kColCW + tx;
      tx1[o] = x * g.dpr2[k - 1] * mu;
      ty1[o] = y * g.dpr2[k - 1] * mu;
    }
  }
  __syncthreads();
  if (!ok || kl != 0) return;
  R su = 0, sv = 0;
  for (int k = 1; k <= nk; ++k) {
    su = su + tx1[(k - 1) * kColCW + tx];
    sv = sv + ty1[(k - 1) * kColCW + tx];
  }
  ut[p] = (double)su / g.prh_u[p];
  vt[p] = (double)sv / g.prh_u[p];
}

// ---------------------------------------------------------------------------
// int_clin (mod_int.f90:314-368): baroclinic momentum at one point,
// updated in place (uc == up; every read of uc/vc here is pointwise).  The
// pressure gradient is evaluated in E on the E grid; the AB3 combination,
// friction, Coriolis and the update of the FP64 state in double.
template <typename E>
struct ClinArgs {
  Grid<E> g;
  double *uc, *vc;
  const void *auc, *aup, *aupp, *avc, *avp, *avpp, *fx, *fy;  // in R
  const E *adp, *alpha;
  const double *mch, *schm, *pax, *pay, *cor;
  double bc;
};

// Workgroups take kBlock points of one level, levels fastest: the nk
// workgroups of a stretch of columns run together on one XCD, whose L2 then
// serves the 2-D fields and the neighbouring rows every level reads.
template <typename R, typename E>
__global__ void k_clin(ClinArgs<E> a) {
  const Grid<E>& g = a.g;
  const unsigned b = xcd_block();
  const long p2 = (long)(b / g.nk) * blockDim.x + threadIdx.x;
  if (p2 >= g.n2) return;
  long p = p2 + g.n2 * (long)(b % g.nk);
  int i, j, k;
  ijk(g, p, i, j, k);
  const double c_tc = (double)(23.0f / 12.0f), c_tp = (double)(-(16.0f / 12.0f)), c_tpp = (double)(5.0f / 12.0f);
  const long q = g.ix(i, j);
  const long ko = g.n2 * (k - 1);
  auto MCH = [&](int m, int jj) -> E { return (E)a.mch[g.ix(m, jj)]; };
  const double sch = a.schm[q];
  const E* ad = a.adp + ko;
  const E* al = a.alpha + ko;
  const E pr = g.pr2[k - 1];
  auto WK = [&](int m, int jj) -> E {
    long r = g.ix(m, jj);
    return (E)g.phih_t[r] + (E)a.mch[r] * ad[r];
  };
  auto AL = [&](int m, int jj) -> E { return al[g.ix(m, jj)]; };
  auto PG1 = [&](int m, int jj) -> E {
    return grax<1, 2>(g, WK, m, jj) + ter_ew<1>(g, AL, m, jj) * pr * grax<1, 2>(g, MCH, m, jj);
  };
  auto PG2 = [&](int m, int jj) -> E {
    return gray<1, 4>(g, WK, m, jj) + ter_ns<1>(g, AL, m, jj, (E)0) * pr * gray<1, 4>(g, MCH, m, jj);
  };
  const R* auc = static_cast<const R*>(a.auc);
  const R* aup = static_cast<const R*>(a.aup);
  const R* aupp = static_cast<const R*>(a.aupp);
  const R* avc = static_cast<const R*>(a.avc);
  const R* avp = static_cast<const R*>(a.avp);
  const R* avpp = static_cast<const R*>(a.avpp);
  double wk = (double)ter_ns<2>(g, PG1, i, j, (E)0);
  double pu1 = -((wk + A0 * a.pax[q]) * sch);
  wk = (double)ter_ew<4>(g, PG2, i, j);
  double pu2 = -((wk + A0 * a.pay[q]) * sch);
  pu1 = pu1 - (c_tc * (double)auc[p] + c_tp * (double)aup[p] + c_tpp * (double)aupp[p]);
  pu2 = pu2 - (c_tc * (double)avc[p] + c_tp * (double)avp[p] + c_tpp * (double)avpp[p]);
  pu1 = pu1 + (double)static_cast<const R*>(a.fx)[p];
  pu2 = pu2 + (double)static_cast<const R*>(a.fy)[p];
  const double cor = a.cor[q];
  const double u0 = a.uc[p], v0 = a.vc[p];
  pu1 = pu1 + cor * v0;
  pu2 = pu2 - cor * u0;
  const double x1 = pu1, x2 = pu2;
  const double c = 0.5 * cor * a.bc;
  pu1 = x1 + c * x2;
  pu2 = x2 - c * x1;
  const double d = 1.0 + c * c;
  const double mu = (double)g.msku(i, j, k);
  pu1 = pu1 * mu / d;
  pu2 = pu2 * mu / d;
  a.uc[p] = u0 + pu1 * a.bc;
  a.vc[p] = v0 + pu2 * a.bc;
}

// barotropic / baroclinic mode coupling (mod_int.f90:372-380), on the
// FP64 state
__global__ void __launch_bounds__(kColCW * kColKL)
    k_clin_cpl(Grid<double> g, double* uc, double* vc, const double* __restrict__ buc,
               const double* __restrict__ bvc) {
  extern __shared__ double cpl_lds[];  // [2][nk][kColCW]
  __shared__ double m1s[kColCW], m2s[kColCW];
  const int tx = threadIdx.x, kl = threadIdx.y, nk = g.nk;
  double* t1 = cpl_lds;
  double* t2 = t1 + nk * kColCW;
  const long p = (long)xcd_block() * kColCW + tx;
  const bool ok = p < g.n2;
  const int i = ok ? (int)(p % g.ni) + 1 : 1, j = ok ? (int)(p / g.ni) + 1 : 1;
  if (ok) {
    for (int k = 1 + kl; k <= nk; k += kColKL) {
      const long q = p + g.n2 * (k - 1);
      const double mu = (double)g.msku(i, j, k);
      const int o = (k - 1) * kColCW + tx;
      t1[o] = uc[q] * g.dpr2[k - 1] * mu;
      t2[o] = vc[q] * g.dpr2[k - 1] * mu;
    }
  }
  __syncthreads();
  if (ok && kl == 0) {
    double s1 = 0, s2 = 0;
    for (int k = 1; k <= nk; ++k) {
      s1 += t1[(k - 1) * kColCW + tx];
      s2 += t2[(k - 1) * kColCW + tx];
    }
    m1s[tx] = s1 / g.prh_u_r[p];
    m2s[tx] = s2 / g.prh_u_r[p];
  }
  __syncthreads();
  if (!ok) return;
  const double m1 = m1s[tx], m2 = m2s[tx];
  const double bu = buc[p], bv = bvc[p];
  for (int k = 1 + kl; k <= nk; k += kColKL) {
    if (g.msku(i, j, k) > 0) {
      const long q = p + g.n2 * (k - 1);
      uc[q] = uc[q] - m1 + bu;
      vc[q] = vc[q] - m2 + bv;
    }
  }
}

// unweighted velocity, its accumulator and smag (mod_int.f90:384-440)
// The unweighted velocity u*sqrt(mch2)/mch2 is staged in LDS for a
// kSmagTX x kSmagTY tile of canonical columns at one level plus a one-point
// rim, so each point's five-point Smagorinsky stencil reads it instead of
// recomputing it (one division per value instead of five).  Halo columns
// take their accumulator term from their own values and am from their
// canonical column, as the per-point composition does.
constexpr int kSmagTX = 64, kSmagTY = 4;

template <typename R>
__global__ void __launch_bounds__(kSmagTX * kSmagTY)
    k_smag(Grid<R> g, const double* __restrict__ uc, const double* __restrict__ vc,
           const double* __restrict__ mch2, const double* __restrict__ schs, R* acu, R* acv, R* am, int dtbc) {
  const int G = g.ni - 2;
  const int ntx = (G + kSmagTX - 1) / kSmagTX;
  const int b = (int)xcd_block();
  // levels fastest: a tile's levels run together on one XCD, whose L2 then
  // serves the vertical neighbours each level reads
  const int k = b % g.nk + 1;
  const int r = b / g.nk;
  const int c0 = 2 + (r % ntx) * kSmagTX, j0 = 1 + (r / ntx) * kSmagTY;
  const int tx = threadIdx.x % kSmagTX, ty = threadIdx.x / kSmagTX;
  const int ic = c0 + tx, j = j0 + ty;
  const long ko = g.n2 * (k - 1);
  auto S2 = [&](int m, int jj) -> R { return (R)schs[g.ix(m, jj)]; };
  auto U1 = [&](int m, int jj) -> R { return (R)uc[ko + g.ix(m, jj)] * S2(m, jj) / (R)mch2[g.ix(m, jj)]; };
  auto U2 = [&](int m, int jj) -> R { return (R)vc[ko + g.ix(m, jj)] * S2(m, jj) / (R)mch2[g.ix(m, jj)]; };
  // [row][col]: raw columns c0-1 .. c0+kSmagTX, rows j0-1 .. j0+kSmagTY
  __shared__ R su1[kSmagTY + 2][kSmagTX + 2], su2[kSmagTY + 2][kSmagTX + 2];
  for (int t = threadIdx.x; t < (kSmagTY + 2) * (kSmagTX + 2); t += blockDim.x) {
    const int rr = t / (kSmagTX + 2), cc = t % (kSmagTX + 2);
    const int m = c0 - 1 + cc, jj = j0 - 1 + rr;
    if (m <= g.ni && jj >= 1 && jj <= g.nj) {
      su1[rr][cc] = U1(m, jj);
      su2[rr][cc] = U2(m, jj);
    }
  }
  __syncthreads();
  if (ic > g.nim || j > g.nj) return;
  const int rr = ty + 1, cc = tx + 1;
  R amv = 0;
  const bool rows = j >= 2 && j <= g.nj - 1;
  if (rows) {
    R dx = (R)ARAD * (g.RH(4, ic, j) * g.DX1(4, ic, j) + g.RH(4, ic + 1, j) * g.DX1(4, ic + 1, j));
    R dy = (R)ARAD * (g.DX2(2, ic, j) + g.DX2(2, ic, j + 1));
    R tn = g.tn[3][g.ix(ic, j)];
    R dt = dv((R)4 * (su1[rr][cc + 1] - su1[rr][cc - 1]), dx) - dv((R)4 * (su2[rr + 1][cc] - su2[rr - 1][cc]), dy) +
           dv(su2[rr][cc] * tn, (R)ARAD);
    R ds = dv((R)4 * (su2[rr][cc + 1] - su2[rr][cc - 1]), dx) + dv((R)4 * (su1[rr + 1][cc] - su1[rr - 1][cc]), dy) -
           dv(su1[rr][cc] * tn, (R)ARAD);
    R cs = dv((R)0.5 * (R)SMAG_C, (R)PI_F) * ((R)0.5 * dx + (R)0.5 * dy);
    R c = cs * cs;
    R wk = c * sqrt(dt * dt + ds * ds);
    R amax = dv(dx * dx + dy * dy, (R)(4 * 8 * dtbc));
    amv = (amax < wk) ? amax : wk;
  }
  auto cell = [&](int i, bool own) {
    const long p = ko + g.ix(i, j);
    acu[p] = acu[p] + (own ? su1[rr][cc] : U1(i, j));
    acv[p] = acv[p] + (own ? su2[rr][cc] : U2(i, j));
    if (rows) am[p] = amv;
  };
  cell(ic, true);
  if (ic == g.nim) cell(1, false);
  if (ic == 2) cell(g.ni, false);
}

// int_ssh (mod_int.f90:602-624): alpha from (tp, sp, mch), column SSH
template <typename R>
__global__ void __launch_bounds__(kColCW * kColKL) k_ssh(Grid<R> g, const double* __restrict__ tp,
                                                         const double* __restrict__ sp,
                                                         const double* __restrict__ mch, R* alpha, double* ssh,
                                                         double* acssh) {
  extern __shared__ unsigned char ssh_lds[];
  R(*term)[kColCW] = reinterpret_cast<R(*)[kColCW]>(ssh_lds);  // [nk][kColCW]
  const int tx = threadIdx.x, kl = threadIdx.y, nk = g.nk;
  const long p = (long)xcd_block() * kColCW + tx;
  if (p < g.n2) {
    const R m = (R)mch[p];
    const int lev = g.lev_t[p];
    for (int k = 1 + kl; k <= nk; k += kColKL) {
      long q = p + g.n2 * (k - 1);
      R al = alpha[q];
      if (lev >= k) {
        al = (R)1.0 / den_rho<R>((R)tp[q], (R)sp[q], m * g.pr2[k - 1]);
        alpha[q] = al;
      }
      term[k - 1][tx] = al * g.dpr2[k - 1] * m * (R)(lev >= k ? 1 : 0);
    }
  }
  __syncthreads();
  if (p >= g.n2 || kl != 0) return;
  R s = 0;
  for (int k = 1; k <= nk; ++k) s += term[k - 1][tx];
  double h;
  if constexpr (sizeof(R) == sizeof(double)) {
    h = s / GRAV;
  } else {
    h = (double)s / GRAV;
  }
  h = g.phih_t[p] / GRAV + h;
  ssh[p] = h;
  acssh[p] = acssh[p] + h;
}

// op_adv_ts (mod_op.f90:280-325) + op_dif (186-222) + restoring + forward
// step + freezing clamp (mod_int.f90:464-491).  Reads the old T/S (tc ==
// tp), writes the new T/S into the other buffer of the pair.
template <typename R>
struct TsArgs {
  Grid<R> g;
  const double *t0, *s0, *uc, *vc;
  const R* wm;
  const double *schs, *mch, *bndt, *bnds;
  double *t1, *s1;
  double bc;
};

// int_ts's tendency and step (mod_int.f90:463-490), with the face fluxes
// staged in LDS.  op_adv_ts and op_dif are
// divergences of fluxes on the faces of a T cell (mod_op.f90:186-222,
// 280-325): the east-west faces (g2 points) carry X = ter(um)*p1(var) and
// VA = ter(ch)*masked p1(var)/(a rh), the north-south faces (g4 points) Y
// and VB*rh.  Computed per cell, every face is evaluated twice (once by
// each cell beside it); here a workgroup owns a kTsTX x kTsTY tile of
// canonical columns 2..nim at one level, evaluates each face of the tile
// once for T and S, and combines them per cell with the same operations
// in the same order.  The halo columns 1 and ni take their faces from
// their canonical columns nim and 2 and everything else from themselves,
// as the per-cell composition does.
constexpr int kTsTX = 64, kTsTY = 4;

template <typename R>
struct TsCellIn {
  double var[2][3];  // T, S at k-1, k, k+1
  R wm0, wm1, kh0, kh1;
  int lev;
  R rh1, dx1c, dx2c, rh1c;
  double mch, bt, bs;
};

// five waves per SIMD (92 VGPRs in FP64, no spills) hide more of the
// kernel's memory latency than four with the compiler's default 124
template <typename R>
__global__ void __launch_bounds__(kTsTX * kTsTY) __attribute__((amdgpu_waves_per_eu(5))) k_ts_tiles(TsArgs<R> a) {
  const Grid<R>& g = a.g;
  const int G = g.ni - 2;
  const int ntx = (G + kTsTX - 1) / kTsTX;
  const int b = (int)xcd_block();
  // levels fastest: a tile's levels run together on one XCD, whose L2 then
  // serves the vertical neighbours each level reads
  const int k = b % g.nk + 1;
  const int r = b / g.nk;
  const int c0 = 2 + (r % ntx) * kTsTX, j0 = 1 + (r / ntx) * kTsTY;
  const int tx = threadIdx.x % kTsTX, ty = threadIdx.x / kTsTX;
  const int ic = c0 + tx, j = j0 + ty;
  const long ko = g.n2 * (k - 1);
  // faces [tracer][row][column]: fx/fa east faces, column 0 the tile's
  // west face; fy/fb north faces, row 0 the tile's south face
  __shared__ R fx[2][kTsTY][kTsTX + 1], fa[2][kTsTY][kTsTX + 1];
  __shared__ R fy[2][kTsTY + 1][kTsTX], fb[2][kTsTY + 1][kTsTX];

  auto SCH = [&](int m, int jj) -> R { return (R)a.schs[g.ix(m, jj)]; };
  auto CH = [&](int m, int jj) -> R { return (R)a.mch[g.ix(m, jj)]; };
  auto VAR = [&](int v, int m, int jj) -> R { return (R)(v ? a.s0 : a.t0)[ko + g.ix(m, jj)]; };
  auto UM1 = [&](int m, int jj) -> R { return (R)a.uc[ko + g.ix(m, jj)] * SCH(m, jj); };
  auto UM2 = [&](int m, int jj) -> R { return (R)a.vc[ko + g.ix(m, jj)] * SCH(m, jj) * g.RH(3, m, jj); };
  // east-west face at g2 point (m, jj), m canonical: X (op_adv_ts) and
  // VA (op_dif) of T and S
  auto face_ew = [&](int m, int jj, int row, int col) {
    const R um = (jj < 2 || jj > g.nj) ? (R)0 : (R)0.5 * (UM1(m, jj - 1) + UM1(m, jj));  // ter_ns<3>
    const R chx = (R)0.5 * (CH(m, jj) + CH(m + 1, jj));                                 // ter_ew<1>
    for (int v = 0; v < 2; ++v) {
      const R dvar = VAR(v, m + 1, jj) - VAR(v, m, jj);
      fx[v][row][col] = um * dv(dvar, g.DX1(2, m, jj));  // * p1<1,2>
      R d = dvar * (R)g.mskt(m + 1, jj, k) * (R)g.mskt(m, jj, k);
      d = dv(d, g.DX1(2, m, jj));
      fa[v][row][col] = chx * dv(d, (R)ARAD * g.RH(2, m, jj));
    }
  };
  // north-south face at g4 point (m, jj): Y and VB * rh_g4
  auto face_ns = [&](int m, int jj, int row, int col) {
    const R vm = (R)0.5 * (UM2(m - 1, jj) + UM2(m, jj));                                     // ter_ew<3>
    const R chy = (jj < 1 || jj > g.njm) ? (R)0 : (R)0.5 * (CH(m, jj) + CH(m, jj + 1));      // ter_ns<1>
    for (int v = 0; v < 2; ++v) {
      const R dvar = (jj < 1 || jj > g.njm) ? (R)0 : VAR(v, m, jj + 1) - VAR(v, m, jj);
      fy[v][row][col] = vm * dv(dvar, g.DX2(4, m, jj));  // * p2<1,4>
      R d = (jj > g.njm) ? (R)0 : (VAR(v, m, jj + 1) - VAR(v, m, jj)) * (R)g.mskt(m, jj + 1, k) * (R)g.mskt(m, jj, k);
      d = dv(d, g.DX2(4, m, jj));
      fb[v][row][col] = chy * dv(d, (R)ARAD) * g.RH(4, m, jj);
    }
  };
  // the column inputs of cell (i, j): T/S at k-1, k, k+1 (0 outside 1..nk),
  // wm and kh at k and k+1, the mask depth, metrics and forcing
  auto cell_in = [&](int i) -> TsCellIn<R> {
    TsCellIn<R> c;
    const long q = g.ix(i, j);
    for (int v = 0; v < 2; ++v) {
      const double* varp = v ? a.s0 : a.t0;
      for (int o = 0; o < 3; ++o) {
        const int kk = k - 1 + o;
        c.var[v][o] = (kk >= 1 && kk <= g.nk) ? varp[q + g.n2 * (kk - 1)] : 0.0;
      }
    }
    c.wm0 = a.wm[g.ix(i, j, k)];
    c.wm1 = a.wm[g.ix(i, j, k + 1)];
    c.kh0 = g.kh[g.ix(i, j, k)];
    c.kh1 = g.kh[g.ix(i, j, k + 1)];
    c.lev = g.lev_t[q];
    c.rh1 = g.RH(1, i, j);
    c.dx1c = g.DX1(1, ic, j);
    c.dx2c = g.DX2(1, ic, j);
    c.rh1c = g.RH(1, ic, j);
    c.mch = a.mch[q];
    c.bt = k == 1 ? a.bndt[q] : 0.0;
    c.bs = k == 1 ? a.bnds[q] : 0.0;
    return c;
  };
  const bool in = ic <= g.nim && j <= g.nj;
  TsCellIn<R> own{};
  if (in) {
    own = cell_in(ic);
    face_ew(ic, j, ty, tx + 1);
    face_ns(ic, j, ty + 1, tx);
    if (tx == 0) face_ew(c0 - 1 == 1 ? g.nim : c0 - 1, j, ty, 0);
    if (ty == 0 && j0 > 1) face_ns(ic, j0 - 1, 0, tx);
  }
  __syncthreads();
  if (!in) return;

  // one cell (i, j) from the faces of its canonical column ic and the
  // column inputs `in` (loaded by cell_in, for the thread's own cell before
  // the face phase so that both rounds of loads are in flight together)
  auto cell = [&](int i, const TsCellIn<R>& in) {
    const long p = ko + g.ix(i, j);
    R pt[2];
    for (int v = 0; v < 2; ++v) {
      // op_adv_ts (mod_op.f90:280-325)
      const R rh1 = in.rh1;
      R ans = dv((R)0.5 * (fx[v][ty][tx] + fx[v][ty][tx + 1]), (R)ARAD * rh1);                      // ter_ew<2>
      ans = ans + dv((j < 2 || j > g.nj) ? (R)0 : (R)0.5 * (fy[v][ty][tx] + fy[v][ty + 1][tx]),    // ter_ns<4>
                     (R)ARAD * rh1);
      // var at kk - 1 minus var at kk (dx3_r3d_b), kk = k or k + 1
      auto DX3 = [&](int kk) -> R {
        if (kk < 2 || kk > g.nk) return (R)0;
        const int o = kk - k;
        return (R)in.var[v][o] - (R)in.var[v][o + 1];
      };
      R wka = (in.wm0 * DX3(k) + in.wm1 * DX3(k + 1)) * (R)0.5;
      wka = dv(wka, g.dpr2[k - 1]);
      pt[v] = -(ans + wka);
    }
    for (int v = 0; v < 2; ++v) {
      // op_dif (mod_op.f90:186-222): divg<2,4,1> of (VA, VB*rh)
      const R a1 = dv(fa[v][ty][tx + 1] - fa[v][ty][tx], in.dx1c);
      const R d = (j < 2 || j > g.nj) ? (R)0 : (fb[v][ty + 1][tx] - fb[v][ty][tx]);
      const R b1 = dv(d, in.dx2c);
      R ans = (R)AH_C * dv(a1 + b1, (R)ARAD * in.rh1c);
      auto WKP = [&](int kk) -> R {
        R w = (R)0;
        if (kk >= 2 && kk <= g.nk) {
          const int o = kk - k;
          w = dv(((R)in.var[v][o] - (R)in.var[v][o + 1]) * (R)(in.lev >= kk - 1 ? 1 : 0) * (R)(in.lev >= kk ? 1 : 0),
                 g.dz1[kk - 1]);
        }
        return (R)GRAV * (R)RHO0 * (kk == k ? in.kh0 : in.kh1) * w;
      };
      pt[v] = pt[v] + (ans + dv(WKP(k) - WKP(k + 1), g.dpr2[k - 1]));
    }
    const R mk = (R)(in.lev >= k ? 1 : 0);
    const R m = (R)in.mch;
    const double t0 = in.var[0][1], s0 = in.var[1][1];
    if (k == 1) {
      pt[0] = pt[0] + dv((R)GAMMA_T * ((R)in.bt - (R)t0) * mk, g.dpr2[0]);
      pt[1] = pt[1] + (R)GAMMA_S * ((R)in.bs - (R)s0) * mk * m;
    }
    // the forward step of the FP64 state (mod_int.f90:484-487)
    double tn = t0 + (double)pt[0] * a.bc * (double)mk / (double)m;
    double sn = s0 + (double)pt[1] * a.bc * (double)mk / (double)m;
    if (in.lev >= k && tn < TICE) tn = TICE;
    a.t1[p] = tn;
    a.s1[p] = sn;
  };
  if (in) cell(ic, own);
  if (in && ic == g.nim) cell(1, cell_in(1));
  if (in && ic == 2) cell(g.ni, cell_in(g.ni));
}

// convect / adjust / mix (mod_int.f90:509-600) on one column, in place.
// rhoa(k) = prho(k).(t(k), s(k)) and rhob(k) = prho(k+1).(t(k), s(k)) are
// kept equal to those expressions of the current column by mix(), and
// every entry the search reads is one mix() has kept current, so they are
// evaluated on demand instead of stored.  The column lives in LDS (stride
// `st` between levels) while the data-dependent search runs.
template <typename E>
struct Column {
  double* t;
  double* s;
  const E* p1;
  const E* p2;
  const double* dpr;
  int st;
  __device__ double rhoa(int k) const {
    return (double)p1[st * (k - 1)] * t[st * (k - 1)] + (double)p2[st * (k - 1)] * s[st * (k - 1)];
  }
  __device__ double rhob(int k) const {
    return (double)p1[st * k] * t[st * (k - 1)] + (double)p2[st * k] * s[st * (k - 1)];
  }
  // the three sums of the reference's three loops in one pass: each is
  // still accumulated over k = na..nb in increasing k
  __device__ void mix(int na, int nb) {
    double sumt = 0, sumd = 0, sums = 0;
    for (int k = na; k <= nb; ++k) {
      const double d = dpr[k - 1];
      sumt += t[st * (k - 1)] * d;
      sumd += d;
      sums += s[st * (k - 1)] * d;
    }
    const double tv = sumt / sumd;
    const double sv = sums / sumd;
    for (int k = na; k <= nb; ++k) {
      t[st * (k - 1)] = tv;
      s[st * (k - 1)] = sv;
    }
  }
};

// The same search for the mixed forms, which need not reproduce the
// reference's summation order: a range mixed to (tv, sv) has sum t d =
// tv sum d, so extending it by a level costs one update instead of the
// reference's new sums over the whole range, and the mixed values are
// written once per range instead of once per extension.  Deep convection
// then costs O(levels) per column instead of O(levels^2).  The tests
// (comparisons and mixing order) are the reference's, on the values the
// range holds.
template <typename E>
__device__ void convect_linear(const Column<E>& c, int nlev) {
  int kcnt = 1;
  while (kcnt < nlev) {
    if (!(c.rhob(kcnt) > c.rhoa(kcnt + 1))) {
      kcnt = kcnt + 1;
      continue;
    }
    int na = kcnt, nb = kcnt + 1;
    const double da = c.dpr[na - 1], db = c.dpr[nb - 1];
    double sumd = da + db;
    double tv = (c.t[c.st * (na - 1)] * da + c.t[c.st * (nb - 1)] * db) / sumd;
    double sv = (c.s[c.st * (na - 1)] * da + c.s[c.st * (nb - 1)] * db) / sumd;
    // rho of the range's own levels, with its mixed values
    auto ra = [&](int k) { return (double)c.p1[c.st * (k - 1)] * tv + (double)c.p2[c.st * (k - 1)] * sv; };
    auto rb = [&](int k) { return (double)c.p1[c.st * k] * tv + (double)c.p2[c.st * k] * sv; };
    auto take = [&](int k) {
      const double d = c.dpr[k - 1];
      const double nd = sumd + d;
      tv = (tv * sumd + c.t[c.st * (k - 1)] * d) / nd;
      sv = (sv * sumd + c.s[c.st * (k - 1)] * d) / nd;
      sumd = nd;
    };
    while (nb < nlev) {
      if (rb(nb) > c.rhoa(nb + 1)) {
        nb = nb + 1;
        take(nb);
      } else {
        if (na <= 1 || c.rhob(na - 1) <= ra(na)) break;
        na = na - 1;
        take(na);
      }
    }
    while (na > 1) {
      if (c.rhob(na - 1) <= ra(na)) break;
      na = na - 1;
      take(na);
    }
    for (int k = na; k <= nb; ++k) {
      c.t[c.st * (k - 1)] = tv;
      c.s[c.st * (k - 1)] = sv;
    }
    kcnt = nb;
  }
}

constexpr int kConvCW = 32, kConvKL = 8;

// kConvCW columns per workgroup (fewer than the other column kernels: the
// LDS per workgroup then allows 5 workgroups per CU, whose memory sweeps and
// searches overlap); the kConvKL k-lanes stage t, s (FP64 state) and
// prho (E) of each column into LDS (dynamic: 2*nk*kConvCW doubles + 2*nk*kConvCW
// E), lane 0 runs the convective adjustment, then all lanes write back and
// accumulate act, acs and acr (mod_int.f90:493-505) in R, and int_ts's w
// accumulator acw (mod_int.f90:458-461: wm * alpha / (g mch2) with
// int_ssh's alpha; every point once, as convect visits them).  The 1/rho that
// acr takes is the next step's den_alpha (mod_den.f90, main.f90:205: the
// same T, S and chc, which nothing changes in between); with `alphn` set
// the kernel stores it there (alpha's own value at levels below the
// bottom, which den_alpha leaves as they are) and int_tobot's adp of it in
// `adpn`, in the prho1 LDS array the search no longer needs.
template <typename R, typename E>
__global__ void __launch_bounds__(kConvCW * kConvKL) k_conv(Grid<E> g, const double* __restrict__ dpr, double* tc,
                                                    double* sc, const E* __restrict__ prho1,
                                                    const E* __restrict__ prho2, const double* __restrict__ chc,
                                                    R* act, R* acs, R* acr, R* acw, const R* __restrict__ wm,
                                                    const double* __restrict__ mch2, const E* __restrict__ alpha,
                                                    E* alphn, E* adpn) {
  extern __shared__ double conv_lds[];
  __shared__ double ldpr[64];  // dpr per level (nk <= 64), read by the searches
  const int tx = threadIdx.x, kl = threadIdx.y, nk = g.nk;
  if (kl == 0 && tx < nk) ldpr[tx] = dpr[tx];
  if (kl == 1 && tx + kConvCW < nk) ldpr[tx + kConvCW] = dpr[tx + kConvCW];
  double* lt = conv_lds;
  double* ls = lt + nk * kConvCW;
  E* l1 = reinterpret_cast<E*>(ls + nk * kConvCW);
  E* l2 = l1 + nk * kConvCW;
  const long p = (long)xcd_block() * kConvCW + tx;
  const bool ok = p < g.n2;
  const int nlev = ok ? g.lev_t[p] : 0;
  if (ok) {
    for (int k = 1 + kl; k <= nk; k += kConvKL) {
      const long q = p + g.n2 * (k - 1);
      const int o = (k - 1) * kConvCW + tx;
      lt[o] = tc[q];
      ls[o] = sc[q];
      l1[o] = prho1[q];
      l2[o] = prho2[q];
    }
  }
  __syncthreads();
  // the searches of the kConvCW columns are spread over the kConvKL k-lanes
  // (column tx + kPer * kl on lane tx < kPer of k-lane kl), so every wave of
  // the workgroup runs a few searches and waits only for the longest of them
  constexpr int kPer = kConvCW / kConvKL;
  const int col = tx + kPer * kl;
  const long pcol = (long)xcd_block() * kConvCW + col;
  const int ncol = tx < kPer && pcol < g.n2 ? g.lev_t[pcol] : 0;
  // the FP64 form (R = double) follows the reference's sums; the mixed
  // forms (FP32 tendencies) take the linear search
  if (ncol > 1 && !std::is_same<R, double>::value) {
    const Column<E> c{lt + col, ls + col, l1 + col, l2 + col, ldpr, kConvCW};
    convect_linear(c, ncol);
  } else if (ncol > 1) {
    Column<E> c{lt + col, ls + col, l1 + col, l2 + col, ldpr, kConvCW};
    const int nlev = ncol;
    int kcnt = 1;
    while (kcnt < nlev) {
      if (c.rhob(kcnt) > c.rhoa(kcnt + 1)) {
        int na = kcnt, nb = kcnt + 1;
        c.mix(na, nb);
        while (nb < nlev) {
          if (c.rhob(nb) > c.rhoa(nb + 1)) {
            nb = nb + 1;
            c.mix(na, nb);
          } else {
            if (na <= 1 || c.rhob(na - 1) <= c.rhoa(na)) break;
            na = na - 1;
            c.mix(na, nb);
          }
        }
        while (na > 1) {
          if (c.rhob(na - 1) <= c.rhoa(na)) break;
          na = na - 1;
          c.mix(na, nb);
        }
        kcnt = nb;
      } else {
        kcnt = kcnt + 1;
      }
    }
  }
  __syncthreads();  // the column's new T/S are visible to the other k-lanes
  const bool fuse = alphn != nullptr;
  if (ok) {
    const E ch = (E)chc[p];
    for (int k = 1 + kl; k <= nk; k += kConvKL) {
      const long q = p + g.n2 * (k - 1);
      const int o = (k - 1) * kConvCW + tx;
      const double t = lt[o], s = ls[o];
      tc[q] = t;
      sc[q] = s;
      act[q] = act[q] + (R)t;
      acs[q] = acs[q] + (R)s;
      acw[q] = acw[q] + dv(wm[q] * (R)alpha[q], (R)GRAV * (R)mch2[p]);
      if (nlev >= k) {
        E wk = dv((E)1.0, den_rho<E>((E)t, (E)s, ch * g.pr2[k - 1]));
        acr[q] = acr[q] + (R)dv((E)1.0, wk);
        if (fuse) {
          alphn[q] = wk;
          l1[o] = wk * g.dpr2[k - 1] * (E)1;
        }
      } else if (fuse) {
        const E a = alpha[q];
        alphn[q] = a;
        l1[o] = a * g.dpr2[k - 1] * (E)0;
      }
    }
  }
  if (!fuse) return;
  __syncthreads();  // the column's int_tobot terms are in l1
  if (!ok) return;
  for (int k = 1 + kl; k <= nk; k += kConvKL) adpn[p + g.n2 * (k - 1)] = tobot(l1 + tx, kConvCW, k, nk);
}

// ---------------------------------------------------------------------------
// diagnostics: FP64 global sums (deterministic two-level reduction) and a
// non-finite scan.
template <typename R>
__global__ void k_sums_partial(Grid<R> g, const R* __restrict__ t, const R* __restrict__ s,
                               const double* __restrict__ chc, const double* __restrict__ w2, double* part) {
  __shared__ double sh[3][kBlock];
  long p = (long)xcd_block() * blockDim.x + threadIdx.x;
  double a = 0, b = 0, c = 0;
  if (p < g.n2) {
    int i = (int)(p % g.ni) + 1, j = (int)(p / g.ni) + 1;
    if (i >= 2 && i <= g.nim && j >= 2 && j <= g.njm) {
      double area = w2[p];
      for (int k = 1; k <= g.nk; ++k) {
        if (g.mskt(i, j, k)) {
          long q = p + g.n2 * (k - 1);
          double v = area * (double)g.dpr2[k - 1];
          a += v * (double)t[q];
          b += v * (double)s[q];
        }
      }
      if (g.lev_t[p] > 0) c += area * chc[p];
    }
  }
  sh[0][threadIdx.x] = a;
  sh[1][threadIdx.x] = b;
  sh[2][threadIdx.x] = c;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int v = 0; v < 3; ++v) sh[v][threadIdx.x] += sh[v][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int v = 0; v < 3; ++v) part[3 * blockIdx.x + v] = sh[v][0];
}

__global__ void k_sums_final(const double* part, int nb, double* out) {
  if (threadIdx.x < 3) {
    double s = 0;
    for (int b = 0; b < nb; ++b) s += part[3 * b + threadIdx.x];
    out[threadIdx.x] = s;
  }
}

template <typename R>
__global__ void k_finite(const R* a, long n, int* flag) {
  long p = (long)xcd_block() * blockDim.x + threadIdx.x;
  if (p < n && !isfinite((double)a[p])) atomicOr(flag, 1);
}

__global__ void k_finite_d(const double* a, long n, int* flag) {
  long p = (long)xcd_block() * blockDim.x + threadIdx.x;
  if (p < n && !isfinite(a[p])) atomicOr(flag, 1);
}

__global__ void k_cell_area(Grid<double> g, double* w) {
  long p = (long)xcd_block() * blockDim.x + threadIdx.x;
  if (p >= g.n2) return;
  w[p] = g.rh[1][p] * g.dx1[1][p] * g.dx2[1][p] * ARAD * ARAD;
}

// ---------------------------------------------------------------------------
// time-mean output (main.f90:726-760, mod_mympi.f90:178-343): interior
// points of acc / n into a (glo_ni, glo_nj, nk) array, land masked
template <typename R>
__global__ void k_mean3(Grid<R> g, const R* __restrict__ acc, int n, int mask, double* out) {
  const long gni = g.ni - 2, gnj = g.nj - 2;
  long p = (long)xcd_block() * blockDim.x + threadIdx.x;
  if (p >= gni * gnj * g.nk) return;
  const int i = (int)(p % gni) + 2;
  const long r = p / gni;
  const int j = (int)(r % gnj) + 2, k = (int)(r / gnj) + 1;
  double v = (double)acc[g.ix(i, j, k)] / (double)n;
  if ((mask == 1 && g.mskt(i, j, k) == 0) || (mask == 2 && g.msku(i, j, k) == 0)) v = (double)PCOM_MISSING_FLOAT;
  out[p] = v;
}

__global__ void k_mean2(Grid<double> g, const double* __restrict__ acc, int n, double* out) {
  const long gni = g.ni - 2, gnj = g.nj - 2;
  long p = (long)xcd_block() * blockDim.x + threadIdx.x;
  if (p >= gni * gnj) return;
  const int i = (int)(p % gni) + 2, j = (int)(p / gni) + 2;
  double v = acc[g.ix(i, j)] / (double)n;
  if (g.lev_t[g.ix(i, j)] < 1) v = (double)PCOM_MISSING_FLOAT;
  out[p] = v;
}

// sum(glo_var*glo_wg) and sum(glo_wg), wg = rh_g1*msk, each accumulated in
// Fortran's column-major order by one lane so the FP64 result is the
// reference's; the two sums run on two lanes side by side
__global__ void k_ssh_mean(Grid<double> g, const double* __restrict__ var, double* sums) {
  const int t = threadIdx.x;
  if (t > 1) return;
  const long gni = g.ni - 2, gnj = g.nj - 2;
  double s = 0.0;
  for (long jj = 0; jj < gnj; ++jj)
    for (long ii = 0; ii < gni; ++ii) {
      const long q = g.ix((int)ii + 2, (int)jj + 2);
      const double w = g.rh[1][q] * (double)(g.lev_t[q] >= 1 ? 1 : 0);
      s += t == 0 ? var[ii + gni * jj] * w : w;
    }
  sums[t] = s;
}

__global__ void k_ssh_anom(long n, const double* __restrict__ sums, double* var) {
  long p = (long)xcd_block() * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double mean = sums[0] / sums[1];
  if (var[p] != (double)PCOM_MISSING_FLOAT) var[p] = var[p] - mean;
}

template <typename R>
__global__ void k_fill(R* a, long n) {
  long p = (long)xcd_block() * blockDim.x + threadIdx.x;
  if (p < n) a[p] = (R)0;
}

template <typename R>
inline R* P(void* v) {
  return static_cast<R*>(v);
}

template <typename R>
const Grid<R>& grid_of(const Ctx& c);
template <>
const Grid<double>& grid_of<double>(const Ctx& c) {
  return c.gd;
}
template <>
const Grid<float>& grid_of<float>(const Ctx& c) {
  return c.gf;
}

#define LAUNCH(kern, n, ...)                                                        \
  do {                                                                              \
    hipLaunchKernelGGL(kern, dim3(nblocks(n)), dim3(kBlock), 0, c.stream, __VA_ARGS__); \
    hipError_t e_ = hipGetLastError();                                              \
    if (e_ != hipSuccess) return hip_err(e_, #kern);                                \
  } while (0)

#define LAUNCH_ON(st, kern, n, ...)                                                 \
  do {                                                                              \
    hipLaunchKernelGGL(kern, dim3(nblocks(n)), dim3(kBlock), 0, st, __VA_ARGS__);   \
    hipError_t e_ = hipGetLastError();                                              \
    if (e_ != hipSuccess) return hip_err(e_, #kern);                                \
  } while (0)

// a column kernel of kColCW columns x kColKL k-lanes with one [nk][kColCW]
// LDS array of `elem`-byte entries
#define LAUNCH_COLS_ON(st, kern, elem, ...)                                         \
  do {                                                                              \
    hipLaunchKernelGGL(kern, dim3((unsigned)((c.n2 + kColCW - 1) / kColCW)), dim3(kColCW, kColKL), \
                       (size_t)c.nk * kColCW * (elem), st, __VA_ARGS__);            \
    hipError_t e_ = hipGetLastError();                                              \
    if (e_ != hipSuccess) return hip_err(e_, #kern);                                \
  } while (0)

#define LAUNCH_UPW() LAUNCH_COLS_ON(c.stream, k_upw<R>, sizeof(R), g, U, V, h.sqp, P<R>(f.wm))

// work on c.side after everything queued so far on c.stream (fork), and
// c.stream's later work after c.side's (join)
inline hipError_t fork_side(Ctx& c, int w) {
  hipError_t e = hipEventRecord(c.fork[w], c.stream);
  return e != hipSuccess ? e : hipStreamWaitEvent(c.side, c.fork[w], 0);
}
inline hipError_t join_side(Ctx& c, int w) {
  hipError_t e = hipEventRecord(c.join[w], c.side);
  return e != hipSuccess ? e : hipStreamWaitEvent(c.stream, c.join[w], 0);
}

// main.f90:164-211: daily forcing, den_alpha, int_pgra, int_readyc (the
// step up to the barotropic subcycle)
template <typename R, typename E>
std::string step_pre_t(Ctx& c) {
  const Grid<R>& g = grid_of<R>(c);
  const Grid<E>& ge = grid_of<E>(c);
  Fields3& f = c.f3;
  Fields2& h = c.f2;
  const size_t n2 = c.n2, n3 = c.n3;
  double* T = P<double>(f.T[c.cur]);
  double* S = P<double>(f.S[c.cur]);
  double* U = P<double>(f.U);
  double* V = P<double>(f.V);
  hipError_t e;

  c.step_i += 1;
  if (c.timing) (void)hipEventRecord(c.ev[0], c.stream);
  // daily forcing (main.f90:164-180): int_bnd writes the restoring fields
  // int_ts reads, and den_prho reads T and S and writes the coefficients
  // convect reads, so the previous step's int_ts and convect must be done
  // (skipped when the caller applied this step's refresh with its own
  // forcing records, pcom_set_forcing)
  const bool daily = (c.step_i == 1 || c.ct.d != c.pt.d) && c.forcing_taken != c.step_i;
  if (daily) {
    std::string es = settle_side(c);
    if (!es.empty()) return es;
    int m1, m2;
    double fw;
    bnd_weights(c.ct, &m1, &m2, &fw);
    Six s6;
    for (int v = 0; v < 6; ++v) {
      s6.a[v] = h.frc[v];
      s6.b[v] = h.bnd[v];
    }
    LAUNCH(k_bnd, n2, (long)n2, s6, m1, m2, fw);
    LAUNCH(k_prho<E>, n3, ge, T, S, h.chc, P<E>(f.prho1), P<E>(f.prho2));
  }
  if (daily || c.pa_stale) {
    LAUNCH(k_pa, n2, c.gd, h.bnd[4], h.pax, h.pay);
    c.pa_stale = false;
  }
  // den_alpha, int_pgra on the side stream: they read T, S, chc and write
  // alpha, adp, bphi, bgraphi, none of which int_readyc reads or writes.
  // While the previous step's int_ts is queued there (side_pending), the
  // side stream is already past everything of c.stream's that they read
  // (it forked after int_ts's upwelling); otherwise it forks here.
  if (!c.side_pending && (e = fork_side(c, 0)) != hipSuccess) return hip_err(e, "fork");
  const bool ready = c.next_ready;  // the previous step formed schr, wm, alpha, adp
  c.next_ready = false;
  if (ready) {
    std::swap(f.alpha, f.alphn);
    std::swap(f.adp, f.adpn);
  } else {
    LAUNCH_ON(c.side, k_alpha<E>, n3, ge, T, S, h.chc, P<E>(f.alpha));
    LAUNCH_COLS_ON(c.side, k_adp<E>, sizeof(E), ge, P<E>(f.alpha), P<E>(f.adp));
  }
  {
    const size_t lds = 4 * (size_t)c.nk * kColCW * sizeof(E);
    hipError_t eb = hipSuccess;
    if (lds > 65536)
      eb = hipFuncSetAttribute((const void*)k_bphi<E>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (eb != hipSuccess) return hip_err(eb, "k_bphi LDS size");
    hipLaunchKernelGGL(k_bphi<E>, dim3((unsigned)((c.n2 + kColCW - 1) / kColCW)), dim3(kColCW, kColKL), lds, c.side,
                       ge, P<E>(f.adp), P<E>(f.alpha), h.bphi[0], h.bphi[1], h.bphi[2], h.bphi[3], h.bgra[0],
                       h.bgra[1], h.bgra[2], h.bgra[3]);
    if ((eb = hipGetLastError()) != hipSuccess) return hip_err(eb, "k_bphi");
  }
  // int_readyc.  Its sqrt(ch) factors and upwelling come from the same
  // chc, chp, mch, uc and vc as the previous step's int_ts ones, which that
  // step formed (and its int_ts, on the side stream, may still read)
  if (!ready) {
    if (c.side_pending) {  // nothing may overwrite wm and sqp under int_ts
      std::string es = settle_side(c);
      if (!es.empty()) return es;
    }
    LAUNCH(k_sch, n2, c.gd, h.chc, h.chp, h.mch, h.schr, h.sqp, (double*)nullptr, (double*)nullptr);
    LAUNCH_UPW();
  }
  c.ab = (c.ab + 1) % 3;
  R* au = P<R>(f.AU[c.ab]);
  R* av = P<R>(f.AV[c.ab]);
  R* aup = P<R>(f.AU[(c.ab + 2) % 3]);
  R* avp = P<R>(f.AV[(c.ab + 2) % 3]);
  R* aupp = P<R>(f.AU[(c.ab + 1) % 3]);
  R* avpp = P<R>(f.AV[(c.ab + 1) % 3]);
  auto launch_adv = [&]() -> std::string {
    {
      const int G = c.ni - 2;
      const unsigned nb = (unsigned)(((G + kAdvTX - 1) / kAdvTX) * ((c.nj + kAdvTY - 1) / kAdvTY) * c.nk);
      hipLaunchKernelGGL(k_adv_tiles<R>, dim3(nb), dim3(kAdvTX * kAdvTY), 0, c.stream, g, U, V, P<R>(f.wm), h.chc,
                         h.schr, au, av);
      hipError_t ea = hipGetLastError();
      if (ea != hipSuccess) return hip_err(ea, "k_adv_tiles");
    }
    if (c.rst != 1) {  // AB3 bootstrap (mod_int.f90:122-132)
      size_t bytes = n3 * sizeof(R);
      if (c.step_i == 1) {
        if ((e = hipMemcpyAsync(aup, au, bytes, hipMemcpyDeviceToDevice, c.stream)) != hipSuccess) return hip_err(e, "copy");
        if ((e = hipMemcpyAsync(aupp, au, bytes, hipMemcpyDeviceToDevice, c.stream)) != hipSuccess) return hip_err(e, "copy");
        if ((e = hipMemcpyAsync(avp, av, bytes, hipMemcpyDeviceToDevice, c.stream)) != hipSuccess) return hip_err(e, "copy");
        if ((e = hipMemcpyAsync(avpp, av, bytes, hipMemcpyDeviceToDevice, c.stream)) != hipSuccess) return hip_err(e, "copy");
      } else if (c.step_i == 2) {
        if ((e = hipMemcpyAsync(aupp, aup, bytes, hipMemcpyDeviceToDevice, c.stream)) != hipSuccess) return hip_err(e, "copy");
        if ((e = hipMemcpyAsync(avpp, avp, bytes, hipMemcpyDeviceToDevice, c.stream)) != hipSuccess) return hip_err(e, "copy");
      }
    }
    return "";
  };
  auto launch_fri = [&]() -> std::string {
    {
      const int G = c.ni - 2;
      const unsigned nb = (unsigned)(((G + kFriTX - 1) / kFriTX) * ((c.nj + kFriTY - 1) / kFriTY) * c.nk);
      hipLaunchKernelGGL(k_fri_tiles<R>, dim3(nb), dim3(kFriTX * kFriTY), 0, c.stream, g, U, V, P<R>(f.am), h.chc,
                         h.schr, h.bnd[0], h.bnd[1], P<R>(f.fx), P<R>(f.fy));
      hipError_t ef = hipGetLastError();
      if (ef != hipSuccess) return hip_err(ef, "k_fri_tiles");
    }
    return "";
  };
  // op_adv and op_fri are independent; which goes first decides what each
  // shares the device with while the other stream runs int_ts and convect
  // (measured: FP64 adv first, FP32 stencils fri first, 1.2% per step)
  std::string el;
  if constexpr (sizeof(R) == sizeof(float)) {
    if (!(el = launch_fri()).empty() || !(el = launch_adv()).empty()) return el;
  } else {
    if (!(el = launch_adv()).empty() || !(el = launch_fri()).empty()) return el;
  }
  {
    const size_t lds = 2 * (size_t)c.nk * kColCW * sizeof(R);
    hipLaunchKernelGGL(k_ut<R>, dim3((unsigned)((c.n2 + kColCW - 1) / kColCW)), dim3(kColCW, kColKL), lds,
                       c.stream, g, au, av, P<R>(f.fx), P<R>(f.fy), h.schr, h.pax, h.pay, h.but, h.bvt);
    hipError_t eu = hipGetLastError();
    if (eu != hipSuccess) return hip_err(eu, "k_ut");
  }
  // the subcycle's set-up kernel needs nothing from the side stream: it
  // runs while the join waits
  std::string ev = launch_vz(c);
  if (!ev.empty()) return ev;
  if ((e = join_side(c, 0)) != hipSuccess) return hip_err(e, "join");
  c.side_pending = false;  // the join follows the previous step's int_ts too
  return "";
}

// main.f90:214-226 after int_trop: int_clin, int_ssh, int_ts, calendar
template <typename R, typename E>
std::string step_post_t(Ctx& c) {
  const Grid<R>& g = grid_of<R>(c);
  const Grid<E>& ge = grid_of<E>(c);
  Fields3& f = c.f3;
  Fields2& h = c.f2;
  const size_t n2 = c.n2, n3 = c.n3;
  double* T = P<double>(f.T[c.cur]);
  double* S = P<double>(f.S[c.cur]);
  double* U = P<double>(f.U);
  double* V = P<double>(f.V);
  R* au = P<R>(f.AU[c.ab]);
  R* av = P<R>(f.AV[c.ab]);
  R* aup = P<R>(f.AU[(c.ab + 2) % 3]);
  R* avp = P<R>(f.AV[(c.ab + 2) % 3]);
  R* aupp = P<R>(f.AU[(c.ab + 1) % 3]);
  R* avpp = P<R>(f.AV[(c.ab + 1) % 3]);
  (void)n3;
  if (c.timing) (void)hipEventRecord(c.ev[2], c.stream);
  // (schr too: the next step's int_readyc would form the same)
  LAUNCH(k_sch, n2, c.gd, h.chc, h.chp, h.mch, h.schr, h.sqp, h.schs, h.schm);
  // int_clin
  {
    ClinArgs<E> ca;
    ca.g = ge; ca.uc = U; ca.vc = V;
    ca.auc = au; ca.aup = aup; ca.aupp = aupp; ca.avc = av; ca.avp = avp; ca.avpp = avpp;
    ca.fx = f.fx; ca.fy = f.fy; ca.adp = P<E>(f.adp); ca.alpha = P<E>(f.alpha);
    ca.mch = h.mch; ca.schm = h.schm; ca.pax = h.pax; ca.pay = h.pay; ca.cor = c.gd.cor;
    ca.bc = (double)c.dtbc;
    hipLaunchKernelGGL((k_clin<R, E>), dim3((unsigned)(nblocks(n2) * c.nk)), dim3(kBlock), 0, c.stream, ca);
    hipError_t ec = hipGetLastError();
    if (ec != hipSuccess) return hip_err(ec, "k_clin");
  }
  {
    const size_t lds = 2 * (size_t)c.nk * kColCW * sizeof(double);
    hipLaunchKernelGGL(k_clin_cpl, dim3((unsigned)((c.n2 + kColCW - 1) / kColCW)), dim3(kColCW, kColKL), lds,
                       c.stream, c.gd, U, V, h.buc, h.bvc);
    hipError_t ep = hipGetLastError();
    if (ep != hipSuccess) return hip_err(ep, "k_clin_cpl");
  }
  // int_ts's upwelling (its w accumulator is convect's)
  LAUNCH_UPW();
  // int_ts and convect on the side stream once the upwelling is done, and
  // int_ssh ahead of them there or beside them on this stream: int_ts does
  // not read the alpha that int_ssh writes, convect does, and on this
  // stream int_ssh is followed by fork 2, which convect waits for.  Until
  // the next step's int_readyc join nothing on c.stream writes what the
  // side stream reads (T, S, uc, vc, wm, alpha after int_ssh, mch, mch2,
  // chc, schs, bnd, prho) or touches what it writes (T, S, the T/S/rho/w
  // accumulators, alphn, adpn), so int_ts and convect overlap smag and the
  // next step's int_readyc; a daily forcing update (which writes bnd and
  // prho) settles the side stream first.  The side stream is the longer in
  // the mixed forms (their int_ts keeps the FP64 state, their int_readyc
  // kernels are FP32), the main stream in FP64, so int_ssh joins the
  // shorter one.
  hipError_t e;
  if ((e = fork_side(c, 1)) != hipSuccess) return hip_err(e, "fork");
  c.side_pending = true;
  constexpr bool kSshMain = !std::is_same<R, double>::value;
  if (!kSshMain) LAUNCH_COLS_ON(c.side, k_ssh<E>, sizeof(E), ge, T, S, h.mch, P<E>(f.alpha), h.ssh, h.acssh);
  // int_ts
  c.n_w += 1;
  {
    TsArgs<R> a;
    a.g = g; a.t0 = T; a.s0 = S; a.wm = P<R>(f.wm); a.uc = U; a.vc = V;
    a.schs = h.schs; a.mch = h.mch; a.bndt = h.bnd[2]; a.bnds = h.bnd[3];
    a.t1 = P<double>(f.T[1 - c.cur]); a.s1 = P<double>(f.S[1 - c.cur]); a.bc = (double)c.dtbc;
    const int G = c.ni - 2;
    const unsigned nb = (unsigned)(((G + kTsTX - 1) / kTsTX) * ((c.nj + kTsTY - 1) / kTsTY) * c.nk);
    hipLaunchKernelGGL(k_ts_tiles<R>, dim3(nb), dim3(kTsTX * kTsTY), 0, c.side, a);
    hipError_t et = hipGetLastError();
    if (et != hipSuccess) return hip_err(et, "k_ts_tiles");
  }
  // int_ssh reads the T, S that int_ts leaves in place (it writes the other
  // buffer of each pair)
  if (kSshMain) {
    LAUNCH_COLS_ON(c.stream, k_ssh<E>, sizeof(E), ge, T, S, h.mch, P<E>(f.alpha), h.ssh, h.acssh);
    if ((e = fork_side(c, 2)) != hipSuccess) return hip_err(e, "fork");
  }
  c.n_ssh += 1;
  {
    const int G = c.ni - 2;
    const unsigned nb = (unsigned)(((G + kSmagTX - 1) / kSmagTX) * ((c.nj + kSmagTY - 1) / kSmagTY) * c.nk);
    hipLaunchKernelGGL(k_smag<R>, dim3(nb), dim3(kSmagTX * kSmagTY), 0, c.stream, g, U, V, h.mch2, h.schs,
                       P<R>(f.acu), P<R>(f.acv), P<R>(f.am), c.dtbc);
    hipError_t es = hipGetLastError();
    if (es != hipSuccess) return hip_err(es, "k_smag");
  }
  c.n_uv += 1;
  c.cur = 1 - c.cur;
  {
    const size_t lds = 2 * (size_t)c.nk * kConvCW * (sizeof(double) + sizeof(E));
    if (lds > 65536) {
      hipError_t ea =
          hipFuncSetAttribute((const void*)k_conv<R, E>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      if (ea != hipSuccess) return hip_err(ea, "k_conv LDS size");
    }
    // convect and the T/S/rho accumulators (reads chc, which nothing
    // writes before the next subcycle)
    hipLaunchKernelGGL((k_conv<R, E>), dim3((unsigned)((c.n2 + kConvCW - 1) / kConvCW)), dim3(kConvCW, kConvKL), lds,
                       c.side, ge,
                       c.gd.dpr2, P<double>(f.T[c.cur]), P<double>(f.S[c.cur]), P<E>(f.prho1), P<E>(f.prho2), h.chc,
                       P<R>(f.act), P<R>(f.acs), P<R>(f.acr), P<R>(f.acw), (const R*)P<R>(f.wm), h.mch2,
                       (const E*)P<E>(f.alpha), P<E>(f.alphn), P<E>(f.adpn));
    hipError_t e2 = hipGetLastError();
    if (e2 != hipSuccess) return hip_err(e2, "k_conv");
    c.next_ready = true;
  }
  c.n_ts += 1;
  if (c.timing) {
    std::string es = settle_side(c);
    if (!es.empty()) return es;
    (void)hipEventRecord(c.ev[3], c.stream);
    (void)hipEventSynchronize(c.ev[3]);
    float a = 0, b = 0, d = 0;
    (void)hipEventElapsedTime(&a, c.ev[0], c.ev[3]);
    (void)hipEventElapsedTime(&b, c.ev[1], c.ev[2]);
    (void)hipEventElapsedTime(&d, c.ev[0], c.ev[1]);
    float d2 = 0;
    (void)hipEventElapsedTime(&d2, c.ev[2], c.ev[3]);
    c.t_ms[0] += a;
    c.t_ms[1] += d + d2;
    c.t_ms[2] += b;
    c.t_steps += 1;
  }
  // calendar (main.f90:225-226)
  c.pt = c.ct;
  c.ct = time_plus(c.ct, c.dtbc);
  return "";
}

// ---- single operators (FP64) for parity tests ------------------------------
template <int GA, int GB>
__global__ void k_op_ter(Grid<double> g, int nd3, double dft, const double* in, double* out) {
  long p = (long)xcd_block() * blockDim.x + threadIdx.x;
  if (p >= g.n2 * nd3) return;
  int i = (int)(p % g.ni) + 1;
  long r = p / g.ni;
  int j = (int)(r % g.nj) + 1, k = (int)(r / g.nj) + 1;
  const double* v = in + g.n2 * (k - 1);
  auto F = [&](int m, int jj) -> double { return v[g.ix(m, jj)]; };
  double o;
  if constexpr (gew(GA) == GB) o = ter_ew<GA>(g, F, i, j);
  else if constexpr (gns(GA) == GB) o = ter_ns<GA>(g, F, i, j, dft);
  else o = ter_di<GA>(g, F, i, j, dft);
  out[p] = o;
}

template <int DIR, int GA, int GB>
__global__ void k_op_px(Grid<double> g, int nd3, const double* in, double* out) {
  long p = (long)xcd_block() * blockDim.x + threadIdx.x;
  if (p >= g.n2 * nd3) return;
  int i = (int)(p % g.ni) + 1;
  long r = p / g.ni;
  int j = (int)(r % g.nj) + 1, k = (int)(r / g.nj) + 1;
  const double* v = in + g.n2 * (k - 1);
  auto F = [&](int m, int jj) -> double { return v[g.ix(m, jj)]; };
  out[p] = DIR == 1 ? p1<GA, GB>(g, F, i, j) : p2<GA, GB>(g, F, i, j);
}

template <int GA, int GB, int GC>
__global__ void k_op_div(Grid<double> g, int nd3, const double* va, const double* vb, double* out) {
  long p = (long)xcd_block() * blockDim.x + threadIdx.x;
  if (p >= g.n2 * nd3) return;
  int i = (int)(p % g.ni) + 1;
  long r = p / g.ni;
  int j = (int)(r % g.nj) + 1, k = (int)(r / g.nj) + 1;
  const double* a = va + g.n2 * (k - 1);
  const double* b = vb + g.n2 * (k - 1);
  auto FA = [&](int m, int jj) -> double { return a[g.ix(m, jj)]; };
  auto FB = [&](int m, int jj) -> double { return b[g.ix(m, jj)]; };
  out[p] = divg<GA, GB, GC>(g, FA, FB, i, j);
}

__global__ void k_op_lap33(Grid<double> g, const double* in, double* out) {
  long p = (long)xcd_block() * blockDim.x + threadIdx.x;
  if (p >= g.n2) return;
  int i = (int)(p % g.ni) + 1, j = (int)(p / g.ni) + 1;
  auto F = [&](int m, int jj) -> double { return in[g.ix(m, jj)]; };
  out[p] = lap<3, 3>(g, F, i, j);
}

__global__ void k_den_rho(long n, const double* t, const double* s, const double* pr, double* rho) {
  long p = (long)xcd_block() * blockDim.x + threadIdx.x;
  if (p < n) rho[p] = den_rho<double>(t[p], s[p], pr[p]);
}

}  // namespace

std::string launch_step_pre(Ctx& c) {
  if (c.precision == PCOM_FP64) return step_pre_t<double, double>(c);
  if (c.precision == PCOM_MIXED) return step_pre_t<float, double>(c);
  return step_pre_t<float, float>(c);
}

std::string launch_step_post(Ctx& c) {
  if (c.precision == PCOM_FP64) return step_post_t<double, double>(c);
  if (c.precision == PCOM_MIXED) return step_post_t<float, double>(c);
  return step_post_t<float, float>(c);
}

std::string settle_side(Ctx& c) {
  if (!c.side_pending) return "";
  hipError_t e = join_side(c, 2);
  if (e != hipSuccess) return hip_err(e, "join");
  c.side_pending = false;
  return "";
}

namespace {
template <typename E>
std::string set_forcing_t(Ctx& c, const double* const rec[6], bool interpolate) {
  Fields2& h = c.f2;
  const size_t n2 = c.n2, n3 = c.n3;
  if (interpolate) {
    int m1, m2;
    double fw;
    bnd_weights(c.ct, &m1, &m2, &fw);
    Six s6;
    for (int v = 0; v < 6; ++v) {
      s6.a[v] = h.frc[v];
      s6.b[v] = h.bnd[v];
    }
    LAUNCH(k_bnd, n2, (long)n2, s6, m1, m2, fw);
  }
  for (int v = 0; v < 6; ++v)
    if (rec[v]) LAUNCH(k_put_forcing, n2, c.gd, rec[v], h.bnd[v]);
  LAUNCH(k_prho<E>, n3, grid_of<E>(c), P<double>(c.f3.T[c.cur]), P<double>(c.f3.S[c.cur]), h.chc,
         P<E>(c.f3.prho1), P<E>(c.f3.prho2));
  c.pa_stale = true;
  return "";
}
}  // namespace

std::string launch_set_forcing(Ctx& c, const double* const rec[6], bool interpolate) {
  if (c.precision == PCOM_MIXED_EOS32) return set_forcing_t<float>(c, rec, interpolate);
  return set_forcing_t<double>(c, rec, interpolate);
}

std::string launch_step(Ctx& c) {
  std::string e = launch_step_pre(c);
  if (!e.empty()) return e;
  e = c.nn ? nn_trop_step(c, launch_trop_kernel) : launch_trop_kernel(c);
  if (!e.empty()) return e;
  c.n_ch += 1;
  return launch_step_post(c);
}

std::string launch_global_sums(Ctx& c, double out[3]) {
  const size_t n2 = c.n2;
  unsigned nb = nblocks(n2);
  double* area = c.f2.red;              // n2
  double* part = c.f2.red + n2;         // 3 * nb
  double* res = part + 3 * (size_t)nb;  // 3
  LAUNCH(k_cell_area, n2, c.gd, area);
  LAUNCH(k_sums_partial<double>, n2, c.gd, P<double>(c.f3.T[c.cur]), P<double>(c.f3.S[c.cur]), c.f2.chc, area, part);
  hipLaunchKernelGGL(k_sums_final, dim3(1), dim3(64), 0, c.stream, part, (int)nb, res);
  hipError_t e = hipMemcpyAsync(out, res, 3 * sizeof(double), hipMemcpyDeviceToHost, c.stream);
  if (e != hipSuccess) return hip_err(e, "copy sums");
  e = hipStreamSynchronize(c.stream);
  return hip_err(e, "sync sums");
}

std::string launch_check_finite(Ctx& c, int* bad) {
  hipError_t e = hipMemsetAsync(c.f2.flag, 0, sizeof(int), c.stream);
  if (e != hipSuccess) return hip_err(e, "memset flag");
  void* f3s[] = {c.f3.T[c.cur], c.f3.S[c.cur], c.f3.U, c.f3.V};
  for (void* a : f3s) LAUNCH(k_finite<double>, c.n3, P<double>(a), (long)c.n3, c.f2.flag);
  double* f2s[] = {c.f2.chc, c.f2.buc, c.f2.bvc};
  for (double* a : f2s) LAUNCH(k_finite_d, c.n2, a, (long)c.n2, c.f2.flag);
  e = hipMemcpyAsync(bad, c.f2.flag, sizeof(int), hipMemcpyDeviceToHost, c.stream);
  if (e != hipSuccess) return hip_err(e, "copy flag");
  return hip_err(hipStreamSynchronize(c.stream), "sync flag");
}

template <typename R>
std::string output_means_t(Ctx& c, pcom_means* o, double* stage) {
  const Grid<R>& g = grid_of<R>(c);
  Fields3& f = c.f3;
  const long gni = c.ni - 2, gnj = c.nj - 2;
  const size_t m3 = (size_t)(gni * gnj * c.nk), m2 = (size_t)(gni * gnj);
  auto out3 = [&](void* acc, int n, int mask, double* host) -> std::string {
    if (!host) return "";
    LAUNCH(k_mean3<R>, m3, g, P<R>(acc), n, mask, stage);
    hipError_t e = hipMemcpyAsync(host, stage, m3 * sizeof(double), hipMemcpyDeviceToHost, c.stream);
    if (e != hipSuccess) return hip_err(e, "copy means");
    return hip_err(hipStreamSynchronize(c.stream), "sync means");
  };
  std::string err;
  if (!(err = out3(f.act, c.n_ts, 0, o->pt)).empty()) return err;
  if (!(err = out3(f.acs, c.n_ts, 0, o->sa)).empty()) return err;
  if (!(err = out3(f.acr, c.n_ts, 1, o->rho)).empty()) return err;
  if (!(err = out3(f.acu, c.n_uv, 2, o->u)).empty()) return err;
  if (!(err = out3(f.acv, c.n_uv, 2, o->v)).empty()) return err;
  if (!(err = out3(f.acw, c.n_w, 2, o->w)).empty()) return err;
  auto out2 = [&](const double* acc, int n, bool anomaly, double* host) -> std::string {
    if (!host) return "";
    LAUNCH(k_mean2, m2, c.gd, acc, n, stage);
    if (anomaly) {
      double* sums = stage + m2;
      hipLaunchKernelGGL(k_ssh_mean, dim3(1), dim3(64), 0, c.stream, c.gd, stage, sums);
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) return hip_err(e, "k_ssh_mean");
      LAUNCH(k_ssh_anom, m2, (long)m2, sums, stage);
    }
    hipError_t e = hipMemcpyAsync(host, stage, m2 * sizeof(double), hipMemcpyDeviceToHost, c.stream);
    if (e != hipSuccess) return hip_err(e, "copy means");
    return hip_err(hipStreamSynchronize(c.stream), "sync means");
  };
  if (!(err = out2(c.f2.acch, c.n_ch, false, o->ch)).empty()) return err;
  if (!(err = out2(c.f2.acssh, c.n_ssh, true, o->ssh)).empty()) return err;
  // reset accumulators and counts
  void* acc3[] = {f.act, f.acs, f.acr, f.acu, f.acv, f.acw};
  for (void* a : acc3) LAUNCH(k_fill<R>, c.n3, P<R>(a), (long)c.n3);
  LAUNCH(k_fill<double>, c.n2, c.f2.acch, (long)c.n2);
  LAUNCH(k_fill<double>, c.n2, c.f2.acssh, (long)c.n2);
  c.n_ts = c.n_uv = c.n_w = c.n_ch = c.n_ssh = 0;
  return hip_err(hipStreamSynchronize(c.stream), "sync reset");
}

std::string launch_output_means(Ctx& c, pcom_means* o, double* stage) {
  if (c.precision == PCOM_FP64) return output_means_t<double>(c, o, stage);
  return output_means_t<float>(c, o, stage);
}

#define TER_CASE(A, B)                                                              \
  case A * 10 + B:                                                                  \
    LAUNCH((k_op_ter<A, B>), n, c.gd, nd3, dft, in, out);                           \
    return "";

std::string launch_op_ter(Ctx& c, int ga, int gb, int nd3, double dft, const double* in, double* out) {
  size_t n = c.n2 * (size_t)nd3;
  switch (ga * 10 + gb) {
    TER_CASE(1, 2) TER_CASE(1, 3) TER_CASE(1, 4)
    TER_CASE(2, 1) TER_CASE(2, 3) TER_CASE(2, 4)
    TER_CASE(3, 1) TER_CASE(3, 2) TER_CASE(3, 4)
    TER_CASE(4, 1) TER_CASE(4, 2) TER_CASE(4, 3)
    default:
      return "op_ter: staggers must differ and lie in 1..4";
  }
}

#define PX_CASE(A, B)                                                                       \
  case A * 10 + B:                                                                          \
    if (dir == 1) LAUNCH((k_op_px<1, A, B>), n, c.gd, nd3, in, out);                        \
    else LAUNCH((k_op_px<2, A, B>), n, c.gd, nd3, in, out);                                 \
    return "";

std::string launch_op_px(Ctx& c, int dir, int ga, int gb, int nd3, const double* in, double* out) {
  size_t n = c.n2 * (size_t)nd3;
  switch (ga * 10 + gb) {
    PX_CASE(1, 1) PX_CASE(1, 2) PX_CASE(1, 3) PX_CASE(1, 4)
    PX_CASE(2, 1) PX_CASE(2, 2) PX_CASE(2, 3) PX_CASE(2, 4)
    PX_CASE(3, 1) PX_CASE(3, 2) PX_CASE(3, 3) PX_CASE(3, 4)
    PX_CASE(4, 1) PX_CASE(4, 2) PX_CASE(4, 3) PX_CASE(4, 4)
    default:
      return "op_px: staggers must lie in 1..4";
  }
}

std::string launch_op_div(Ctx& c, int ga, int gb, int gc, int nd3, const double* va, const double* vb, double* out) {
  size_t n = c.n2 * (size_t)nd3;
  int code = ga * 100 + gb * 10 + gc;
  switch (code) {
    case 331: LAUNCH((k_op_div<3, 3, 1>), n, c.gd, nd3, va, vb, out); return "";
    case 423: LAUNCH((k_op_div<4, 2, 3>), n, c.gd, nd3, va, vb, out); return "";
    case 241: LAUNCH((k_op_div<2, 4, 1>), n, c.gd, nd3, va, vb, out); return "";
    default:
      return "op_div: supported (ga,gb,gc) are the reference's (3,3,1), (4,2,3), (2,4,1)";
  }
}

std::string launch_op_lap(Ctx& c, int ga, int gb, const double* in, double* out) {
  if (ga != 3 || gb != 3) return "op_lap: the reference applies it on g3 -> g3 only (mod_int.f90:198-199)";
  LAUNCH(k_op_lap33, c.n2, c.gd, in, out);
  return "";
}

std::string launch_den_rho(hipStream_t st, size_t n, const double* t, const double* s, const double* p, double* rho) {
  hipLaunchKernelGGL(k_den_rho, dim3(nblocks(n)), dim3(kBlock), 0, st, (long)n, t, s, p, rho);
  return hip_err(hipGetLastError(), "k_den_rho");
}

}  // namespace pcom
