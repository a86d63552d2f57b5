// pcom_dev.h - device-side grid description, constants, equation of state
// and the staggered-grid operators of the PCOM B-grid (ouyuyuan/pcom v2.2,
// mod_op.f90) evaluated pointwise.
//
// Every reference operator fills a whole halo'd array, then the one-rank
// halo swap copies column ni-1 into column 1 and column 2 into column ni
// (mod_mympi.f90:669-746), leaving the wall rows 1 and nj as computed or
// prefilled.  An operator's value at (i, j) is therefore its formula at the
// canonical column ic = canon(i) (1 -> ni-1, ni -> 2, else i) whenever the
// inputs are zonally periodic, and the prefill (dft) on the rows outside
// its write range.  The functions below return exactly that value, so that
// fused kernels can compose operators without materialising intermediates
// and still perform, per point, the same IEEE operations in the same order
// as the reference's array statements.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pcom {

// ---- constants: mod_con.f90:16-41 as the reference's active build
// (gfortran without -r8) evaluates them: default-REAL literals are float.
constexpr double RHO0 = 1029.0;
constexpr double A0 = 1.0 / 1029.0;
constexpr double GRAV = (double)9.8f;
constexpr double ARAD = 6370e3;
constexpr double PI_F = (double)3.141592653589793f;
constexpr double OMEGA = PI_F / 43082.0;
constexpr double TORAD = PI_F / 180.0;
constexpr double GAMMA_T = GRAV * 40.0 / 3901.0;
constexpr double GAMMA_S = (double)(1.0f / (float)(120 * 24 * 60 * 60));
constexpr double GAMMA_B = (double)0.2f;
constexpr double DELTAT = (double)1.0e-2f;
constexpr double DELTAS = (double)1.0e-3f;
constexpr double CDBOT = (double)2.6f;
constexpr double TICE = -1.5;
constexpr double KM_C = (double)1.0e-5f;
constexpr double AM_C = 1.0e3;
constexpr double SMAG_C = (double)0.56f;
constexpr double AH_C = 1.0e3;

// ---- XCD-aware workgroup order.  MI355X deals the workgroups of a launch
// round-robin to its 8 XCDs (workgroup b runs on XCD b % 8), each with its
// own L2.  Stencil kernels read the rows next to their own, so consecutive
// workgroup ids are remapped to consecutive tiles of ONE XCD: XCD x takes
// the contiguous range [start(x), start(x+1)) of logical tiles.  The map is
// a bijection of [0, n) for any n.
constexpr unsigned kXcds = 8;
// a / b in the stencils' type: IEEE division in FP64 (the FP64 form is the
// reference's arithmetic), and in FP32 a hardware reciprocal refined by one
// residual step (Markstein, IBM J. Res. Dev. 34(1), 1990: q = a r,
// q + (a - q b) r), within an ulp of the IEEE quotient for the normal,
// nonzero divisors of the stencils, at a third of the instructions
__device__ __forceinline__ double dv(double a, double b) { return a / b; }
__device__ __forceinline__ float dv(float a, float b) {
  const float r = __builtin_amdgcn_rcpf(b);
  const float q = a * r;
  return __builtin_fmaf(__builtin_fmaf(-q, b, a), r, q);
}

__device__ __forceinline__ unsigned xcd_block() {
  const unsigned b = blockIdx.x, n = gridDim.x;
  const unsigned q = n / kXcds, rem = n % kXcds;
  const unsigned x = b % kXcds, r = b / kXcds;
  return x * q + (x < rem ? x : rem) + r;
}

// ---- stagger tables (mod_arrays.f90:148-166, link_grid 524-552):
// offsets g%i, g%j and the ew / ns / di neighbours of g1..g4.
__host__ __device__ constexpr int goi(int g) { return (g == 2 || g == 3) ? 1 : 0; }
__host__ __device__ constexpr int goj(int g) { return (g == 3 || g == 4) ? 1 : 0; }
__host__ __device__ constexpr int gew(int g) { return g == 1 ? 2 : g == 2 ? 1 : g == 3 ? 4 : 3; }
__host__ __device__ constexpr int gns(int g) { return g == 1 ? 4 : g == 2 ? 3 : g == 3 ? 2 : 1; }
__host__ __device__ constexpr int gdi(int g) { return g == 1 ? 3 : g == 2 ? 4 : g == 3 ? 1 : 2; }

// Static fields of the grid, resident on the device.  R is the arithmetic
// type of the 3-D stencils (double, or float in the mixed form); metric
// arrays exist in both types.
template <typename R>
struct Grid {
  using real = R;
  int ni, nj, nk, nkp, nim, njm;
  long n2;
  const R *rh[5], *tn[5], *dx1[5], *dx2[5];  // g1j..g4j (index 1..4)
  const int *lev_t, *lev_u;                   // g12%lev, g32%lev
  const double *phih_t, *prh_t, *prh_u;       // g12%phih, g12%prh, g32%prh
  const R *prh_t_r, *prh_u_r;                 // same, in R
  const R *cv1, *cv2, *cor;                   // calc_cf, Coriolis (g3)
  const R *pr2, *dpr2, *dz1;                  // gi2%pr, gi2%dpr, gi1%dz
  const R *km, *kh;                           // (ni,nj,nk+1)

  __device__ __forceinline__ long ix(int i, int j) const { return (long)(i - 1) + (long)ni * (j - 1); }
  __device__ __forceinline__ long ix(int i, int j, int k) const { return ix(i, j) + n2 * (long)(k - 1); }
  // canonical column: the halo columns are copies of ni-1 and 2
  __device__ __forceinline__ int canon(int i) const { return i == 1 ? nim : (i == ni ? 2 : i); }
  // g12%msk and g32%msk equal (lev >= k): the reference sets them from lev
  // on the interior and swaps (main.f90:601-625); wall rows have lev = 0.
  __device__ __forceinline__ int mskt(int i, int j, int k) const { return lev_t[ix(i, j)] >= k ? 1 : 0; }
  __device__ __forceinline__ int msku(int i, int j, int k) const { return lev_u[ix(i, j)] >= k ? 1 : 0; }
  __device__ __forceinline__ R RH(int g, int i, int j) const { return rh[g][ix(i, j)]; }
  __device__ __forceinline__ R DX1(int g, int i, int j) const { return dx1[g][ix(i, j)]; }
  __device__ __forceinline__ R DX2(int g, int i, int j) const { return dx2[g][ix(i, j)]; }
};

// ---- mod_den.f90:57-118, UNESCO-1981 rho0 with the Jackett-McDougall
// (1995) secant bulk modulus; p0 in Pa without the surface atmosphere.
// Coefficients are the float32 values of the reference's literals.
template <typename R>
__device__ __host__ inline R den_rho(R t, R s, R p0) {
  R p = p0 * (R)1e-5f;
  R p2 = p * p;
  R t2 = t * t;
  R t3 = t2 * t;
  R t4 = t3 * t;
  R t5 = t4 * t;
  R s32 = s * sqrt(s);
  R s2 = s * s;
  R r0 = (R)999.842594f + (R)6.793952e-2f * t - (R)9.095290e-3f * t2 + (R)1.001685e-4f * t3 -
         (R)1.120083e-6f * t4 + (R)6.536336e-9f * t5;
  R a = (R)8.24493e-1f - (R)4.0899e-3f * t + (R)7.6438e-5f * t2 - (R)8.2467e-7f * t3 + (R)5.3875e-9f * t4;
  R b = (R)-5.72466e-3f + (R)1.0227e-4f * t - (R)1.6546e-6f * t2;
  R r = r0 + a * s + b * s32 + (R)4.8314e-4f * s2;
  R k = (R)1.965933e4f + (R)1.444304e2f * t - (R)1.706103f * t2 + (R)9.648704e-3f * t3 - (R)4.190253e-5f * t4 +
        ((R)5.284855e1f - (R)3.101089e-1f * t + (R)6.283263e-3f * t2 - (R)5.084188e-5f * t3) * s +
        ((R)3.886640e-1f + (R)9.085835e-3f * t - (R)4.619924e-4f * t2) * s32 +
        ((R)3.186519f + (R)2.212276e-2f * t - (R)2.984642e-4f * t2 + (R)1.956415e-6f * t3) * p +
        ((R)6.704388e-3f - (R)1.847318e-4f * t + (R)2.059331e-7f * t2) * s * p + (R)1.480266e-4f * s32 * p +
        ((R)2.102898e-4f - (R)1.202016e-5f * t + (R)1.394680e-7f * t2) * p2 +
        ((R)-2.040237e-6f + (R)6.128773e-8f * t + (R)6.207323e-10f * t2) * s * p2;
  return r / ((R)1.0 - p / k);
}

// ---- pointwise operators.  F is a callable f(i, j) -> value of the
// (periodic) input on its own stagger at column i in [1, ni], row j.

// ter_r3d/ter_r2d, hga -> hga%ew (mod_op.f90:376-378): all rows written.
template <int GA, typename G, typename F, typename R = typename G::real>
__device__ __forceinline__ R ter_ew(const G& g, F f, int i, int j) {
  constexpr int oi = goi(GA);
  int ic = g.canon(i);
  return (R)0.5 * (f(ic - oi, j) + f(ic - oi + 1, j));
}
// hga -> hga%ns (mod_op.f90:379-381): rows 1+oj .. njm+oj, else dft.
template <int GA, typename G, typename F, typename R = typename G::real>
__device__ __forceinline__ R ter_ns(const G& g, F f, int i, int j, R dft) {
  constexpr int oj = goj(GA);
  if (j < 1 + oj || j > g.njm + oj) return dft;
  int ic = g.canon(i);
  return (R)0.5 * (f(ic, j - oj) + f(ic, j - oj + 1));
}
// hga -> hga%di (mod_op.f90:382-385)
template <int GA, typename G, typename F, typename R = typename G::real>
__device__ __forceinline__ R ter_di(const G& g, F f, int i, int j, R dft) {
  constexpr int oi = goi(GA), oj = goj(GA);
  if (j < 1 + oj || j > g.njm + oj) return dft;
  int ic = g.canon(i);
  return (f(ic - oi, j - oj) + f(ic - oi, j - oj + 1) + f(ic - oi + 1, j - oj) + f(ic - oi + 1, j - oj + 1)) *
         (R)0.25;
}

// p1_r3d/p1_r2d when gb%ew == ga (no pre-interpolation), mod_op.f90:
// 496-527: written on all rows, i in 1+ga%i .. nim+ga%i; divided by
// gb%dx%x(1) everywhere.  `tf` is the source on stagger GA.
template <int GA, int GB, typename G, typename F, typename R = typename G::real>
__device__ __forceinline__ R p1_direct(const G& g, F tf, int i, int j) {
  static_assert(gew(GB) == GA, "p1_direct needs gb%ew == ga");
  constexpr int oi = goi(GA);
  int ic = g.canon(i);
  return (tf(ic - oi + 1, j) - tf(ic - oi, j)) / g.DX1(GB, ic, j);
}
// p1 with pre-interpolation ter(var, ga -> gb%ew) (mod_op.f90:510-511).
template <int GA, int GB, typename G, typename F, typename R = typename G::real>
__device__ __forceinline__ R p1_interp(const G& g, F f, int i, int j) {
  constexpr int GT = gew(GB);
  static_assert(GT != GA, "use p1_direct");
  constexpr int oi = goi(GA);
  int ic = g.canon(i);
  auto tmp = [&](int m, int jj) -> R {
    if constexpr (gew(GA) == GT) return ter_ew<GA>(g, f, m, jj);
    else if constexpr (gns(GA) == GT) return ter_ns<GA>(g, f, m, jj, (R)0);
    else return ter_di<GA>(g, f, m, jj, (R)0);
  };
  return (tmp(ic - oi + 1, j) - tmp(ic - oi, j)) / g.DX1(GB, ic, j);
}
// p2_r3d/p2_r2d when gb%ns == ga (mod_op.f90:622-652): rows
// 1+ga%j .. njm+ga%j, zero elsewhere, divided by gb%dx%x(2).
template <int GA, int GB, typename G, typename F, typename R = typename G::real>
__device__ __forceinline__ R p2_direct(const G& g, F tf, int i, int j) {
  static_assert(gns(GB) == GA, "p2_direct needs gb%ns == ga");
  constexpr int oj = goj(GA);
  int ic = g.canon(i);
  R d = (j < 1 + oj || j > g.njm + oj) ? (R)0 : (tf(ic, j - oj + 1) - tf(ic, j - oj));
  return d / g.DX2(GB, ic, j);
}
template <int GA, int GB, typename G, typename F, typename R = typename G::real>
__device__ __forceinline__ R p2_interp(const G& g, F f, int i, int j) {
  constexpr int GT = gns(GB);
  static_assert(GT != GA, "use p2_direct");
  constexpr int oj = goj(GA);
  int ic = g.canon(i);
  auto tmp = [&](int m, int jj) -> R {
    if constexpr (gew(GA) == GT) return ter_ew<GA>(g, f, m, jj);
    else if constexpr (gns(GA) == GT) return ter_ns<GA>(g, f, m, jj, (R)0);
    else return ter_di<GA>(g, f, m, jj, (R)0);
  };
  R d = (j < 1 + oj || j > g.njm + oj) ? (R)0 : (tmp(ic, j - oj + 1) - tmp(ic, j - oj));
  return d / g.DX2(GB, ic, j);
}

template <int GA, int GB, typename G, typename F, typename R = typename G::real>
__device__ __forceinline__ R p1(const G& g, F f, int i, int j) {
  if constexpr (gew(GB) == GA) return p1_direct<GA, GB>(g, f, i, j);
  else return p1_interp<GA, GB>(g, f, i, j);
}
template <int GA, int GB, typename G, typename F, typename R = typename G::real>
__device__ __forceinline__ R p2(const G& g, F f, int i, int j) {
  if constexpr (gns(GB) == GA) return p2_direct<GA, GB>(g, f, i, j);
  else return p2_interp<GA, GB>(g, f, i, j);
}

// div_r3d / div_r2d (mod_op.f90:813-848): (va on GA, vb on GB) -> GC.
template <int GA, int GB, int GC, typename G, typename FA, typename FB, typename R = typename G::real>
__device__ __forceinline__ R divg(const G& g, FA fa, FB fb, int i, int j) {
  auto vr = [&](int m, int jj) -> R { return fb(m, jj) * g.RH(GB, m, jj); };
  R a1 = p1<GA, GC>(g, fa, i, j);
  R b1 = p2<GB, GC>(g, vr, i, j);
  int ic = g.canon(i);
  return (a1 + b1) / ((R)ARAD * g.RH(GC, ic, j));
}

// gra x and y parts (mod_op.f90:850-902): p1/(a*gb%rh), p2/a.
template <int GA, int GB, typename G, typename F, typename R = typename G::real>
__device__ __forceinline__ R grax(const G& g, F f, int i, int j) {
  int ic = g.canon(i);
  return p1<GA, GB>(g, f, i, j) / ((R)ARAD * g.RH(GB, ic, j));
}
template <int GA, int GC, typename G, typename F, typename R = typename G::real>
__device__ __forceinline__ R gray(const G& g, F f, int i, int j) {
  return p2<GA, GC>(g, f, i, j) / (R)ARAD;
}

// op_lap (mod_op.f90:904-925) for 2-D var on GA, result on GB.
template <int GA, int GB, typename G, typename F, typename R = typename G::real>
__device__ __forceinline__ R lap(const G& g, F f, int i, int j) {
  constexpr int GE = gew(GB), GN = gns(GB);
  auto wa = [&](int m, int jj) -> R {
    int mc = g.canon(m);
    return p1<GA, GE>(g, f, m, jj) / g.RH(GE, mc, jj);
  };
  auto wb = [&](int m, int jj) -> R {
    int mc = g.canon(m);
    return p2<GA, GN>(g, f, m, jj) * g.RH(GN, mc, jj);
  };
  R x = p1<GE, GB>(g, wa, i, j);
  R y = p2<GN, GB>(g, wb, i, j);
  int ic = g.canon(i);
  return (x + y) / ((R)ARAD * (R)ARAD * g.RH(GB, ic, j));
}

}  // namespace pcom
