// setup.cpp - the start-up chain of main.f90 (set_cf, calc_hpos, set_hgrid,
// set_vgrid, set_grid, calc_cf, Coriolis, inistat), run once on the host
// when a context is created.  Arithmetic follows the reference statements
// in order (FP64, float32 literals), so the static fields are the ones the
// reference's own start-up produces for the same inputs.
#include "setup.h"

#include <cmath>
#include <cstring>

#include "pcom_dev.h"

namespace pcom {

int days_in_month(int y, int m) {
  static const int dm[12] = {31, 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31};
  int d = dm[m - 1];
  if (m == 2) {
    bool leap = (y % 100 == 0) ? (y % 400 == 0) : (y % 4 == 0);
    if (leap) d = 29;
  }
  return d;
}

Calendar time_plus(Calendar t, int dt) {
  t.s += dt;
  if (t.s >= 60) { t.mi += t.s / 60; t.s %= 60; }
  if (t.mi >= 60) { t.h += t.mi / 60; t.mi %= 60; }
  if (t.h >= 24) { t.d += t.h / 24; t.h %= 24; }
  if (t.d > t.mm) { t.m += 1; t.d = 1; }
  if (t.m > 12) { t.y += 1; t.m = 1; }
  t.mm = days_in_month(t.y, t.m);
  return t;
}

void bnd_weights(const Calendar& t, int* m1, int* m2, double* f) {
  float ff;
  if (t.d < 15) {
    *m2 = t.m;
    *m1 = t.m == 1 ? 12 : t.m - 1;
    int days = days_in_month(t.y, *m1);
    ff = ((float)days - 15.0f + (float)t.d) / ((float)days - 15.0f + 15.0f);
  } else {
    *m1 = t.m;
    *m2 = t.m == 12 ? 1 : t.m + 1;
    int days = days_in_month(t.y, *m1);
    ff = ((float)t.d - 15.0f) / ((float)(days - 15) + 15.0f);
  }
  *f = (double)ff;
}

namespace {

struct Dims {
  int ni, nj, nim, njm;
  size_t n2;
  size_t at(int i, int j) const { return (size_t)(i - 1) + (size_t)ni * (size_t)(j - 1); }
  size_t at(int i, int j, int k) const { return at(i, j) + n2 * (size_t)(k - 1); }
};

// E/W halo copy of one rank (mod_mympi.f90:669-746)
template <typename T>
void halo(const Dims& d, T* a, int levels) {
  for (int k = 1; k <= levels; ++k)
    for (int j = 1; j <= d.nj; ++j) a[d.at(1, j, k)] = a[d.at(d.ni - 1, j, k)];
  for (int k = 1; k <= levels; ++k)
    for (int j = 1; j <= d.nj; ++j) a[d.at(d.ni, j, k)] = a[d.at(2, j, k)];
}

// mympi_div: global interior block into (2:ni-1, 2:nj-1), then halo.
void scatter(const Dims& d, int glo_ni, int glo_nj, double* dst, const double* src, int levels) {
  for (int k = 1; k <= levels; ++k)
    for (int j = 2; j <= d.nj - 1; ++j)
      for (int i = 2; i <= d.ni - 1; ++i)
        dst[d.at(i, j, k)] = src[(size_t)(i - 2) + (size_t)glo_ni * ((size_t)(j - 2) + (size_t)glo_nj * (k - 1))];
  halo(d, dst, levels);
}

// 2-D host versions of ter (g1 -> g2 / g4) and p1 / p2 used by inistat
// (mod_op.f90:395-427, 597-620, 689-713, 864-876).
void ter2_ew_from_g1(const Dims& d, std::vector<double>& ans, const std::vector<double>& v, int oi) {
  std::fill(ans.begin(), ans.end(), 0.0);
  for (int j = 1; j <= d.nj; ++j)
    for (int i = 1 + oi; i <= d.nim + oi; ++i) ans[d.at(i, j)] = 0.5 * (v[d.at(i - oi, j)] + v[d.at(i - oi + 1, j)]);
  halo(d, ans.data(), 1);
}
void ter2_ns(const Dims& d, std::vector<double>& ans, const std::vector<double>& v, int oj) {
  std::fill(ans.begin(), ans.end(), 0.0);
  for (int j = 1 + oj; j <= d.njm + oj; ++j)
    for (int i = 1; i <= d.ni; ++i) ans[d.at(i, j)] = 0.5 * (v[d.at(i, j - oj)] + v[d.at(i, j - oj + 1)]);
  halo(d, ans.data(), 1);
}

}  // namespace

std::string build_setup(const pcom_inputs& in, HostSetup& hs) {
  if (in.glo_ni < 3 || in.glo_nj < 3 || in.nk < 2) return "grid too small (need glo_ni, glo_nj >= 3, nk >= 2)";
  if ((long long)(in.glo_ni + 2) * (in.glo_nj + 2) >= (1ll << 31))
    return "grid too large (the halo'd 2-D size must stay below 2^31 points)";
  if (!in.lon || !in.lat || !in.z || !in.itn || !in.tmn || !in.smn || !in.pt || !in.sa || !in.taux ||
      !in.tauy || !in.bct || !in.bcs || !in.pa || !in.fw)
    return "pcom_inputs has a NULL array (only vmix may be NULL)";
  hs.glo_ni = in.glo_ni;
  hs.glo_nj = in.glo_nj;
  hs.nk = in.nk;
  hs.nkp = in.nk + 1;
  hs.ni = in.glo_ni + 2;
  hs.nj = in.glo_nj + 2;
  Dims d{hs.ni, hs.nj, hs.ni - 1, hs.nj - 1, (size_t)hs.ni * hs.nj};
  const int ni = hs.ni, nj = hs.nj, nk = hs.nk, nkp = hs.nkp;
  hs.n2 = d.n2;
  hs.n3 = d.n2 * nk;
  hs.n3p = d.n2 * nkp;
  const size_t n2 = hs.n2;

  // ---- set_cf (main.f90:312-345).  pf(12) is written only while
  // k <= nk+1: the reference writes km(:,:,12) past the array end when
  // nk < 11 (main.f90:339-341); that store is skipped here.
  {
    static const float pf0[12] = {0.1f, 50.0f, 50.0f, 50.0f, 40.0f, 20.0f, 10.0f, 5.0f, 2.0f, 1.0f, 0.5f, 0.2f};
    hs.km.assign(hs.n3p, KM_C);
    for (int k = 1; k <= 12 && k <= nkp; ++k) {
      double pf = (double)pf0[k - 1] * (double)1.0e-4f;
      for (size_t p = 0; p < n2; ++p) hs.km[p + n2 * (k - 1)] = pf;
    }
    size_t gn = (size_t)in.glo_ni * in.glo_nj;
    std::vector<double> gk(gn * nkp, 0.0);
    for (int k = 2; k <= nkp; ++k)
      for (size_t p = 0; p < gn; ++p) gk[p + gn * (k - 1)] = in.vmix ? in.vmix[p + gn * (k - 2)] : 9.99999e30;
    for (double& v : gk) {
      if (v > (double)(9.0e+30f - 1.0f)) v = (double)0.1f;
      if (v > 200.0) v = 200.0;
      v = v * (double)1.0e-4f;
    }
    hs.kh.assign(hs.n3p, 0.0);
    scatter(d, in.glo_ni, in.glo_nj, hs.kh.data(), gk.data(), nkp);
  }

  // ---- coordinates (main.f90:292-308)
  std::vector<double> lon(ni), lat(nj);
  for (int i = 2; i <= ni - 1; ++i) lon[i - 1] = in.lon[i - 2];
  lon[0] = lon[ni - 2];
  lon[ni - 1] = lon[1];
  for (int j = 2; j <= nj - 1; ++j) lat[j - 1] = in.lat[j - 2];
  lat[0] = 2 * lat[1] - lat[2];
  lat[nj - 1] = 2 * lat[nj - 2] - lat[nj - 3];

  // ---- calc_hpos (main.f90:347-406): hposx/hposy (ni, nj, 4)
  std::vector<double> hx(n2 * 4, 0.0), hy(n2 * 4, 0.0);
  auto L = [&](int i) { return lon[i - 1]; };
  auto P = [&](int j) { return lat[j - 1]; };
  for (int j = 2; j <= d.njm; ++j)
    for (int i = 2; i <= d.nim; ++i) {
      double mid = (L(i) + L(i + 1)) * 0.5;
      if (L(i + 1) < L(i)) mid = (L(i) + L(i + 1) + 360) * 0.5;
      double latm = (P(j) + P(j - 1)) * 0.5;
      hx[d.at(i, j, 1)] = L(i) * TORAD;  hy[d.at(i, j, 1)] = latm * TORAD;
      hx[d.at(i, j, 2)] = mid * TORAD;   hy[d.at(i, j, 2)] = latm * TORAD;
      hx[d.at(i, j, 3)] = mid * TORAD;   hy[d.at(i, j, 3)] = P(j) * TORAD;
      hx[d.at(i, j, 4)] = L(i) * TORAD;  hy[d.at(i, j, 4)] = P(j) * TORAD;
    }
  halo(d, hx.data(), 4);
  halo(d, hy.data(), 4);
  for (int g = 1; g <= 4; ++g)
    for (int i = 1; i <= ni; ++i) {
      hx[d.at(i, 1, g)] = 2 * hx[d.at(i, 2, g)] - hx[d.at(i, 3, g)];
      hy[d.at(i, 1, g)] = 2 * hy[d.at(i, 2, g)] - hy[d.at(i, 3, g)];
    }
  for (int g = 1; g <= 4; ++g)
    for (int i = 1; i <= ni; ++i) {
      hx[d.at(i, nj, g)] = 2 * hx[d.at(i, nj - 1, g)] - hx[d.at(i, nj - 2, g)];
      hy[d.at(i, nj, g)] = 2 * hy[d.at(i, nj - 1, g)] - hy[d.at(i, nj - 2, g)];
    }

  // ---- set_hgrid (main.f90:408-497)
  for (int g = 1; g <= 4; ++g) {
    hs.rh[g].assign(n2, 0.0); hs.tn[g].assign(n2, 0.0);
    hs.dx1[g].assign(n2, 0.0); hs.dx2[g].assign(n2, 0.0);
  }
  auto wrap = [](double v) { return v < 0 ? v + 360.0 * TORAD : v; };
  for (int j = 2; j <= d.njm; ++j)
    for (int i = 2; i <= d.nim; ++i) {
      size_t p = d.at(i, j);
      hs.dx1[1][p] = wrap(hx[d.at(i, j, 2)] - hx[d.at(i - 1, j, 2)]);
      hs.dx2[1][p] = hy[d.at(i, j, 4)] - hy[d.at(i, j - 1, 4)];
      hs.dx1[2][p] = wrap(hx[d.at(i + 1, j, 1)] - hx[d.at(i, j, 1)]);
      hs.dx2[2][p] = hy[d.at(i, j, 3)] - hy[d.at(i, j - 1, 3)];
      hs.dx1[3][p] = wrap(hx[d.at(i + 1, j, 4)] - hx[d.at(i, j, 4)]);
      hs.dx2[3][p] = hy[d.at(i, j + 1, 2)] - hy[d.at(i, j, 2)];
      hs.dx1[4][p] = wrap(hx[d.at(i, j, 3)] - hx[d.at(i - 1, j, 3)]);
      hs.dx2[4][p] = hy[d.at(i, j + 1, 1)] - hy[d.at(i, j, 1)];
    }
  for (int g = 1; g <= 4; ++g) {
    halo(d, hs.dx1[g].data(), 1);
    halo(d, hs.dx2[g].data(), 1);
  }
  for (int g = 1; g <= 4; ++g)
    for (int i = 1; i <= ni; ++i) {
      auto& a = hs.dx1[g];
      auto& b = hs.dx2[g];
      a[d.at(i, 1)] = 2 * a[d.at(i, 2)] - a[d.at(i, 3)];
      b[d.at(i, 1)] = 2 * b[d.at(i, 2)] - b[d.at(i, 3)];
    }
  for (int g = 1; g <= 4; ++g)
    for (int i = 1; i <= ni; ++i) {
      auto& a = hs.dx1[g];
      auto& b = hs.dx2[g];
      a[d.at(i, nj)] = 2 * a[d.at(i, nj - 1)] - a[d.at(i, nj - 2)];
      b[d.at(i, nj)] = 2 * b[d.at(i, nj - 1)] - b[d.at(i, nj - 2)];
    }
  for (int g = 1; g <= 4; ++g)
    for (size_t p = 0; p < n2; ++p) {
      hs.rh[g][p] = std::cos(hy[p + n2 * (g - 1)]);
      hs.tn[g][p] = std::tan(hy[p + n2 * (g - 1)]);
    }

  // ---- set_vgrid (main.f90:499-583): hydrostatic fixed point for the
  // layer-interface pressures with den_rho of the mean profile.
  {
    std::vector<double> dz(nk), rmn(nk, RHO0), pmn(nk, 0.0);
    dz[0] = 2 * in.z[0];
    for (int k = 2; k <= nk; ++k) dz[k - 1] = 2 * (in.z[k - 1] - in.z[k - 2]) - dz[k - 2];
    double pmnbot = 1.0;
    int iter = 0;
    while (std::fabs(pmn[nk - 1] - pmnbot) > (double)1e-6f) {
      if (++iter > 100000) return "set_vgrid pressure iteration did not converge";
      pmnbot = pmn[nk - 1];
      rmn[0] = den_rho<double>(in.tmn[0], in.smn[0], pmn[0] * 0.5);
      for (int k = 2; k <= nk; ++k)
        rmn[k - 1] = den_rho<double>(in.tmn[k - 1], in.smn[k - 1], (pmn[k - 2] + pmn[k - 1]) * 0.5);
      pmn[0] = rmn[0] * GRAV * dz[0];
      for (int k = 2; k <= nk; ++k) pmn[k - 1] = rmn[k - 1] * GRAV * dz[k - 1] + pmn[k - 2];
    }
    hs.z1.assign(nkp, 0.0); hs.pr1.assign(nkp, 0.0); hs.dz1.assign(nkp, 0.0); hs.dpr1.assign(nkp, 0.0);
    hs.z2.assign(nk, 0.0); hs.pr2.assign(nk, 0.0); hs.dz2.assign(nk, 0.0); hs.dpr2.assign(nk, 0.0);
    for (int k = 2; k <= nkp; ++k) hs.z1[k - 1] = hs.z1[k - 2] + dz[k - 2];
    for (int k = 0; k < nk; ++k) hs.z2[k] = in.z[k];
    for (int k = 2; k <= nkp; ++k) hs.pr1[k - 1] = pmn[k - 2];
    hs.pr2[0] = pmn[0] * 0.5;
    for (int k = 2; k <= nk; ++k) hs.pr2[k - 1] = (pmn[k - 1] + pmn[k - 2]) * 0.5;
    hs.dz1[0] = 0.5 * dz[0];
    for (int k = 2; k <= nk; ++k) hs.dz1[k - 1] = in.z[k - 1] - in.z[k - 2];
    hs.dz1[nkp - 1] = hs.dz1[nk - 1];
    for (int k = 0; k < nk; ++k) hs.dz2[k] = dz[k];
    hs.dpr1[0] = 0.5 * pmn[0];
    for (int k = 2; k <= nk; ++k) hs.dpr1[k - 1] = hs.pr2[k - 1] - hs.pr2[k - 2];
    hs.dpr1[nkp - 1] = hs.dpr1[nk - 1];
    hs.dpr2[0] = pmn[0];
    for (int k = 2; k <= nk; ++k) hs.dpr2[k - 1] = pmn[k - 1] - pmn[k - 2];
  }

  // ---- set_grid (main.f90:585-659): lev, masks as (lev >= k), phih, prh
  {
    hs.lev_t.assign(n2, 0);
    for (int j = 2; j <= nj - 1; ++j)
      for (int i = 2; i <= ni - 1; ++i) {
        int v = in.itn[(size_t)(i - 2) + (size_t)in.glo_ni * (size_t)(j - 2)];
        if (v < 0 || v > nk) return "itn out of range [0, nk]";
        hs.lev_t[d.at(i, j)] = v;
      }
    halo(d, hs.lev_t.data(), 1);
    // g12%msk(i,j,k) = (lev >= k) on the interior then swapped; g32%msk =
    // product of the 4 surrounding T masks on 1..nim x 1..njm then swapped;
    // g32%lev = sum over k of g32%msk.
    std::vector<int> mt(hs.n3, 0), mu(hs.n3, 0);
    for (int k = 1; k <= nk; ++k)
      for (int j = 2; j <= d.njm; ++j)
        for (int i = 2; i <= d.nim; ++i) mt[d.at(i, j, k)] = hs.lev_t[d.at(i, j)] >= k ? 1 : 0;
    halo(d, mt.data(), nk);
    for (int k = 1; k <= nk; ++k)
      for (int j = 1; j <= d.njm; ++j)
        for (int i = 1; i <= d.nim; ++i)
          mu[d.at(i, j, k)] = mt[d.at(i, j, k)] * mt[d.at(i + 1, j, k)] * mt[d.at(i, j + 1, k)] * mt[d.at(i + 1, j + 1, k)];
    halo(d, mu.data(), nk);
    hs.lev_u.assign(n2, 0);
    for (size_t p = 0; p < n2; ++p) {
      int s = 0;
      for (int k = 0; k < nk; ++k) s += mu[p + n2 * k];
      hs.lev_u[p] = s;
    }
    // the kernels derive the masks from lev; confirm that is exact
    for (size_t p = 0; p < n2; ++p)
      for (int k = 1; k <= nk; ++k) {
        if (mt[p + n2 * (k - 1)] != (hs.lev_t[p] >= k ? 1 : 0)) return "internal: T mask is not (lev >= k)";
        if (mu[p + n2 * (k - 1)] != (hs.lev_u[p] >= k ? 1 : 0)) return "internal: U mask is not (lev >= k)";
      }
    std::vector<double> phimn(nk);
    phimn[0] = -(GRAV * hs.dz2[0]);
    for (int k = 2; k <= nk; ++k) phimn[k - 1] = phimn[k - 2] - GRAV * hs.dz2[k - 1];
    hs.phih_t.assign(n2, 0.0); hs.prh_t.assign(n2, 0.0);
    hs.phih_u.assign(n2, 0.0); hs.prh_u.assign(n2, 0.0);
    for (size_t p = 0; p < n2; ++p) {
      int k = hs.lev_t[p];
      if (k > 0) { hs.phih_t[p] = phimn[k - 1]; hs.prh_t[p] = hs.pr1[k]; }
      k = hs.lev_u[p];
      if (k > 0) { hs.phih_u[p] = phimn[k - 1]; hs.prh_u[p] = hs.pr1[k]; }
    }
    for (size_t p = 0; p < n2; ++p) {
      if (hs.lev_t[p] == 0) hs.prh_t[p] = (double)0.1f;
      if (hs.lev_u[p] == 0) hs.prh_u[p] = (double)0.1f;
    }
  }

  // ---- calc_cf (main.f90:661-673) and Coriolis (main.f90:119), on g3
  hs.cv1.assign(n2, 0.0); hs.cv2.assign(n2, 0.0); hs.cor.assign(n2, 0.0);
  for (size_t p = 0; p < n2; ++p) {
    double y3 = hy[p + n2 * 2];
    double c = std::cos(y3);
    double t = std::sqrt(1 - c * c) / std::cos(y3);
    hs.cv1[p] = (1 - t * t) / (ARAD * ARAD);
    hs.cv2[p] = 2 * t / ARAD;
    hs.cor[p] = 2.0 * OMEGA * std::sin(y3);
  }

  // ---- inistat (main.f90:675-704)
  {
    const double* src[6] = {in.taux, in.tauy, in.bct, in.bcs, in.pa, in.fw};
    for (int v = 0; v < 6; ++v) {
      hs.frc[v].assign(n2 * 12, 0.0);
      scatter(d, in.glo_ni, in.glo_nj, hs.frc[v].data(), src[v], 12);
    }
    const double missing = (double)(-9.99999e30f);
    for (size_t q = 0; q < n2 * 12; ++q)
      if (hs.frc[0][q] == missing)
        for (int v = 0; v < 6; ++v) hs.frc[v][q] = 0.0;
    hs.tc.assign(hs.n3, 0.0);
    hs.sc.assign(hs.n3, 0.0);
    scatter(d, in.glo_ni, in.glo_nj, hs.tc.data(), in.pt, nk);
    scatter(d, in.glo_ni, in.glo_nj, hs.sc.data(), in.sa, nk);
    // graphih = ter(gra(phih)): x = p1(phih, g1, g2)/(a*rh2) -> g3 (ns,
    // g2%j = 0); y = p2(phih, g1, g4)/a -> g3 (ew, g4%i = 0)
    std::vector<double> wx(n2, 0.0), wy(n2, 0.0);
    for (int j = 1; j <= nj; ++j)
      for (int i = 1; i <= d.nim; ++i) wx[d.at(i, j)] = hs.phih_t[d.at(i + 1, j)] - hs.phih_t[d.at(i, j)];
    for (size_t p = 0; p < n2; ++p) wx[p] /= hs.dx1[2][p];
    halo(d, wx.data(), 1);
    for (size_t p = 0; p < n2; ++p) wx[p] = wx[p] / (ARAD * hs.rh[2][p]);
    for (int j = 1; j <= d.njm; ++j)
      for (int i = 1; i <= ni; ++i) wy[d.at(i, j)] = hs.phih_t[d.at(i, j + 1)] - hs.phih_t[d.at(i, j)];
    for (size_t p = 0; p < n2; ++p) wy[p] /= hs.dx2[4][p];
    halo(d, wy.data(), 1);
    for (size_t p = 0; p < n2; ++p) wy[p] = wy[p] / ARAD;
    hs.graphihx.assign(n2, 0.0);
    hs.graphihy.assign(n2, 0.0);
    ter2_ns(d, hs.graphihx, wx, 0);
    ter2_ew_from_g1(d, hs.graphihy, wy, 0);
  }
  return "";
}

}  // namespace pcom
