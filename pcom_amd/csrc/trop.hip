// trop.hip - int_trop (mod_int.f90:153-296): the barotropic predictor-
// corrector subcycle of 2*dtbc/dtbt substeps, in FP64, as one cooperative
// launch of LDS-resident tiles with temporal blocking.
//
// Each workgroup owns a TW x TH tile of the (i, j) index space and keeps
// the tile plus a halo of 2S (west/south) and 3S (east/north) points of
// the prognostic free surface ch and barotropic velocity in LDS.  One
// substep needs ch within [-2, +3] and ub within [-2, +2] points of a point
// (predictor div -> sch -> viscosity / pressure gradient -> corrector div),
// so S substeps run on LDS alone, recomputing the halo redundantly; then
// tiles exchange their owned points through double-buffered global arrays
// and meet at one grid barrier.  The arithmetic per point is the same
// sequence of IEEE operations as the reference's array statements, so the
// result does not depend on the tiling.
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <string>

#include "context.h"

namespace cg = cooperative_groups;

namespace pcom {

namespace {

constexpr int kTropBlock = 1024;

// Tile coordinates: x is an interior column counted from 0 (Fortran
// i = 2 + (x mod glo_ni)), j is the Fortran row.  Periodicity is exact in
// these coordinates, so canon() is the identity.  Everything a substep reads
// lives in LDS: the horizontal metrics as 1-D vectors (rh and dx%x(2) depend
// on the row only, dx%x(1) on the column only, because the start-up chain
// builds them from the 1-D lon/lat inputs, main.f90:347-497; the host
// checks this), the 2-D statics and the prognostic fields on the region.
struct TileGrid {
  using real = double;
  int ni, njm, bx0, by0, RW, RHt;
  const int* colmap;   // [RW] Fortran column - 1 of each region column
  const double* rhj;   // [5][RHt]
  const double* dx2j;  // [5][RHt]
  const double* dx1i;  // [5][RW]
  __device__ __forceinline__ int canon(int x) const { return x; }
  __device__ __forceinline__ int cx(int x) const {
    int b = x - bx0;
    return b < 0 ? 0 : (b >= RW ? RW - 1 : b);
  }
  __device__ __forceinline__ int cy(int j) const {
    int b = j - by0;
    return b < 0 ? 0 : (b >= RHt ? RHt - 1 : b);
  }
  __device__ __forceinline__ int lid(int x, int j) const { return cx(x) + RW * cy(j); }
  __device__ __forceinline__ double RH(int g, int x, int j) const { return rhj[g * RHt + cy(j)]; }
  __device__ __forceinline__ double DX1(int g, int x, int j) const { return dx1i[g * RW + cx(x)]; }
  __device__ __forceinline__ double DX2(int g, int x, int j) const { return dx2j[g * RHt + cy(j)]; }
};

struct TropTileArgs {
  int ni, nj, G;
  const double *rh[5], *dx1[5], *dx2[5];
  const int *lev_t, *lev_u;
  const double *prh_t, *prh_u, *cor, *ghx, *ghy, *but, *bvt, *vz1, *vz2, *am1;
  const double *bphi[4], *bgra[4];
  double *chc, *chp, *mch, *mch2, *acch, *buc, *bvc, *bup, *bvp;
  double *xc[2], *xu[2], *xv[2];
  unsigned long long* flags;  // per tile, kFlagStride apart: epoch * 2^16 + round
  int* error;                 // set when a neighbour wait exceeds its bound
  unsigned long long epoch;
  int nstep, S, TW, TH, tiles_x, tiles_y;
  double bt;
};

constexpr int kFlagStride = 32;                 // 256 B between flags
constexpr unsigned long long kSpinLimit = 1ull << 22;  // ~seconds; a normal wait is microseconds

// LDS layout of one tile: region arrays of the prognostic fields and of
// every intermediate a substep reuses, owned-point accumulators, 1-D metrics
struct TileLds {
  int NR, NO, RW, RHt;
  static constexpr int kArr = 15;  // CHP PR Ua Va Ub Vb SC S1 S2 WAu WBu WAv WBv GX GY
  size_t bytes() const {
    size_t d = (size_t)kArr * NR + 4 * (size_t)NO + 11 * (size_t)RHt + 5 * (size_t)RW;
    return d * 8 + (2 * (size_t)NR + RW) * 4;
  }
};

// Publish this tile's round and wait until every tile whose owned points
// lie in this tile's region has published it.  Waits are bounded: past
// kSpinLimit polls the tile records an error and stops waiting, so the
// launch always completes (the host then reports the failure).
__device__ void tile_sync(const TropTileArgs& a, int tx, int ty, int nx_lo, int nx_hi, int ny_lo, int ny_hi,
                          unsigned long long value, int* abort_flag) {
  __syncthreads();  // this tile's exchange stores are complete
  if (threadIdx.x == 0) {
    __threadfence();
    __hip_atomic_store(&a.flags[(size_t)(tx + a.tiles_x * ty) * kFlagStride], value, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
  const int nxs = nx_hi - nx_lo + 1;
  const int nys = ny_hi - ny_lo + 1;
  for (int n = threadIdx.x; n < nxs * nys; n += blockDim.x) {
    int ox = nx_lo + n % nxs, oy = ny_lo + n / nxs;
    int y = ty + oy;
    if (y < 0 || y >= a.tiles_y) continue;
    int x = ((tx + ox) % a.tiles_x + a.tiles_x) % a.tiles_x;
    const unsigned long long* f = &a.flags[(size_t)(x + a.tiles_x * y) * kFlagStride];
    unsigned long long it = 0;
    while (__hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < value) {
      if (++it > kSpinLimit) {
        atomicOr(a.error, 1);
        *abort_flag = 1;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    __threadfence();
  }
  __syncthreads();
}

__global__ void __launch_bounds__(kTropBlock) k_trop_tiles(TropTileArgs a) {
  extern __shared__ double lds[];
  __shared__ int abort_flag;
  if (threadIdx.x == 0) abort_flag = 0;
  const int G = a.G, nj = a.nj, ni = a.ni;
  const int TW = a.TW, TH = a.TH;
  const int tx = blockIdx.x % a.tiles_x, ty = blockIdx.x / a.tiles_x;
  const int x0 = tx * TW;
  const int y0 = 1 + ty * TH;
  const int HL = 2 * a.S, HR = 3 * a.S;
  // tiles whose owned points fall in this tile's region
  const int nx_lo = -((HL + TW - 1) / TW), nx_hi = (HR + TW - 1) / TW;
  const int ny_lo = -((HL + TH - 1) / TH), ny_hi = (HR + TH - 1) / TH;
  const int nx_span = nx_hi - nx_lo + 1 > a.tiles_x ? 0 : 1;  // region wraps the whole circle
  const int wx_lo = nx_span ? nx_lo : 0, wx_hi = nx_span ? nx_hi : a.tiles_x - 1;
  const unsigned long long base = a.epoch << 16;
  const int RW = TW + HL + HR, RHt = TH + HL + HR;
  const int bx0 = x0 - HL, by0 = y0 - HL;
  const int NR = RW * RHt, NO = TW * TH;
  double* CHP = lds;
  double* PR = CHP + NR;  // predictor ch
  double* Ua = PR + NR;
  double* Va = Ua + NR;
  double* Ub = Va + NR;
  double* Vb = Ub + NR;
  double* SC = Vb + NR;   // sqrt(ter(ch)) on U points
  double* S1 = SC + NR;   // prh*sch*ub fluxes
  double* S2 = S1 + NR;
  double* WAu = S2 + NR;  // Laplacian face fluxes of u and v
  double* WBu = WAu + NR;
  double* WAv = WBu + NR;
  double* WBv = WAv + NR;
  double* GXa = WBv + NR; // gradients of the predictor ch
  double* GYa = GXa + NR;
  double* AM = GYa + NR;  // owned: mch, mch2, upb1, upb2
  double* AM2 = AM + NO;
  double* AU = AM2 + NO;
  double* AV = AU + NO;
  double* rhj = AV + NO;         // [5][RHt]
  double* dx2j = rhj + 5 * RHt;  // [5][RHt]
  double* corj = dx2j + 5 * RHt; // [RHt]
  double* dx1i = corj + RHt;     // [5][RW]
  int* levt = reinterpret_cast<int*>(dx1i + 5 * RW);
  int* levu = levt + NR;
  int* colmap = levu + NR;

  TileGrid g;
  g.ni = ni; g.njm = nj - 1; g.bx0 = bx0; g.by0 = by0; g.RW = RW; g.RHt = RHt;
  g.colmap = colmap; g.rhj = rhj; g.dx2j = dx2j; g.dx1i = dx1i;
  const double bt = a.bt;
  const int nstep = a.nstep;

  // ---- stage the 1-D metrics, masks and column map
  for (int b = threadIdx.x; b < RW; b += blockDim.x) {
    int w = (bx0 + b) % G;
    if (w < 0) w += G;
    colmap[b] = w + 1;
  }
  __syncthreads();
  for (int b = threadIdx.x; b < RW; b += blockDim.x)
    for (int s = 1; s <= 4; ++s) dx1i[s * RW + b] = a.dx1[s][colmap[b] + (long)ni * 1];  // row 2
  for (int b = threadIdx.x; b < RHt; b += blockDim.x) {
    int y = by0 + b;
    if (y < 1 || y > nj) continue;
    long p = 1 + (long)ni * (y - 1);  // column 2
    for (int s = 1; s <= 4; ++s) {
      rhj[s * RHt + b] = a.rh[s][p];
      dx2j[s * RHt + b] = a.dx2[s][p];
    }
    corj[b] = a.cor[p];
  }
  auto gidx = [&](int q) -> long {
    int bx = q % RW, by = q / RW;
    return colmap[bx] + (long)ni * (by0 + by - 1);
  };
  for (int q = threadIdx.x; q < NR; q += blockDim.x) {
    int y = by0 + q / RW;
    bool in = y >= 1 && y <= nj;
    levt[q] = in ? a.lev_t[gidx(q)] : 0;
    levu[q] = in ? a.lev_u[gidx(q)] : 0;
  }

  auto load = [&](const double* c, const double* u, const double* v) {
    for (int q = threadIdx.x; q < NR; q += blockDim.x) {
      int y = by0 + q / RW;
      if (y < 1 || y > nj) continue;
      long p = gidx(q);
      CHP[q] = c[p];
      Ua[q] = u[p];
      Va[q] = v[p];
    }
  };
  auto owned = [&](int o, int& x, int& y) -> bool {
    x = x0 + o % TW;
    y = y0 + o / TW;
    return x < G && y <= nj;
  };

  load(a.chp, a.bup, a.bvp);
  for (int o = threadIdx.x; o < NO; o += blockDim.x) {
    int x, y;
    if (!owned(o, x, y)) continue;
    long p = (long)(1 + x) + (long)ni * (y - 1);
    AU[o] = a.buc[p];
    AV[o] = a.bvc[p];
    AM[o] = a.chc[p];
    AM2[o] = a.chc[p];
  }
  __syncthreads();

  double* PU = Ua;
  double* PV = Va;
  double* QU = Ub;
  double* QV = Vb;
  int round = 0;
  auto L = [&](const double* arr) { return [=, &g](int m, int jj) -> double { return arr[g.lid(m, jj)]; }; };
  auto CH = L(CHP);
  auto PRED = L(PR);
  auto F1 = L(S1);
  auto F2 = L(S2);
  // every stage visits the region points on rows 1..nj
#define FOR_REGION(...)                                        \
  for (int q = threadIdx.x; q < NR; q += blockDim.x) {        \
    const int x = bx0 + q % RW, y = by0 + q / RW;             \
    if (y < 1 || y > nj) continue;                            \
    (void)x;                                                  \
    __VA_ARGS__                                               \
  }                                                           \
  __syncthreads();

  for (int nt = 1; nt <= nstep; ++nt) {
    auto FU = L(PU);
    auto FV = L(PV);
    // A) predictor (mod_int.f90:184-193): sch = sqrt(ter(chp)) (1 on land),
    // fluxes prh*sch*ub_p, chc = chp - gamma_b*dtbt*div/prh on sea
    FOR_REGION({
      const double sc = levu[q] > 0 ? sqrt(ter_di<1>(g, CH, x, y, 0.0)) : 1.0;
      const double pu = a.prh_u[gidx(q)];
      S1[q] = pu * sc * PU[q];
      S2[q] = pu * sc * PV[q];
    })
    FOR_REGION({
      double wk = divg<3, 3, 1>(g, F1, F2, x, y);
      wk = CHP[q] - wk * bt * GAMMA_B / a.prh_t[gidx(q)];
      // land keeps its ch: the entry value in the first substep, afterwards
      // the corrector's chp - 0 (mod_int.f90:193, 267)
      PR[q] = levt[q] > 0 ? wk : (nt == 1 ? a.chc[gidx(q)] : CHP[q]);
    })
    // B) barotropic momentum (mod_int.f90:195-261).  Face fluxes of op_lap
    // (mod_op.f90:913-920) and gradients of the predictor ch per point.
    FOR_REGION({
      SC[q] = levu[q] > 0 ? sqrt(ter_di<1>(g, PRED, x, y, 0.0)) : 1.0;
      WAu[q] = p1<3, 4>(g, FU, x, y) / g.RH(4, x, y);
      WBu[q] = p2<3, 2>(g, FU, x, y) * g.RH(2, x, y);
      WAv[q] = p1<3, 4>(g, FV, x, y) / g.RH(4, x, y);
      WBv[q] = p2<3, 2>(g, FV, x, y) * g.RH(2, x, y);
      GXa[q] = grax<1, 2>(g, PRED, x, y);
      GYa[q] = gray<1, 4>(g, PRED, x, y);
    })
    {
      auto WA1 = L(WAu);
      auto WB1 = L(WBu);
      auto WA2 = L(WAv);
      auto WB2 = L(WBv);
      FOR_REGION({
        const long p = gidx(q);
        const double su = levu[q] >= 1 ? 1.0 : 0.0;
        const double den = ARAD * ARAD * g.RH(3, x, y);
        double l1 = (p1<4, 3>(g, WA1, x, y) + p2<2, 3>(g, WB1, x, y)) / den;
        double l2 = (p1<4, 3>(g, WA2, x, y) + p2<2, 3>(g, WB2, x, y)) / den;
        double am1 = a.am1[p];
        double vs1 = su * am1 * l1;
        double vs2 = su * am1 * l2;
        const double z1 = a.vz1[p], z2 = a.vz2[p];
        double gi1 = 0.0, ci1 = 0.0;  // row nj of grachint/chgraint%x(1) is never written
        if (y <= g.njm) {
          const int qn = g.lid(x, y + 1);
          const long pn = gidx(qn);
          gi1 = 0.5 * (GXa[q] * a.bphi[0][p] + GXa[qn] * a.bphi[1][pn]);
          ci1 = 0.5 * (ter_ew<1>(g, PRED, x, y) * a.bgra[0][p] + ter_ew<1>(g, PRED, x, y + 1) * a.bgra[1][pn]);
        }
        const int qe = g.lid(x + 1, y);
        const long pe = gidx(qe);
        double gi2 = 0.5 * (GYa[q] * a.bphi[2][p] + GYa[qe] * a.bphi[3][pe]);
        double ci2 = 0.5 * (ter_ns<1>(g, PRED, x, y, 0.0) * a.bgra[2][p] +
                            ter_ns<1>(g, PRED, x + 1, y, 0.0) * a.bgra[3][pe]);
        const double sch = SC[q];
        const double cor = corj[g.cy(y)];
        double pg1 = (gi1 + ci1 + a.ghx[p]) * su * sch;
        double pg2 = (gi2 + ci2 + a.ghy[p]) * su * sch;
        const double up = PU[q], vp = PV[q];
        const double ut = a.but[p], vt = a.bvt[p];
        double pu1 = su * (vs1 - z1 + ut - pg1 + 0.5 * cor * vp);
        double pu2 = su * (vs2 - z2 + vt - pg2 - 0.5 * cor * up);
        double w1 = pu1 * bt + up;
        double w2 = pu2 * bt + vp;
        double xx = 0.5 * cor * bt;
        double uc = (w1 + 0.5 * cor * bt * w2) / (1 + xx * xx);
        double vc = (w2 - 0.5 * cor * bt * w1) / (1 + xx * xx);
        uc = 0.5 * (uc + up);
        vc = 0.5 * (vc + vp);
        pu1 = su * (vs1 - z1 - pg1 + ut + cor * vc);
        pu2 = su * (vs2 - z2 - pg2 + vt - cor * uc);
        const double un = pu1 * bt + up, vn = pu2 * bt + vp;
        QU[q] = un;
        QV[q] = vn;
        // corrector fluxes prh*sch*ub (mod_int.f90:264-265)
        const double pr = a.prh_u[p];
        S1[q] = pr * sch * un;
        S2[q] = pr * sch * vn;
      })
    }
    // C) corrector ch and the time-filter accumulators (mod_int.f90:263-273)
    FOR_REGION({
      double wk = divg<3, 3, 1>(g, F1, F2, x, y);
      double mt = levt[q] >= 1 ? 1.0 : 0.0;
      CHP[q] = CHP[q] - mt * wk * bt / a.prh_t[gidx(q)];
    })
    for (int o = threadIdx.x; o < NO; o += blockDim.x) {
      int x, y;
      if (!owned(o, x, y)) continue;
      int q = g.lid(x, y);
      double c = CHP[q];
      if (nt <= nstep / 2) AM[o] = AM[o] + c;
      AU[o] = AU[o] + QU[q];
      AV[o] = AV[o] + QV[q];
      AM2[o] = AM2[o] + c;
    }
    // the new velocity is the next substep's ub_p
    double* t;
    t = PU; PU = QU; QU = t;
    t = PV; PV = QV; QV = t;
    if (nt < nstep && nt % a.S == 0) {
      double* xc = a.xc[round & 1];
      double* xu = a.xu[round & 1];
      double* xv = a.xv[round & 1];
      for (int o = threadIdx.x; o < NO; o += blockDim.x) {
        int x, y;
        if (!owned(o, x, y)) continue;
        int q = g.lid(x, y);
        long p = (long)(1 + x) + (long)ni * (y - 1);
        xc[p] = CHP[q];
        xu[p] = PU[q];
        xv[p] = PV[q];
      }
      tile_sync(a, tx, ty, wx_lo, wx_hi, ny_lo, ny_hi, base + (unsigned long long)(round + 1), &abort_flag);
      if (abort_flag) return;
      // reload the whole region (own points included) into the A buffers
      PU = Ua; PV = Va; QU = Ub; QV = Vb;
      load(xc, xu, xv);
      ++round;
    }
    __syncthreads();
  }
#undef FOR_REGION

  // every tile of the region has read its entry fields before they are
  // overwritten below
  tile_sync(a, tx, ty, wx_lo, wx_hi, ny_lo, ny_hi, base + (unsigned long long)(round + 1), &abort_flag);
  if (abort_flag) return;
  // end of the cycle (mod_int.f90:276-287), owned points and their halo
  // copies (column 1 mirrors column ni-1, column ni mirrors column 2)
  for (int o = threadIdx.x; o < NO; o += blockDim.x) {
    int x, y;
    if (!owned(o, x, y)) continue;
    double u = AU[o] / (double)(nstep + 1);
    double v = AV[o] / (double)(nstep + 1);
    double m1 = AM[o] / (double)(nstep / 2 + 1);
    double m2 = AM2[o] / (double)(nstep + 1);
    int i = 2 + x;
    int cols[2] = {i, i == G + 1 ? 1 : (i == 2 ? G + 2 : 0)};
    for (int cidx = 0; cidx < 2; ++cidx) {
      if (cols[cidx] == 0) continue;
      long p = (long)(cols[cidx] - 1) + (long)ni * (y - 1);
      a.buc[p] = u;
      a.bvc[p] = v;
      a.bup[p] = u;
      a.bvp[p] = v;
      a.mch[p] = m1;
      a.mch2[p] = m2;
      a.chc[p] = m2;
      a.chp[p] = m2;
      a.acch[p] = a.acch[p] + m2;
    }
  }
}

// vs at the first substep (mod_int.f90:198-207): seau*am(:,:,1)*lap(ub),
// and am(:,:,1) in FP64 for the substeps.
template <typename R>
__global__ void k_vz(Grid<double> g, const R* __restrict__ am, const double* __restrict__ bup,
                     const double* __restrict__ bvp, double* vz1, double* vz2, double* am1) {
  long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.n2) return;
  int i = (int)(p % g.ni) + 1, j = (int)(p / g.ni) + 1;
  const double su = g.lev_u[p] >= 1 ? 1.0 : 0.0;
  auto PU = [&](int m, int jj) -> double { return bup[g.ix(m, jj)]; };
  auto PV = [&](int m, int jj) -> double { return bvp[g.ix(m, jj)]; };
  double a1 = (double)am[p];
  vz1[p] = su * a1 * lap<3, 3>(g, PU, i, j);
  vz2[p] = su * a1 * lap<3, 3>(g, PV, i, j);
  am1[p] = a1;
}

size_t trop_lds_bytes(int TW, int TH, int S) {
  TileLds l;
  l.RW = TW + 5 * S;
  l.RHt = TH + 5 * S;
  l.NR = l.RW * l.RHt;
  l.NO = TW * TH;
  return l.bytes();
}

}  // namespace

namespace {
// co-resident workgroups of k_trop_tiles with `lds` bytes of LDS each
std::string trop_resident(const Ctx& c, size_t lds, long* resident) {
  hipDeviceProp_t prop;
  hipError_t e = hipGetDeviceProperties(&prop, c.device);
  if (e != hipSuccess) return std::string("hipGetDeviceProperties: ") + hipGetErrorString(e);
  *resident = 0;
  if (lds > (size_t)prop.sharedMemPerBlock) return "";
  e = hipFuncSetAttribute((const void*)k_trop_tiles, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return std::string("hipFuncSetAttribute: ") + hipGetErrorString(e);
  int per_cu = 0;
  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_trop_tiles, kTropBlock, lds);
  if (e != hipSuccess) return std::string("occupancy of k_trop_tiles: ") + hipGetErrorString(e);
  *resident = (long)per_cu * prop.multiProcessorCount;
  return "";
}
}  // namespace

std::string trop_configure(Ctx& c) {
  const int G = c.ni - 2;
  // the LDS tiles keep rh, dx%x(2) and cor per row and dx%x(1) per column
  const HostSetup& hs = c.hs;
  for (int s = 1; s <= 4; ++s)
    for (int j = 1; j <= c.nj; ++j)
      for (int i = 1; i <= c.ni; ++i) {
        size_t p = (size_t)(i - 1) + (size_t)c.ni * (j - 1);
        size_t pr = 1 + (size_t)c.ni * (j - 1), pc = (size_t)(i - 1) + (size_t)c.ni;
        if (hs.rh[s][p] != hs.rh[s][pr] || hs.dx2[s][p] != hs.dx2[s][pr] || hs.dx1[s][p] != hs.dx1[s][pc] ||
            hs.cor[p] != hs.cor[pr])
          return "internal: horizontal metrics are not separable in (i, j)";
      }
  std::string err;
  if (c.trop_tw > 0 || c.trop_th > 0 || c.trop_s > 0) {
    if (c.trop_tw <= 0) c.trop_tw = 24;
    if (c.trop_th <= 0) c.trop_th = 12;
    if (c.trop_s <= 0) c.trop_s = 2;
  } else {
    // default: 2 substeps per exchange and the tile with the fewest region
    // points (owned + halo) whose tiles are all co-resident
    long best = -1;
    for (int tw = 4; tw <= 64; tw += 2)
      for (int th = 4; th <= 32; th += 2) {
        size_t lds = trop_lds_bytes(tw, th, 2);
        long resident;
        if (!(err = trop_resident(c, lds, &resident)).empty()) return err;
        long tiles = (long)((G + tw - 1) / tw) * ((c.nj + th - 1) / th);
        if (tiles > resident) continue;
        long region = (long)(tw + 10) * (th + 10);
        if (best < 0 || region < best) {
          best = region;
          c.trop_tw = tw;
          c.trop_th = th;
        }
      }
    if (best < 0) return "no barotropic tiling fits this device";
    c.trop_s = 2;
  }
  if (c.trop_s > 16) return "at most 16 barotropic substeps per exchange";
  size_t lds = trop_lds_bytes(c.trop_tw, c.trop_th, c.trop_s);
  long resident;
  if (!(err = trop_resident(c, lds, &resident)).empty()) return err;
  if (resident == 0)
    return "barotropic tile " + std::to_string(c.trop_tw) + "x" + std::to_string(c.trop_th) + "x" +
           std::to_string(c.trop_s) + " needs " + std::to_string(lds) + " B of LDS";
  long tiles_x = (G + c.trop_tw - 1) / c.trop_tw;
  long tiles_y = (c.nj + c.trop_th - 1) / c.trop_th;
  long need = tiles_x * tiles_y;
  if (need > resident)
    return "barotropic tiling needs " + std::to_string(need) + " co-resident workgroups, the device holds " +
           std::to_string(resident) + "; choose larger tiles";
  c.trop_tiles_x = (int)tiles_x;
  c.trop_blocks = (int)need;
  c.trop_lds = lds;
  return "";
}

template <typename R>
std::string launch_trop_t(Ctx& c) {
  Fields2& h = c.f2;
  const Grid<double>& gd = c.gd;
  {
    hipLaunchKernelGGL(k_vz<R>, dim3((unsigned)((c.n2 + 255) / 256)), dim3(256), 0, c.stream, gd,
                       static_cast<const R*>(c.f3.am), h.bup, h.bvp, h.vz1, h.vz2, h.am1);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return std::string("k_vz: ") + hipGetErrorString(e);
  }
  TropTileArgs a;
  a.ni = c.ni;
  a.nj = c.nj;
  a.G = c.ni - 2;
  for (int s = 0; s < 5; ++s) {
    a.rh[s] = gd.rh[s];
    a.dx1[s] = gd.dx1[s];
    a.dx2[s] = gd.dx2[s];
  }
  a.lev_t = gd.lev_t;
  a.lev_u = gd.lev_u;
  a.prh_t = gd.prh_t;
  a.prh_u = gd.prh_u;
  a.cor = gd.cor;
  a.ghx = h.graphihx;
  a.ghy = h.graphihy;
  a.but = h.but;
  a.bvt = h.bvt;
  a.vz1 = h.vz1;
  a.vz2 = h.vz2;
  a.am1 = h.am1;
  for (int q = 0; q < 4; ++q) {
    a.bphi[q] = h.bphi[q];
    a.bgra[q] = h.bgra[q];
  }
  a.chc = h.chc; a.chp = h.chp; a.mch = h.mch; a.mch2 = h.mch2; a.acch = h.acch;
  a.buc = h.buc; a.bvc = h.bvc; a.bup = h.bup; a.bvp = h.bvp;
  for (int q = 0; q < 2; ++q) {
    a.xc[q] = h.xc[q];
    a.xu[q] = h.xu[q];
    a.xv[q] = h.xv[q];
  }
  a.nstep = c.dtbc / c.dtbt * 2;
  a.S = c.trop_s;
  a.TW = c.trop_tw;
  a.TH = c.trop_th;
  a.tiles_x = c.trop_tiles_x;
  a.tiles_y = c.trop_blocks / c.trop_tiles_x;
  a.flags = c.trop_flags;
  a.error = c.trop_error;
  a.epoch = ++c.trop_epoch;
  a.bt = (double)c.dtbt;
  void* args[] = {&a};
  hipError_t e = hipLaunchCooperativeKernel((const void*)k_trop_tiles, dim3(c.trop_blocks), dim3(kTropBlock), args,
                                            (unsigned)c.trop_lds, c.stream);
  if (e != hipSuccess) return std::string("k_trop_tiles (cooperative launch): ") + hipGetErrorString(e);
  return "";
}

std::string launch_trop(Ctx& c) {
  if (c.precision == PCOM_FP64) return launch_trop_t<double>(c);
  return launch_trop_t<float>(c);
}

}  // namespace pcom
