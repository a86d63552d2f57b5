// trop.hip - int_trop (mod_int.f90:153-296): the barotropic predictor-
// corrector subcycle of 2*dtbc/dtbt substeps, in FP64, as one launch of
// co-resident LDS tiles with temporal blocking.
//
// Each workgroup owns a TW x TH tile of the (i, j) index space and keeps
// the tile plus a halo of 2S points on every side of the prognostic free
// surface ch and barotropic velocity in LDS.  One substep needs ch within
// [-2, +2] and ub within [-2, +1] points of a point (predictor fluxes ->
// predictor div -> sch, viscosity and pressure gradient -> momentum ->
// corrector div), so S substeps run on LDS alone, recomputing the halo
// redundantly, each stage only where a later stage of the round reads it;
// then tiles exchange their owned points through double-buffered global
// arrays and wait for their neighbour tiles' flags.  That needs every tile
// co-resident, which holds up to about 1 degree; larger grids run the same
// kernel in launch mode instead: one launch per S substeps over all tiles,
// each tile loading its region from the previous launch's buffer and
// storing its owned points for the next, the kernel boundary being the
// exchange.  The arithmetic per point is the same sequence of IEEE
// operations as the reference's array statements, so the result depends on
// neither the tiling nor the mode.  The mixed forms' variant (kFast)
// replaces the correctly rounded divisions and square roots by the
// reciprocal multiply and the hardware root.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "context.h"

namespace pcom {

namespace {

// one region point per thread: RW * RHt <= the workgroup size, which is
// kTropSmall for regions that fit it (four fewer idle waves per barrier and
// 168 instead of 128 VGPRs per thread) and kTropBlock otherwise
constexpr int kTropBlock = 1024;
constexpr int kTropSmall = 768;
inline int trop_block_for(long region) { return region <= kTropSmall ? kTropSmall : kTropBlock; }

inline std::string hip_error_string(hipError_t e, const char* what) {
  return std::string(what) + ": " + hipGetErrorString(e);
}

// a / b for b with rb = RN(1/b): q = RN(a*rb) is within one ulp of a/b, and
// one FMA residual step rounds it correctly (Markstein, IBM J. Res. Dev.
// 34(1), 1990), so the result is the IEEE quotient for every finite,
// non-subnormal quotient.  All divisors of a substep are static, so their
// reciprocals are computed once per launch and each division costs a
// multiply and two FMAs instead of the iterative IEEE sequence.  The mixed
// forms, which need accuracy but not the reference's bits, take a * rb
// alone (within an ulp of the quotient): a third of the substep's FP64
// operations are these corrections.
template <bool kFast>
__device__ __forceinline__ double qdiv(double a, double b, double rb) {
  if constexpr (kFast) return a * rb;
  const double q = a * rb;
  const double r = __builtin_fma(-q, b, a);
  return __builtin_fma(r, rb, q);
}

// sqrt(t) for a positive normal t above 2^-767: the hardware reciprocal
// square root refined by two Newton steps in FMA form, which is the
// compiler's own correctly rounded sequence for sqrt without its range
// scaling (inputs below 2^-767) and its special cases (0, infinities, NaN),
// so it gives the same bits for such t; 0 gives 0.  The subcycle's roots
// are of averages of ch, near 1.
__device__ __forceinline__ double sqrt_normal(double t) {
  const double r = __builtin_amdgcn_rsq(t);
  double s = t * r, h = 0.5 * r;
  const double e = __builtin_fma(-h, s, 0.5);
  s = __builtin_fma(s, e, s);
  h = __builtin_fma(h, e, h);
  double d = __builtin_fma(-s, s, t);
  s = __builtin_fma(d, h, s);
  d = __builtin_fma(-s, s, t);
  s = __builtin_fma(d, h, s);
  return t > 0.0 ? s : 0.0;
}

struct TropTileArgs {
  int ni, nj, G;
  const double *rh[5], *dx1[5], *dx2[5];
  const int *lev_t, *lev_u;
  const double *prh_t, *prh_u, *cor, *ghx, *ghy, *but, *bvt, *vz1, *vz2, *am1;
  const double *bphi[4], *bgra[4];
  double *chc, *chp, *mch, *mch2, *acch, *buc, *bvc, *bup, *bvp;
  double *xc[2], *xu[2], *xv[2];
  double* acc[4];             // accumulators between launches (launch mode)
  unsigned long long* flags;  // per tile, kFlagStride apart: epoch * 2^16 + round
  int* error;                 // set when a neighbour wait exceeds its bound
  unsigned long long spin_limit;  // unsuccessful polls that end a neighbour wait
  unsigned long long epoch;
  int nstep, S, TW, TH, tiles_x, tiles_y;
  // launch mode (k_trop_block): substeps nt0..nt1 of this launch; the state
  // enters from buffer `src` of xc/xu/xv (nt0 > 1) and leaves into buffer
  // `dst` (nt1 < nstep)
  int nt0, nt1, src, dst;
  int fast;  // divisions as multiplications by the reciprocal (mixed forms)
  double bt;
};

// The stages of a substep and the area each must cover.  With k substeps
// left in the round after this one, and T the tile side, the stages read
// (per dimension, tile points 0..T-1, east/north positive):
//   A1 predictor fluxes at x   <- ch at x, x+1; ub at x
//   A2 predictor ch at x       <- A1 at x-1, x
//   B1 Laplacian faces, ch gradients and sch at x <- ub at x-1, x; A2 at x, x+1
//   B2 momentum and corrector fluxes at x <- B1 at x, x+1; A2 at x, x+1
//   C  corrector ch at x       <- B2 at x-1, x
// so C covers [-2k, T-1+2k], B2 [-2k-1, T-1+2k], A2 and B1 [-2k-1, T+2k]
// and A1 [-2k-2, T+2k]: nested rectangles, the area of "level" L = 4k + j
// (j = 0 for C, 1 for B2, 2 for A2 and B1, 3 for A1) being
// [-ceil(L/2), T-1+floor(L/2)], T + L points wide; the region is level 4S.
// A point's level is the least L whose area holds it.
__device__ __forceinline__ int ring_level(int d, int T) {
  if (d < 0) return -2 * d - 1;
  if (d > T - 1) return 2 * (d - T + 1);
  return 0;
}
// region points (row-major, RW wide) of the level-M area before (bx, by)
__device__ __forceinline__ int ring_before(int M, int bx, int by, int H, int TW, int TH) {
  if (M < 0) return 0;
  const int xl = H - (M + 1) / 2, yl = H - (M + 1) / 2;
  const int w = TW + M, h = TH + M;
  const int rows = min(max(by - yl, 0), h);
  const int cols = (by >= yl && by < yl + h) ? min(max(bx - xl, 0), w) : 0;
  return rows * w + cols;
}
constexpr int kStageLevel[5] = {3, 2, 2, 1, 0};  // A1, A2, B1, B2, C

constexpr int kFlagStride = 32;                 // 256 B between flags
constexpr unsigned long long kSpinLimit = 1ull << 22;  // ~seconds; a normal wait is microseconds

// LDS of one tile: six region arrays of (double, double) pairs, then a
// table of per-row metrics (rh, dx%x(2) and cor depend on the row only) and
// one of per-column metrics (dx%x(1) depends on the column only): the
// start-up chain builds them from the 1-D lon/lat inputs (main.f90:347-497)
// and the host checks it.  Every array sits at a fixed offset with a fixed
// stride, so an access is one address register per neighbour plus an
// immediate offset, and the two components a stage reads at the same point
// come in one 16-byte access; components read apart are separate arrays, whose
// 8-byte lane stride keeps the LDS banks free of conflicts.
enum RegionPair {
  P_CH,  // corrector ch, then predictor ch: two scalar arrays (read apart)
  P_UV,  // barotropic velocity
  P_SF,  // prh*sch*ub fluxes
  P_GR,  // x, then y gradient of the predictor ch: two scalar arrays
  P_WA,  // Laplacian face fluxes of (u, v), east-west
  P_WB,  // the same, north-south
  P_N
};
enum RowTab {
  R_RH3, R_DX21, R_DX21I, R_ARH1, R_ARH1I, R_RH4, R_RH4I, R_DX22, R_DX22I, R_RH2, R_ARH2, R_ARH2I,
  R_DX24, R_DX24I, R_DX23, R_DX23I, R_DEN, R_DENI, R_COR, R_DXX, R_DXXI, R_N
};
enum ColTab { C_DX11, C_DX11I, C_DX14, C_DX14I, C_DX12, C_DX12I, C_DX13, C_DX13I, C_N };
// entries per metric table: at least the widest region side (a region of at
// most kTropBlock points with sides of at least 6 is at most 170 long)
constexpr int kTabStride = 176;

size_t trop_lds_bytes(int block) { return 16 * (size_t)P_N * block + 8 * (R_N + C_N) * kTabStride; }

// The exchange values and the flags are written and read only with
// agent-scope atomic accesses.  Those carry the device-scope cache policy
// (sc1) that keeps each of them coherent across the XCDs' L2 caches on its
// own, which is what makes relaxed agent-scope atomics coherent at all.
// Order between locations comes from completion: every wave waits for all
// its memory accesses (s_waitcnt with every counter at zero) before the
// workgroup barrier that precedes the flag store, so the tile's exchange
// stores, and its loads of the entry fields and of the last round, are
// complete before its flag can be seen; a consumer issues its data loads
// only after its flag poll returned.  A release store and an acquire fence
// would add an L2 writeback and an L2 invalidate per exchange, which only
// non-atomic accesses need.  The rounds' exchanges are in the substep loop
// (each thread polls the owner of its halo point); tile_sync publishes the
// end of the launch and waits until every tile whose owned points lie in
// this tile's region has published it, before a tile overwrites its owned
// points of the entry fields, which those tiles' regions read at the start.
//
// Waits are bounded: after spin_limit unsuccessful polls (kSpinLimit unless
// the host set another) the tile records an error and
// stops waiting, so the launch always completes (the host then reports the
// failure).
__device__ void tile_sync(const TropTileArgs& a, int tx, int ty, int nx_lo, int nx_hi, int ny_lo, int ny_hi,
                          unsigned long long value, int* abort_flag) {
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0)
    __hip_atomic_store(&a.flags[(size_t)(tx + a.tiles_x * ty) * kFlagStride], value, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
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
      if (++it >= a.spin_limit) {
        atomicOr(a.error, 1);
        *abort_flag = 1;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
  }
  __syncthreads();
}

__device__ __forceinline__ void xstore(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double xload(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Each workgroup owns a TW x TH tile and holds the tile plus its halo (the
// "region", RW x RHt points, one per thread) in LDS.  Tile coordinates: x
// counts interior columns from 0 (Fortran i = 2 + (x mod G)), y is the
// Fortran row; periodicity is exact in x.  Each thread keeps the static
// fields of its own point in registers for the whole launch.  Every
// statement below evaluates the reference's expression for its point with
// the reference's operation order (ter/p1/p2/div/gra inlined), so the
// result is independent of the tiling; points near the region edge read
// clamped neighbours and are garbage, but no owned point depends on them
// within S substeps.
template <bool kLaunches, int kB, bool kFast>
__device__ __forceinline__ void trop_tiles_body(const TropTileArgs& a, const int tile,
                                                const unsigned long long epoch) {
  extern __shared__ double2 lds2[];
  __shared__ int abort_flag;
  __shared__ unsigned short order[kB];  // region points by level, then row-major
  if (threadIdx.x == 0) abort_flag = 0;
  const int G = a.G, nj = a.nj, ni = a.ni, njm = a.nj - 1;
  const int TW = a.TW, TH = a.TH;
  const int tx = tile % a.tiles_x, ty = tile / a.tiles_x;
  const int x0 = tx * TW;
  const int y0 = 1 + ty * TH;
  const int H = 2 * a.S;
  // tiles whose owned points fall in this tile's region; a last column of
  // tiles narrower than a halo lets that halo reach one tile further
  const int wlast = G - (a.tiles_x - 1) * TW;
  const int xext = wlast < H ? 1 : 0;
  const int nx_lo = -((H + TW - 1) / TW) - xext, nx_hi = (H + TW - 1) / TW + xext;
  const int ny_lo = -((H + TH - 1) / TH), ny_hi = (H + TH - 1) / TH;
  const int nx_span = nx_hi - nx_lo + 1 > a.tiles_x ? 0 : 1;  // region wraps the whole circle
  const int wx_lo = nx_span ? nx_lo : 0, wx_hi = nx_span ? nx_hi : a.tiles_x - 1;
  const unsigned long long base = epoch << 16;
  const int RW = TW + 2 * H, RHt = TH + 2 * H;
  const int bx0 = x0 - H, by0 = y0 - H;
  const int NR = RW * RHt;
  // Persistent mode: threads take the region's points in order of level (the
  // stage areas described at ring_level), so a stage that covers levels up
  // to L runs on the first threads and whole wavefronts beyond them skip it.
  // Launch mode loads every region point's statics and state at each launch,
  // so its threads keep the row-major order, whose loads coalesce.
  if (threadIdx.x < NR) {
    const int rx = threadIdx.x % RW, ry = threadIdx.x / RW;
    const int L = max(ring_level(rx - H, TW), ring_level(ry - H, TH));
    const int pos = kLaunches ? (int)threadIdx.x
                              : (L > 0 ? (TW + L - 1) * (TH + L - 1) : 0) + ring_before(L, rx, ry, H, TW, TH) -
                                    ring_before(L - 1, rx, ry, H, TW, TH);
    order[pos] = (unsigned short)threadIdx.x;
  }
  __syncthreads();
  double* CHP = reinterpret_cast<double*>(lds2 + P_CH * kB);
  double* PR = CHP + kB;
  double2* UV = lds2 + P_UV * kB;
  double2* SF = lds2 + P_SF * kB;
  double* GX = reinterpret_cast<double*>(lds2 + P_GR * kB);
  double* GY = GX + kB;
  double2* WA = lds2 + P_WA * kB;
  double2* WB = lds2 + P_WB * kB;
  double* rowt = reinterpret_cast<double*>(lds2 + P_N * kB);  // [R_N][kTabStride]
  double* colt = rowt + R_N * kTabStride;                             // [C_N][kTabStride]
  const double bt = a.bt;
  const int nstep = a.nstep;

  // ---- per-row and per-column metric tables with the reciprocals of
  // every divisor (rows outside 1..nj repeat the nearest row; only
  // region-edge garbage reads them)
  for (int b = threadIdx.x; b < RHt; b += blockDim.x) {
    int y = by0 + b;
    y = y < 1 ? 1 : (y > nj ? nj : y);
    const long p = 1 + (long)ni * (y - 1);  // column 2
    double* r = rowt + b;
    const double cor = a.cor[p];
    const double xx = 0.5 * cor * bt;
    r[R_RH3 * kTabStride] = a.rh[3][p];
    r[R_DX21 * kTabStride] = a.dx2[1][p];
    r[R_ARH1 * kTabStride] = ARAD * a.rh[1][p];
    r[R_RH4 * kTabStride] = a.rh[4][p];
    r[R_DX22 * kTabStride] = a.dx2[2][p];
    r[R_RH2 * kTabStride] = a.rh[2][p];
    r[R_ARH2 * kTabStride] = ARAD * a.rh[2][p];
    r[R_DX24 * kTabStride] = a.dx2[4][p];
    r[R_DX23 * kTabStride] = a.dx2[3][p];
    r[R_DEN * kTabStride] = ARAD * ARAD * a.rh[3][p];
    r[R_COR * kTabStride] = cor;
    r[R_DXX * kTabStride] = 1 + xx * xx;
    const int div[] = {R_DX21, R_ARH1, R_RH4, R_DX22, R_ARH2, R_DX24, R_DX23, R_DEN, R_DXX};
    for (int d : div) r[(d + 1) * kTabStride] = 1.0 / r[d * kTabStride];
  }
  for (int b = threadIdx.x; b < RW; b += blockDim.x) {
    int w = (bx0 + b) % G;
    if (w < 0) w += G;
    const long p = (long)(w + 1) + (long)ni;  // row 2
    double* c = colt + b;
    c[C_DX11 * kTabStride] = a.dx1[1][p];
    c[C_DX14 * kTabStride] = a.dx1[4][p];
    c[C_DX12 * kTabStride] = a.dx1[2][p];
    c[C_DX13 * kTabStride] = a.dx1[3][p];
    const int div[] = {C_DX11, C_DX14, C_DX12, C_DX13};
    for (int d : div) c[(d + 1) * kTabStride] = 1.0 / c[d * kTabStride];
  }

  // ---- this thread's point, its level, its neighbours and its static fields
  const bool inreg = threadIdx.x < NR;
  const int pt = inreg ? (int)order[threadIdx.x] : 0;
  const int bx = inreg ? pt % RW : 0, by = inreg ? pt / RW : 0;
  const int lvl = max(ring_level(bx - H, TW), ring_level(by - H, TH));
  // LDS rows RWp apart: an odd pitch spreads the columns of a level's ring,
  // which neighbouring lanes take, over every bank (where the region still
  // fits the workgroup's arrays)
  const int RWp = (RW | 1) * RHt <= kB ? (RW | 1) : RW;
  const int q = bx + RWp * by;
  // The halo reload takes the region's points in row-major order: the level
  // order puts a column of the ring on neighbouring lanes, one memory line
  // each, where row-major lanes share lines.  This thread's row-major point:
  // its 2-D offset, its LDS slot and the flag of the tile that owns it (-1
  // when the point is owned here or lies outside the rows).
  int hp = 0, hq = 0, hf = -1;
  if (!kLaunches && threadIdx.x < NR) {
    const int hx = threadIdx.x % RW, hy = threadIdx.x / RW;
    const int gx = bx0 + hx, gy = by0 + hy;
    if (gy >= 1 && gy <= nj && !(hx >= H && hx < H + TW && hy >= H && hy < H + TH && gx < G)) {
      int hw = gx % G;
      if (hw < 0) hw += G;
      hp = (hw + 1) + ni * (gy - 1);
      hq = hx + RWp * hy;
      hf = hw / TW + a.tiles_x * ((gy - 1) / TH);
    }
  }
  const int x = bx0 + bx, y = by0 + by;
  const bool live = inreg && y >= 1 && y <= nj;
  const int bxe = bx + 1 < RW ? bx + 1 : RW - 1, bxw = bx > 0 ? bx - 1 : 0;
  const int byn = by + 1 < RHt ? by + 1 : RHt - 1, bys = by > 0 ? by - 1 : 0;
  const int qe = bxe + RWp * by, qw = bxw + RWp * by, qn = bx + RWp * byn, qs = bx + RWp * bys;
  const int qne = bxe + RWp * byn, qsw = bxw + RWp * bys;
  int w = x % G;
  if (w < 0) w += G;
  const int we = w + 1 < G ? w + 1 : 0;
  const int yc = y < 1 ? 1 : (y > nj ? nj : y);
  // 2-D offsets fit in int (ni * nj < 2^31 is checked at create)
  const int p = (w + 1) + ni * (yc - 1);
  const int pe = (we + 1) + ni * (yc - 1);
  const int pn = y < nj ? p + ni : p;
  const bool owned = live && bx >= H && bx < H + TW && by >= H && by < H + TH && x < G;
  int levt = 0, levu = 0;
  double prh_u = 0, prh_t = 1, rprh_t = 1, chc0 = 0, am1 = 0, z1 = 0, z2 = 0, ghx = 0, ghy = 0, ut = 0, vt = 0;
  double bph0 = 0, bph1 = 0, bph2 = 0, bph3 = 0, bgr0 = 0, bgr1 = 0, bgr2 = 0, bgr3 = 0;
  double acm = 0, acm2 = 0, acu = 0, acv = 0;
  if (live) {
    levt = a.lev_t[p];
    levu = a.lev_u[p];
    prh_u = a.prh_u[p];
    prh_t = a.prh_t[p];
    rprh_t = 1.0 / prh_t;
    chc0 = a.chc[p];
    am1 = a.am1[p];
    z1 = a.vz1[p];
    z2 = a.vz2[p];
    ghx = a.ghx[p];
    ghy = a.ghy[p];
    ut = a.but[p];
    vt = a.bvt[p];
    bph0 = a.bphi[0][p];
    bph2 = a.bphi[2][p];
    bgr0 = a.bgra[0][p];
    bgr2 = a.bgra[2][p];
    bph1 = a.bphi[1][pn];
    bgr1 = a.bgra[1][pn];
    bph3 = a.bphi[3][pe];
    bgr3 = a.bgra[3][pe];
    if (!kLaunches || a.nt0 == 1) {
      CHP[q] = a.chp[p];
      UV[q] = make_double2(a.bup[p], a.bvp[p]);
      if (owned) {
        acm = chc0;
        acm2 = chc0;
        acu = a.buc[p];
        acv = a.bvc[p];
      }
    } else {  // launch mode: the previous launch's state
      const bool s1 = a.src != 0;
      CHP[q] = (s1 ? a.xc[1] : a.xc[0])[p];
      UV[q] = make_double2((s1 ? a.xu[1] : a.xu[0])[p], (s1 ? a.xv[1] : a.xv[0])[p]);
      if (owned) {
        acm = a.acc[0][p];
        acm2 = a.acc[1][p];
        acu = a.acc[2][p];
        acv = a.acc[3][p];
      }
    }
  }
  const double* rw = rowt + by;
  const double* rws = rowt + bys;
  const double* cl = colt + bx;
#define ROW(n) rw[(n) * kTabStride]
#define COL(n) cl[(n) * kTabStride]
  // divg<3,3,1>(F1, F2) (div_r2d, mod_op.f90:813-848) of the fluxes in SF:
  // p1 and p2 with the pre-interpolations ter_ns(g3->g2) and ter_ew(g3->g4),
  // / (a*rh_g1)
  auto div331 = [&]() -> double {
    const bool r2 = y >= 2 && y <= nj;
    const double2 fs = SF[qs], fq = SF[q], fsw = SF[qsw], fw = SF[qw];
    const double t1 = r2 ? 0.5 * (fs.x + fq.x) : 0.0;
    const double t0 = r2 ? 0.5 * (fsw.x + fw.x) : 0.0;
    const double a1 = qdiv<kFast>(t1 - t0, COL(C_DX11), COL(C_DX11I));
    double d = 0.0;
    if (r2) {
      const double rh3 = ROW(R_RH3), rh3s = rws[R_RH3 * kTabStride];
      const double u1 = 0.5 * (fw.y * rh3 + fq.y * rh3);
      const double u0 = 0.5 * (fsw.y * rh3s + fs.y * rh3s);
      d = u1 - u0;
    }
    const double b1 = qdiv<kFast>(d, ROW(R_DX21), ROW(R_DX21I));
    return qdiv<kFast>(a1 + b1, ROW(R_ARH1), ROW(R_ARH1I));
  };
  // sqrt(ter(v, g1 -> g3)) on sea U points, 1 on land, from v at the point
  // and its north, east and north-east neighbours
  const bool rin = y >= 1 && y <= njm;
  auto sch_of = [&](double v, double vn, double ve, double vne) -> double {
    if (levu <= 0) return 1.0;
    const double t = rin ? (v + vn + ve + vne) * 0.25 : 0.0;
    // the mixed forms take the hardware square root: one instruction where
    // the correctly rounded one is a refinement sequence, with a relative
    // error of up to about 2e-8 (FP32's accuracy, like their stencils)
    if constexpr (kFast) return __builtin_amdgcn_sqrt(t);
    return sqrt_normal(t);
  };
  const double su = levu >= 1 ? 1.0 : 0.0;
  const double sam1 = su * am1;  // su * am1 * lap, left to right
  const double mt = levt >= 1 ? 1.0 : 0.0;
  __syncthreads();

  int round = 0;
  const int nt1 = kLaunches ? a.nt1 : nstep;
  for (int nt = kLaunches ? a.nt0 : 1; nt <= nt1; ++nt) {
    // the level of the stages of a substep with k substeps left before the
    // next exchange (or the end of the launch): 4k + kStageLevel[stage]
    const int rend = kLaunches ? nt1 : min((nt + a.S - 1) / a.S * a.S, nt1);
    const int k4 = 4 * (rend - nt);
    // A) predictor (mod_int.f90:184-193): sch = sqrt(ter(chp)) (1 on land),
    // fluxes prh*sch*ub_p, chc = chp - gamma_b*dtbt*div/prh on sea
    if (live && lvl <= k4 + kStageLevel[0]) {
      const double sc = sch_of(CHP[q], CHP[qn], CHP[qe], CHP[qne]);
      const double2 uv = UV[q];
      SF[q] = make_double2(prh_u * sc * uv.x, prh_u * sc * uv.y);
    }
    __syncthreads();
    if (live && lvl <= k4 + kStageLevel[1]) {
      double wk = div331();
      const double chp = CHP[q];
      wk = chp - qdiv<kFast>(wk * bt * GAMMA_B, prh_t, rprh_t);
      // land keeps its ch: the entry value in the first substep, afterwards
      // the corrector's chp - 0 (mod_int.f90:193, 267)
      PR[q] = levt > 0 ? wk : (nt == 1 ? a.chc[p] : chp);
    }
    __syncthreads();
    // B) barotropic momentum (mod_int.f90:195-261).  Face fluxes of op_lap
    // (mod_op.f90:913-920) and gradients of the predictor ch per point.
    double sch = 1.0;
    double2 uv = make_double2(0.0, 0.0);
    if (live && lvl <= k4 + kStageLevel[2]) {
      const double pr = PR[q], pre = PR[qe], prn = PR[qn];
      sch = sch_of(pr, prn, pre, PR[qne]);
      const bool r2 = y >= 2 && y <= nj;
      uv = UV[q];
      const double2 uvw = UV[qw], uvs = UV[qs];
      WA[q] = make_double2(qdiv<kFast>(qdiv<kFast>(uv.x - uvw.x, COL(C_DX14), COL(C_DX14I)), ROW(R_RH4), ROW(R_RH4I)),
                           qdiv<kFast>(qdiv<kFast>(uv.y - uvw.y, COL(C_DX14), COL(C_DX14I)), ROW(R_RH4), ROW(R_RH4I)));
      WB[q] = make_double2(qdiv<kFast>(r2 ? uv.x - uvs.x : 0.0, ROW(R_DX22), ROW(R_DX22I)) * ROW(R_RH2),
                           qdiv<kFast>(r2 ? uv.y - uvs.y : 0.0, ROW(R_DX22), ROW(R_DX22I)) * ROW(R_RH2));
      GX[q] = qdiv<kFast>(qdiv<kFast>(pre - pr, COL(C_DX12), COL(C_DX12I)), ROW(R_ARH2), ROW(R_ARH2I));
      GY[q] = qdiv<kFast>(qdiv<kFast>(rin ? prn - pr : 0.0, ROW(R_DX24), ROW(R_DX24I)), ARAD, 1.0 / ARAD);
    }
    __syncthreads();
    double un = 0, vn = 0;
    if (live && lvl <= k4 + kStageLevel[3]) {
      const double2 wa = WA[q], wae = WA[qe], wb = WB[q], wbn = WB[qn];
      const double l1 = qdiv<kFast>(qdiv<kFast>(wae.x - wa.x, COL(C_DX13), COL(C_DX13I)) +
                                 qdiv<kFast>(rin ? wbn.x - wb.x : 0.0, ROW(R_DX23), ROW(R_DX23I)),
                             ROW(R_DEN), ROW(R_DENI));
      const double l2 = qdiv<kFast>(qdiv<kFast>(wae.y - wa.y, COL(C_DX13), COL(C_DX13I)) +
                                 qdiv<kFast>(rin ? wbn.y - wb.y : 0.0, ROW(R_DX23), ROW(R_DX23I)),
                             ROW(R_DEN), ROW(R_DENI));
      const double vs1 = sam1 * l1;
      const double vs2 = sam1 * l2;
      const double gx = GX[q], gxn = GX[qn], gy = GY[q], gye = GY[qe];
      const double pr = PR[q], pre = PR[qe], prn = PR[qn], prne = PR[qne];
      double gi1 = 0.0, ci1 = 0.0;  // row nj of grachint/chgraint%x(1) is never written
      if (y <= njm) {
        gi1 = 0.5 * (gx * bph0 + gxn * bph1);
        ci1 = 0.5 * (0.5 * (pr + pre) * bgr0 + 0.5 * (prn + prne) * bgr1);
      }
      const double gi2 = 0.5 * (gy * bph2 + gye * bph3);
      const double ci2 = 0.5 * ((rin ? 0.5 * (pr + prn) : 0.0) * bgr2 + (rin ? 0.5 * (pre + prne) : 0.0) * bgr3);
      const double cor = ROW(R_COR);
      const double pg1 = (gi1 + ci1 + ghx) * su * sch;
      const double pg2 = (gi2 + ci2 + ghy) * su * sch;
      const double up = uv.x, vp = uv.y;
      double pu1 = su * (vs1 - z1 + ut - pg1 + 0.5 * cor * vp);
      double pu2 = su * (vs2 - z2 + vt - pg2 - 0.5 * cor * up);
      const double w1 = pu1 * bt + up;
      const double w2 = pu2 * bt + vp;
      double uc = qdiv<kFast>(w1 + 0.5 * cor * bt * w2, ROW(R_DXX), ROW(R_DXXI));
      double vc = qdiv<kFast>(w2 - 0.5 * cor * bt * w1, ROW(R_DXX), ROW(R_DXXI));
      uc = 0.5 * (uc + up);
      vc = 0.5 * (vc + vp);
      pu1 = su * (vs1 - z1 - pg1 + ut + cor * vc);
      pu2 = su * (vs2 - z2 - pg2 + vt - cor * uc);
      un = pu1 * bt + up;
      vn = pu2 * bt + vp;
      // the new velocity is the next substep's ub_p; no thread reads UV at
      // another point in this stage
      UV[q] = make_double2(un, vn);
      // corrector fluxes prh*sch*ub (mod_int.f90:264-265)
      SF[q] = make_double2(prh_u * sch * un, prh_u * sch * vn);
    }
    // an exchange follows this substep: the owned velocity goes out now, so
    // that its stores complete while the corrector runs
    const bool exch = !kLaunches && nt < nt1 && nt % a.S == 0;
    if (exch && owned) {
      const bool odd = round & 1;
      xstore((odd ? a.xu[1] : a.xu[0]) + p, un);
      xstore((odd ? a.xv[1] : a.xv[0]) + p, vn);
    }
    __syncthreads();
    // C) corrector ch and the time-filter accumulators (mod_int.f90:263-273)
    double cn = 0;
    if (live && lvl <= k4 + kStageLevel[4]) {
      const double wk = div331();
      cn = CHP[q] - qdiv<kFast>(mt * wk * bt, prh_t, rprh_t);
      CHP[q] = cn;
      if (owned) {
        if (nt <= nstep / 2) acm = acm + cn;
        acu = acu + un;
        acv = acv + vn;
        acm2 = acm2 + cn;
      }
    }
    if (exch) {
      // uniform selects keep the buffer addresses in scalar registers
      const bool odd = round & 1;
      double* const xc = odd ? a.xc[1] : a.xc[0];
      double* const xu = odd ? a.xu[1] : a.xu[0];
      double* const xv = odd ? a.xv[1] : a.xv[0];
      if (owned) xstore(xc + p, cn);
      // publish this round: every wave's accesses complete, then the flag
      const unsigned long long value = base + (unsigned long long)(round + 1);
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
      if (threadIdx.x == 0)
        __hip_atomic_store(&a.flags[(size_t)tile * kFlagStride], value, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      // the halo from the neighbours (owned points already hold these
      // values): each thread waits for the flag of the tile that owns its
      // row-major point, then issues the point's three loads before using
      // either value, so one round trip serves them.  Every tile that reads
      // this tile's points owns points of this tile's halo (the region
      // relation is symmetric), so a tile stores round r + 2 only after every
      // reader of its round r has published round r + 1, which it does once
      // its loads of round r are complete.
      if (hf >= 0) {
        const unsigned long long* f = &a.flags[(size_t)hf * kFlagStride];
        unsigned long long it = 0;
        while (__hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < value) {
          if (++it >= a.spin_limit) {
            atomicOr(a.error, 1);
            abort_flag = 1;
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
        const double hc = xload(xc + hp), hu = xload(xu + hp), hv = xload(xv + hp);
        CHP[hq] = hc;
        UV[hq] = make_double2(hu, hv);
      }
    }
    __syncthreads();
    if (exch) {
      if (__builtin_amdgcn_readfirstlane(abort_flag)) return;
      ++round;
    }
  }
#undef ROW
#undef COL

  if (kLaunches && nt1 < nstep) {  // hand the owned points to the next launch
    if (owned) {
      const bool d1 = a.dst != 0;
      const double2 uv = UV[q];
      (d1 ? a.xc[1] : a.xc[0])[p] = CHP[q];
      (d1 ? a.xu[1] : a.xu[0])[p] = uv.x;
      (d1 ? a.xv[1] : a.xv[0])[p] = uv.y;
      a.acc[0][p] = acm;
      a.acc[1][p] = acm2;
      a.acc[2][p] = acu;
      a.acc[3][p] = acv;
    }
    return;
  }
  // every tile of the region has read its entry fields before they are
  // overwritten below (in launch mode the entry fields were read by an
  // earlier launch)
  if (!kLaunches)
    tile_sync(a, tx, ty, wx_lo, wx_hi, ny_lo, ny_hi, base + (unsigned long long)(round + 1), &abort_flag);
  if (__builtin_amdgcn_readfirstlane(abort_flag) || !owned) return;
  // end of the cycle (mod_int.f90:276-287), owned points and their halo
  // copies (column 1 mirrors column ni-1, column ni mirrors column 2)
  const double u = acu / (double)(nstep + 1);
  const double v = acv / (double)(nstep + 1);
  const double m1 = acm / (double)(nstep / 2 + 1);
  const double m2 = acm2 / (double)(nstep + 1);
  const int i = 2 + x;
  const int cols[2] = {i, i == G + 1 ? 1 : (i == 2 ? G + 2 : 0)};
  for (int cidx = 0; cidx < 2; ++cidx) {
    if (cols[cidx] == 0) continue;
    const long o = (long)(cols[cidx] - 1) + (long)ni * (y - 1);
    a.buc[o] = u;
    a.bvc[o] = v;
    a.bup[o] = u;
    a.bvp[o] = v;
    a.mch[o] = m1;
    a.mch2[o] = m2;
    a.chc[o] = m2;
    a.chp[o] = m2;
    a.acch[o] = a.acch[o] + m2;
  }
}

template <int kB, bool kFast>
__global__ void __launch_bounds__(kB) k_trop_tiles(TropTileArgs a) {
  trop_tiles_body<false, kB, kFast>(a, blockIdx.x, a.epoch);
}

// launch mode: tiles of a launch do not wait for each other, so neighbouring
// tiles go to one XCD, whose L2 then serves their shared halo
template <int kB, bool kFast>
__global__ void __launch_bounds__(kB) k_trop_block(TropTileArgs a) {
  trop_tiles_body<true, kB, kFast>(a, (int)xcd_block(), 0);
}

// The subcycles of several contexts (ensemble members) in one launch:
// workgroups [start[m], start[m+1]) are member m's tiles, each member with
// its own fields, flags and epoch.  Tiles wait only for tiles of their own
// member, so members of any grid sizes and forms share the launch.
constexpr int kMaxPack = 16;
struct MultiLaunch {
  int n;
  int start[kMaxPack + 1];
  unsigned long long epoch[kMaxPack];
};

template <int kB>
__global__ void __launch_bounds__(kB) k_trop_multi(const TropTileArgs* __restrict__ args, MultiLaunch ml) {
  const int b = blockIdx.x;
  int m = 0;
  while (m + 1 < ml.n && b >= ml.start[m + 1]) ++m;
  if (args[m].fast)
    trop_tiles_body<false, kB, true>(args[m], b - ml.start[m], ml.epoch[m]);
  else
    trop_tiles_body<false, kB, false>(args[m], b - ml.start[m], ml.epoch[m]);
}

// the kernels of one workgroup size and division kind
struct TropKernels {
  const void* tiles;
  const void* block;
  const void* multi;
};
template <int kB, bool kFast>
TropKernels trop_kernels_of() {
  return {(const void*)k_trop_tiles<kB, kFast>, (const void*)k_trop_block<kB, kFast>, (const void*)k_trop_multi<kB>};
}
TropKernels trop_kernels(int block, bool fast) {
  if (block == kTropSmall) return fast ? trop_kernels_of<kTropSmall, true>() : trop_kernels_of<kTropSmall, false>();
  return fast ? trop_kernels_of<kTropBlock, true>() : trop_kernels_of<kTropBlock, false>();
}

// vs at the first substep (mod_int.f90:198-207): seau*am(:,:,1)*lap(ub),
// and am(:,:,1) in FP64 for the substeps.
template <typename R>
__global__ void k_vz(Grid<double> g, const R* __restrict__ am, const double* __restrict__ bup,
                     const double* __restrict__ bvp, double* vz1, double* vz2, double* am1) {
  long p = (long)xcd_block() * blockDim.x + threadIdx.x;
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


}  // namespace

namespace {
// co-resident workgroups of k_trop_tiles (and of k_trop_block, whichever
// is fewer) of `block` threads with their LDS
std::string trop_resident(const Ctx& c, int block, long* resident) {
  const size_t lds = trop_lds_bytes(block);
  hipDeviceProp_t prop;
  hipError_t e = hipGetDeviceProperties(&prop, c.device);
  if (e != hipSuccess) return std::string("hipGetDeviceProperties: ") + hipGetErrorString(e);
  *resident = 0;
  if (lds > (size_t)prop.sharedMemPerBlock) return "";
  int per_cu = 1 << 30;
  const TropKernels te = trop_kernels(block, false), tf = trop_kernels(block, true);
  for (const void* k : {te.tiles, te.block, tf.tiles, tf.block}) {
    e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return std::string("hipFuncSetAttribute: ") + hipGetErrorString(e);
    int n = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, block, lds);
    if (e != hipSuccess) return std::string("occupancy of the subcycle kernels: ") + hipGetErrorString(e);
    per_cu = std::min(per_cu, n);
  }
  *resident = (long)per_cu * prop.multiProcessorCount;
  return "";
}
}  // namespace

namespace {
// A candidate tiling and its modelled cost per cycle, in hundredths of one
// wavefront of points through one substep on each of a CU's four SIMDs, in
// integers so that equal costs compare equal.  Persistent mode: a substep's
// stages cover nested areas (ring_level), each costing the wavefronts per
// SIMD of its area weighted by its share (percent) of the substep's
// instructions, the substeps of the 1024-thread kernel (which spills
// registers) at kBigBlockPercent percent; the rounds of S substeps plus
// kExchangeCost between rounds.  Both constants are a least-squares fit to
// 17 tilings timed at 1 degree (an exchange is about 2.1 us).  Launch mode
// (row-major threads): each launch runs its tiles in waves of co-resident
// workgroups, each paying its substeps on the whole region plus about
// kLoadSubsteps substeps for loading its region and statics, and a launch
// adds kLaunchCost.  Equal costs go to the lower tile, then the wider: the
// halo exchange moves region rows, which are contiguous in memory, so fewer
// and longer rows cost less (24x11 < 16x17 < 20x13 < 19x14 < 10x26 at 1
// degree, all of one modelled cost).
struct TropCand {
  long cost;
  int tw, th, s;
};
constexpr long kExchangeCost = 670, kBigBlockPercent = 107, kLoadSubsteps = 2, kLaunchCost = 1400;
constexpr long kStagePercent[5] = {10, 10, 22, 48, 10};  // A1, A2, B1, B2, C

inline long simd_waves(long points) { return 4 * (((points + 63) / 64 + 3) / 4); }

// substeps 1..s of a round (or launch) on a tw x th tile
long round_cost(int tw, int th, int s) {
  long cost = 0;
  for (int k = 0; k < s; ++k)
    for (int st = 0; st < 5; ++st) {
      const int L = 4 * k + kStageLevel[st];
      cost += kStagePercent[st] * simd_waves((long)(tw + L) * (th + L));
    }
  return cost;
}

std::string trop_candidates(const Ctx& c, bool launches, std::vector<TropCand>* out) {
  const int G = c.ni - 2;
  const long nstep = (long)c.dtbc / c.dtbt * 2;
  long res_small, res_big;
  std::string err = trop_resident(c, kTropSmall, &res_small);
  if (err.empty()) err = trop_resident(c, kTropBlock, &res_big);
  if (!err.empty()) return err;
  out->clear();
  for (int s = 1; s <= 4; ++s) {
    if (launches && s >= nstep) continue;  // a cycle takes at least two launches
    for (int tw = 4; tw <= 64; ++tw)
      for (int th = 4; th <= 64; ++th) {
        const long nr = (long)(tw + 4 * s) * (th + 4 * s);
        if (nr > kTropBlock) continue;
        const long resident = trop_block_for(nr) == kTropSmall ? res_small : res_big;
        if (resident == 0) continue;
        const long tiles = (long)((G + tw - 1) / tw) * ((c.nj + th - 1) / th);
        const long nl = (nstep + s - 1) / s;  // rounds (launches) per cycle
        const long rounds_cost = (nstep / s) * round_cost(tw, th, s) + round_cost(tw, th, (int)(nstep % s));
        long cost;
        if (!launches) {
          if (tiles > resident || (c.trop_flag_cap > 0 && tiles > c.trop_flag_cap)) continue;
          cost = rounds_cost * (trop_block_for(nr) == kTropBlock ? kBigBlockPercent : 100) / 100 +
                 (nl - 1) * kExchangeCost;
        } else {  // row-major threads: the whole region's wavefronts every substep
          const long rounds = (tiles + resident - 1) / resident;
          cost = rounds * (nstep + nl * kLoadSubsteps) * simd_waves(nr) * 100 + nl * kLaunchCost;
        }
        out->push_back({cost, tw, th, s});
      }
  }
  std::sort(out->begin(), out->end(), [](const TropCand& a, const TropCand& b) {
    if (a.cost != b.cost) return a.cost < b.cost;
    if (a.th != b.th) return a.th < b.th;
    if (a.tw != b.tw) return a.tw > b.tw;
    return a.s < b.s;
  });
  return "";
}
}  // namespace

std::string trop_configure(Ctx& c) {
  const int G = c.ni - 2;
  const long nstep = (long)c.dtbc / c.dtbt * 2;
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
  bool launches = c.trop_mode == PCOM_TROP_LAUNCHES;
  if (c.trop_tw > 0 || c.trop_th > 0 || c.trop_s > 0) {
    if (c.trop_tw <= 0) c.trop_tw = 24;
    if (c.trop_th <= 0) c.trop_th = 12;
    if (c.trop_s <= 0) c.trop_s = 2;
  } else {
    // default: the candidate of least modelled cost, persistent if any fits
    std::vector<TropCand> cands;
    if (!launches) {
      if (!(err = trop_candidates(c, false, &cands)).empty()) return err;
      if (cands.empty() && c.trop_mode == PCOM_TROP_PERSISTENT)
        return "no persistent barotropic tiling fits this device for this grid";
      launches = cands.empty();
    }
    if (launches && !(err = trop_candidates(c, true, &cands)).empty()) return err;
    if (cands.empty()) return "no barotropic tiling fits this device";
    c.trop_tw = cands[0].tw;
    c.trop_th = cands[0].th;
    c.trop_s = cands[0].s;
  }
  if (c.trop_s > 16) return "at most 16 barotropic substeps per exchange";
  if ((long)(c.trop_tw + 4 * c.trop_s) * (c.trop_th + 4 * c.trop_s) > kTropBlock)
    return "barotropic tile " + std::to_string(c.trop_tw) + "x" + std::to_string(c.trop_th) + "x" +
           std::to_string(c.trop_s) + ": the tile plus its halo of 2 points per substep on each side must hold at most " +
           std::to_string(kTropBlock) + " points";
  if (c.trop_tw + 4 * c.trop_s > kTabStride || c.trop_th + 4 * c.trop_s > kTabStride)
    return "barotropic tile sides plus halo must be at most " + std::to_string(kTabStride);
  const int block = trop_block_for((long)(c.trop_tw + 4 * c.trop_s) * (c.trop_th + 4 * c.trop_s));
  const size_t lds = trop_lds_bytes(block);
  long resident;
  if (!(err = trop_resident(c, block, &resident)).empty()) return err;
  if (resident == 0) return "the subcycle's " + std::to_string(lds) + " B of LDS do not fit this device";
  long tiles_x = (G + c.trop_tw - 1) / c.trop_tw;
  long tiles_y = (c.nj + c.trop_th - 1) / c.trop_th;
  long need = tiles_x * tiles_y;
  const bool fits = need <= resident && !(c.trop_flag_cap > 0 && need > c.trop_flag_cap);
  if (c.trop_mode == PCOM_TROP_AUTO) launches = !fits;
  if (!launches && !fits)
    return "barotropic tiling needs " + std::to_string(need) + " co-resident workgroups with exchange flags, " +
           "the device holds " + std::to_string(std::min(resident, c.trop_flag_cap > 0 ? c.trop_flag_cap : resident)) +
           "; choose larger tiles or PCOM_TROP_LAUNCHES";
  if (launches && c.trop_s >= nstep)
    return "launch mode needs fewer substeps per launch than the " + std::to_string(nstep) + " of a cycle";
  if (need > (1l << 30)) return "too many barotropic tiles";
  c.trop_launches = launches;
  c.trop_tiles_x = (int)tiles_x;
  c.trop_blocks = (int)need;
  c.trop_block = block;
  c.trop_lds = lds;
  return "";
}

// process-wide order of subcycle launches per device (never destroyed: the
// events outlive every context)
struct TropChain {
  hipEvent_t ev = nullptr;
  hipStream_t last = nullptr;  // the stream of the latest launch
};
std::mutex g_trop_mutex;
std::map<int, TropChain> g_trop_chain;

template <typename R>
std::string launch_vz_t(Ctx& c) {
  Fields2& h = c.f2;
  hipLaunchKernelGGL(k_vz<R>, dim3((unsigned)((c.n2 + 255) / 256)), dim3(256), 0, c.stream, c.gd,
                     static_cast<const R*>(c.f3.am), h.bup, h.bvp, h.vz1, h.vz2, h.am1);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return std::string("k_vz: ") + hipGetErrorString(e);
  return "";
}

std::string launch_vz(Ctx& c) {
  if (c.precision == PCOM_FP64) return launch_vz_t<double>(c);
  return launch_vz_t<float>(c);
}

// the subcycle's arguments for the context's current tiling (epoch 0)
TropTileArgs trop_args(const Ctx& c) {
  const Fields2& h = c.f2;
  const Grid<double>& gd = c.gd;
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
  for (int q = 0; q < 4; ++q) a.acc[q] = h.xacc[q];
  a.nstep = c.dtbc / c.dtbt * 2;
  a.nt0 = 1;
  a.nt1 = a.nstep;
  a.src = 0;
  a.dst = 0;
  a.S = c.trop_s;
  a.TW = c.trop_tw;
  a.TH = c.trop_th;
  a.tiles_x = c.trop_tiles_x;
  a.tiles_y = c.trop_blocks / c.trop_tiles_x;
  a.flags = c.trop_flags;
  a.error = c.trop_error;
  a.spin_limit = c.trop_spin_limit ? c.trop_spin_limit : kSpinLimit;
  a.fast = c.precision != PCOM_FP64;
  a.epoch = 0;
  a.bt = (double)c.dtbt;
  return a;
}

// An ordinary launch: trop_configure keeps the grid within the device's
// co-resident capacity (one workgroup per CU by its LDS), so every tile a
// wait depends on gets a CU once other work on the device drains, and every
// wait is bounded.  Two such launches running at once could each hold CUs
// the other's tiles need, so launches of all contexts of this process on a
// device are chained: each waits for the previous one.  (A cooperative
// launch would add nothing but its per-process global-wave-sync resource,
// which fails when another process on the device holds it.)  Launches from
// one stream are ordered by the stream; only a launch from another stream
// than the latest one records that stream's progress (all it has queued,
// which ends with or after the latest launch) and waits for it, so a single
// context pays no event per step.
template <typename F>
std::string chained_launch(int device, hipStream_t st, F launch) {
  std::lock_guard<std::mutex> lock(g_trop_mutex);
  TropChain& ch = g_trop_chain[device];
  if (ch.last && ch.last != st) {
    if (!ch.ev) {
      hipError_t ee = hipEventCreateWithFlags(&ch.ev, hipEventDisableTiming);
      if (ee != hipSuccess) return std::string("hipEventCreate: ") + hipGetErrorString(ee);
    }
    hipError_t e = hipEventRecord(ch.ev, ch.last);
    if (e != hipSuccess) return std::string("hipEventRecord: ") + hipGetErrorString(e);
    if ((e = hipStreamWaitEvent(st, ch.ev, 0)) != hipSuccess)
      return std::string("hipStreamWaitEvent: ") + hipGetErrorString(e);
  }
  std::string err = launch();
  if (!err.empty()) return err;
  ch.last = st;
  return "";
}

// a stream about to be destroyed leaves the chain
void trop_chain_forget(int device, hipStream_t st) {
  std::lock_guard<std::mutex> lock(g_trop_mutex);
  auto it = g_trop_chain.find(device);
  if (it != g_trop_chain.end() && it->second.last == st) it->second.last = nullptr;
}

std::string launch_trop_kernel(Ctx& c) {
  TropTileArgs a = trop_args(c);
  if (c.trop_launches) {
    // one launch per S substeps; tiles of a launch never wait for each
    // other, so the launches need no chaining with other contexts'
    if (c.timing) (void)hipEventRecord(c.ev[1], c.stream);
    for (int b = 0, nt0 = 1; nt0 <= a.nstep; ++b, nt0 += c.trop_s) {
      a.nt0 = nt0;
      a.nt1 = std::min(a.nstep, nt0 + c.trop_s - 1);
      a.src = (b + 1) & 1;
      a.dst = b & 1;
      const void* k = trop_kernels(c.trop_block, a.fast).block;
      void* kargs[] = {&a};
      hipError_t e = hipLaunchKernel(k, dim3(c.trop_blocks), dim3(c.trop_block), kargs, c.trop_lds, c.stream);
      if (e == hipSuccess) e = hipGetLastError();
      if (e != hipSuccess) return std::string("k_trop_block: ") + hipGetErrorString(e);
    }
    return "";
  }
  a.epoch = ++c.trop_epoch;
  return chained_launch(c.device, c.stream, [&]() -> std::string {
    if (c.timing) (void)hipEventRecord(c.ev[1], c.stream);
    const void* k = trop_kernels(c.trop_block, a.fast).tiles;
    void* kargs[] = {&a};
    hipError_t e = hipLaunchKernel(k, dim3(c.trop_blocks), dim3(c.trop_block), kargs, c.trop_lds, c.stream);
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) return std::string("k_trop_tiles: ") + hipGetErrorString(e);
    return "";
  });
}

std::string launch_trop(Ctx& c) {
  std::string e = launch_vz(c);
  if (!e.empty()) return e;
  return launch_trop_kernel(c);
}

}  // namespace pcom

// ---------------------------------------------------------------------------
// Ensembles: several contexts stepped together.  Each member's 3-D kernels
// run on the member's own stream; the members' subcycles run in packed
// k_trop_multi launches on the ensemble's stream, first-fit by tile count
// against the device's co-resident capacity, so small members fill the
// device together instead of running one after another.

namespace pcom {

struct Ensemble {
  std::vector<Ctx*> m;
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<std::vector<int>> packs;  // member indices
  std::vector<int> pack_block;          // workgroup size of each pack's launch
  std::vector<int> pack_of;             // member -> pack (-1: runs alone)
  TropTileArgs* dargs = nullptr;        // device, in pack order
  std::vector<int> pack_first;          // first dargs index of each pack
  std::vector<hipEvent_t> ready, done;
};

namespace {

std::string multi_capacity(int device, int block, long* cap) {
  hipDeviceProp_t prop;
  hipError_t e = hipGetDeviceProperties(&prop, device);
  if (e != hipSuccess) return std::string("hipGetDeviceProperties: ") + hipGetErrorString(e);
  const size_t lds = trop_lds_bytes(block);
  const void* k = trop_kernels(block, false).multi;
  e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return std::string("hipFuncSetAttribute: ") + hipGetErrorString(e);
  int per_cu = 0;
  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, block, lds);
  if (e != hipSuccess) return std::string("occupancy of k_trop_multi: ") + hipGetErrorString(e);
  *cap = (long)per_cu * prop.multiProcessorCount;
  return "";
}

// packs and device arguments for the members' current tilings
std::string ensemble_plan(Ensemble& en) {
  const int n = (int)en.m.size();
  std::vector<int> order;
  for (int q = 0; q < n; ++q)
    if (!en.m[q]->nn && !en.m[q]->trop_launches) order.push_back(q);
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return en.m[a]->trop_blocks > en.m[b]->trop_blocks; });
  en.packs.clear();
  en.pack_block.clear();
  en.pack_of.assign(n, -1);
  std::vector<long> used;
  for (int q : order) {
    const Ctx& c = *en.m[q];
    bool placed = false;
    for (size_t p = 0; p < en.packs.size() && !placed; ++p) {
      if (en.packs[p].size() >= (size_t)kMaxPack) continue;
      const int block = std::max(en.pack_block[p], c.trop_block);
      long cap;
      std::string e = multi_capacity(en.device, block, &cap);
      if (!e.empty()) return e;
      if (used[p] + c.trop_blocks <= cap) {
        en.packs[p].push_back(q);
        en.pack_block[p] = block;
        used[p] += c.trop_blocks;
        en.pack_of[q] = (int)p;
        placed = true;
      }
    }
    if (!placed) {
      long cap;
      std::string e = multi_capacity(en.device, c.trop_block, &cap);
      if (!e.empty()) return e;
      if (c.trop_blocks > cap) return "an ensemble member's subcycle tiling does not fit the device";
      en.packs.push_back({q});
      en.pack_block.push_back(c.trop_block);
      used.push_back(c.trop_blocks);
      en.pack_of[q] = (int)en.packs.size() - 1;
    }
  }
  // device arguments in pack order
  std::vector<TropTileArgs> host;
  en.pack_first.clear();
  for (auto& pk : en.packs) {
    en.pack_first.push_back((int)host.size());
    for (int q : pk) host.push_back(trop_args(*en.m[q]));
  }
  if (en.dargs) (void)hipFree(en.dargs);
  en.dargs = nullptr;
  if (!host.empty()) {
    hipError_t e = hipMalloc(&en.dargs, host.size() * sizeof(TropTileArgs));
    if (e != hipSuccess) return std::string("hipMalloc: ") + hipGetErrorString(e);
    e = hipMemcpy(en.dargs, host.data(), host.size() * sizeof(TropTileArgs), hipMemcpyHostToDevice);
    if (e != hipSuccess) return std::string("upload ensemble arguments: ") + hipGetErrorString(e);
  }
  for (hipEvent_t& ev : en.done) (void)hipEventDestroy(ev);
  en.done.assign(en.packs.size(), nullptr);
  for (hipEvent_t& ev : en.done) {
    hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    if (e != hipSuccess) return std::string("hipEventCreate: ") + hipGetErrorString(e);
  }
  return "";
}

}  // namespace

std::string ensemble_create(const std::vector<Ctx*>& members, Ensemble** out) {
  *out = nullptr;
  if (members.empty()) return "an ensemble needs at least one member";
  Ensemble* en = new Ensemble();
  en->m = members;
  en->device = members[0]->device;
  for (Ctx* c : members)
    if (c->device != en->device) {
      delete en;
      return "all ensemble members must be on one device";
    }
  hipError_t e = hipSetDevice(en->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&en->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete en;
    return std::string("hipStreamCreate: ") + hipGetErrorString(e);
  }
  en->ready.assign(members.size(), nullptr);
  for (hipEvent_t& ev : en->ready) (void)hipEventCreateWithFlags(&ev, hipEventDisableTiming);
  *out = en;
  return "";
}

void ensemble_destroy(Ensemble* en) {
  if (!en) return;
  (void)hipSetDevice(en->device);
  if (en->stream) (void)hipStreamSynchronize(en->stream);
  if (en->stream) trop_chain_forget(en->device, en->stream);
  for (hipEvent_t ev : en->ready) (void)hipEventDestroy(ev);
  for (hipEvent_t ev : en->done) (void)hipEventDestroy(ev);
  if (en->dargs) (void)hipFree(en->dargs);
  if (en->stream) (void)hipStreamDestroy(en->stream);
  delete en;
}

void ensemble_packs(const Ensemble& en, int* npacks, int* pack_of) {
  *npacks = (int)en.packs.size();
  for (size_t q = 0; q < en.m.size(); ++q) pack_of[q] = en.pack_of.empty() ? -1 : en.pack_of[q];
}

std::string ensemble_step(Ensemble& en, int nsteps) {
  hipError_t e = hipSetDevice(en.device);
  if (e != hipSuccess) return std::string("hipSetDevice: ") + hipGetErrorString(e);
  // tilings may have changed since the last call (autotuner)
  for (Ctx* c : en.m) (void)hipStreamSynchronize(c->stream);
  (void)hipStreamSynchronize(en.stream);
  std::string err = ensemble_plan(en);
  if (!err.empty()) return err;
  for (int s = 0; s < nsteps; ++s) {
    // members: the step up to the subcycle
    for (size_t q = 0; q < en.m.size(); ++q) {
      Ctx& c = *en.m[q];
      if (!(err = launch_step_pre(c)).empty()) return err;
      if (c.nn) {  // a member with a surrogate runs its own barotropic part
        if (!(err = nn_trop_step(c, launch_trop_kernel)).empty()) return err;
        continue;
      }
      if (c.trop_launches) {  // so does one whose subcycle runs in launch mode
        if (!(err = launch_trop_kernel(c)).empty()) return err;
        continue;
      }
      if ((e = hipEventRecord(en.ready[q], c.stream)) != hipSuccess) return hip_error_string(e, "hipEventRecord");
    }
    // packed subcycles
    for (size_t p = 0; p < en.packs.size(); ++p) {
      MultiLaunch ml{};
      ml.n = (int)en.packs[p].size();
      int blocks = 0;
      for (int k = 0; k < ml.n; ++k) {
        Ctx& c = *en.m[en.packs[p][k]];
        if ((e = hipStreamWaitEvent(en.stream, en.ready[en.packs[p][k]], 0)) != hipSuccess)
          return hip_error_string(e, "hipStreamWaitEvent");
        ml.start[k] = blocks;
        ml.epoch[k] = ++c.trop_epoch;
        blocks += c.trop_blocks;
      }
      ml.start[ml.n] = blocks;
      const TropTileArgs* args = en.dargs + en.pack_first[p];
      const int block = en.pack_block[p];
      const size_t lds = trop_lds_bytes(block);
      err = chained_launch(en.device, en.stream, [&]() -> std::string {
        hipError_t ea = hipFuncSetAttribute(trop_kernels(block, false).multi, hipFuncAttributeMaxDynamicSharedMemorySize,
                                            (int)lds);
        if (ea != hipSuccess) return hip_error_string(ea, "hipFuncSetAttribute");
        if (block == kTropSmall)
          hipLaunchKernelGGL(k_trop_multi<kTropSmall>, dim3(blocks), dim3(kTropSmall), lds, en.stream, args, ml);
        else
          hipLaunchKernelGGL(k_trop_multi<kTropBlock>, dim3(blocks), dim3(kTropBlock), lds, en.stream, args, ml);
        ea = hipGetLastError();
        if (ea != hipSuccess) return hip_error_string(ea, "k_trop_multi");
        return "";
      });
      if (!err.empty()) return err;
      if ((e = hipEventRecord(en.done[p], en.stream)) != hipSuccess) return hip_error_string(e, "hipEventRecord");
    }
    // members: the rest of the step
    for (size_t q = 0; q < en.m.size(); ++q) {
      Ctx& c = *en.m[q];
      if (en.pack_of[q] >= 0 && (e = hipStreamWaitEvent(c.stream, en.done[en.pack_of[q]], 0)) != hipSuccess)
        return hip_error_string(e, "hipStreamWaitEvent");
      c.n_ch += 1;
      if (!(err = launch_step_post(c)).empty()) return err;
    }
  }
  for (Ctx* c : en.m)
    if (!(err = settle_side(*c)).empty()) return err;
  return "";
}

// ---------------------------------------------------------------------------
// Autotuner: time the subcycle of this context under the tilings the cost
// model ranks best, on the context's own state (snapshotted and restored
// around every trial), and keep the fastest.
std::string tune_subcycle(Ctx& c, int trials, int out[3], double* best_ms) {
  // candidates of the mode in use, best modelled first
  std::vector<TropCand> cands;
  std::string err = trop_candidates(c, c.trop_launches, &cands);
  if (!err.empty()) return err;
  if (cands.empty()) return "no barotropic tiling fits this device";
  if ((int)cands.size() > trials) cands.resize(trials);
  // snapshot of what the subcycle writes
  Fields2& h = c.f2;
  double* fields[] = {h.chc, h.chp, h.mch, h.mch2, h.acch, h.buc, h.bvc, h.bup, h.bvp};
  const int nf = 9;
  double* snap = nullptr;
  hipError_t e = hipMalloc(&snap, sizeof(double) * c.n2 * nf);
  if (e != hipSuccess) return hip_error_string(e, "hipMalloc");
  auto copy = [&](bool save) -> hipError_t {
    for (int q = 0; q < nf; ++q) {
      hipError_t ec = save ? hipMemcpyAsync(snap + q * c.n2, fields[q], sizeof(double) * c.n2,
                                            hipMemcpyDeviceToDevice, c.stream)
                           : hipMemcpyAsync(fields[q], snap + q * c.n2, sizeof(double) * c.n2,
                                            hipMemcpyDeviceToDevice, c.stream);
      if (ec != hipSuccess) return ec;
    }
    return hipSuccess;
  };
  hipEvent_t ea, eb;
  (void)hipEventCreate(&ea);
  (void)hipEventCreate(&eb);
  const int keep[3] = {c.trop_tw, c.trop_th, c.trop_s};
  double best = 1e300;
  TropCand win = cands[0];
  if ((e = copy(true)) != hipSuccess) err = hip_error_string(e, "snapshot");
  for (const TropCand& cd : cands) {
    if (!err.empty()) break;
    c.trop_tw = cd.tw;
    c.trop_th = cd.th;
    c.trop_s = cd.s;
    if (!(err = trop_configure(c)).empty()) break;
    double t = 1e300;
    for (int r = 0; r < 3 && err.empty(); ++r) {
      if ((e = copy(false)) != hipSuccess) { err = hip_error_string(e, "restore"); break; }
      (void)hipEventRecord(ea, c.stream);
      if (!(err = launch_trop(c)).empty()) break;
      (void)hipEventRecord(eb, c.stream);
      if ((e = hipEventSynchronize(eb)) != hipSuccess) { err = hip_error_string(e, "sync"); break; }
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, ea, eb);
      t = std::min(t, (double)ms);
    }
    if (err.empty() && t < best) {
      best = t;
      win = cd;
    }
  }
  // the chosen tiling (or the previous one after a failure), state restored
  c.trop_tw = err.empty() ? win.tw : keep[0];
  c.trop_th = err.empty() ? win.th : keep[1];
  c.trop_s = err.empty() ? win.s : keep[2];
  std::string e2 = trop_configure(c);
  if (copy(false) != hipSuccess && err.empty()) err = "restore after tuning";
  (void)hipStreamSynchronize(c.stream);
  (void)hipEventDestroy(ea);
  (void)hipEventDestroy(eb);
  (void)hipFree(snap);
  if (!err.empty()) return err;
  if (!e2.empty()) return e2;
  out[0] = c.trop_tw;
  out[1] = c.trop_th;
  out[2] = c.trop_s;
  *best_ms = best;
  return "";
}

}  // namespace pcom
