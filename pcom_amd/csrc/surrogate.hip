// surrogate.hip - learned surrogate of the barotropic subcycle (int_trop,
// mod_int.f90:153-296): a 3x3 convolutional network in BF16 with FP32
// accumulation on the gfx950 matrix cores, trained on the device from
// snapshots the plain subcycle produces, and used in pcom_step to skip most
// subcycles.  The plain subcycle stays the fallback and the check: every
// check_every-th step runs it and measures the network's error there.
//
// What the network maps (one baroclinic step, interior glo_nj x glo_ni
// points, zonally periodic, zero padding at the walls).  Over one step the
// subcycle is, up to the small variation of its sqrt(ch) factors, a linear
// recurrence x_{n+1} = A x_n + b of its state x = (ch, u, v), A and b fixed
// for the step (b holds the forcing ut, vt, graphih and the offset of the
// initial viscosity vs1, mod_int.f90:197-248).  Its step means then differ
// from x_0 by a fixed linear map of the first substep's increment
// r = x_1 - x_0 alone: mean_n(x_n) - x_0 = mean_n sum_{k<n} A^k r.  So:
//   inputs (16 channels): r of ch, u, v (computed here exactly as the
//           first substep does it, in FP64) divided by a per-snapshot
//           scale s (the sea RMS of r_u, r_v), which makes the map
//           independent of how far the state is from balance; and what A
//           depends on, each normalised by its mean and standard deviation
//           over the recorded snapshots: the surface viscosity am, the
//           face integrals bphi and bgraphi of int_pgra (xn, xs, ye, yw),
//           the bottom pressure prh, the T- and U-point sea masks and the
//           Coriolis parameter
//   outputs (4 channels): mch2-chc, mch-chc, ubt-buc, vbt-bvc (the
//           step-mean free surface of the full and half cycle and the
//           step-mean barotropic velocity, as increments), divided by s
// The gravity waves of a 2*dtbc cycle travel some 15 cells at 1 degree, so
// the six 3x3 layers are dilated (1, 2, 4, 8, 4, 1): a receptive field of
// +-20 cells.  Applying a prediction writes what int_trop's end writes
// (mod_int.f90:276-287) with the mean of the free-surface increment over
// sea, weighted by area and bottom pressure, removed: the plain cycle's
// continuity equation conserves that mass.
//
// Tensors are NHWC with channels padded to multiples of 32 (the K width of
// v_mfma_f32_16x16x32_bf16).  Lane layout of that instruction (checked on
// the device): A row = lane % 16, k = 8 (lane / 16) + e; B column =
// lane % 16, same k; D row = 4 (lane / 16) + i, column = lane % 16.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "context.h"

namespace pcom {

namespace {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef unsigned short bf16_t;

constexpr int kInC = 32;    // padded input channels (16 used)
constexpr int kHid = 64;    // hidden channels
constexpr int kOutC = 32;   // padded output channels (4 used)
// recorded per snapshot (kDyn raw channels): r_ch, r_u, r_v (kScaled of
// them, divided by the snapshot's scale), am, bphi xn xs ye yw, bgraphi xn
// xs ye yw; network input channels 0-3 and 8-15.  Static: prh 4, sea T 5,
// sea U 6, Coriolis 7.
constexpr int kInUsed = 16, kOutUsed = 4, kDyn = 12, kScaled = 3, kStat = 4;
__host__ __device__ constexpr int dyn_channel(int d) { return d < 4 ? d : d + 4; }
constexpr int kLayers = 6;  // 32->64, 4 x 64->64, 64->32 (last linear)
// dilation of each layer's 3x3 taps
__host__ __device__ constexpr int layer_dil(int l) { return l == 0 || l == 5 ? 1 : (l == 1 ? 2 : (l == 3 ? 8 : 4)); }

__host__ __device__ inline bf16_t f2bf(float f) {
  unsigned u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return (bf16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (bf16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(bf16_t h) {
  unsigned u = (unsigned)h << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

struct Geo {
  int n, H, W, dil;  // images, rows, columns, tap spacing
  __host__ __device__ long npix() const { return (long)n * H * W; }
  // column x + o, zonally periodic (any offset)
  __device__ __forceinline__ int wrap(int x) const {
    int r = x % W;
    return r < 0 ? r + W : r;
  }
};

__device__ __forceinline__ bf8 load8(const bf16_t* p) {
  uint4 v = *reinterpret_cast<const uint4*>(p);
  return __builtin_bit_cast(bf8, v);
}
__device__ __forceinline__ bf8 zero8() { return __builtin_bit_cast(bf8, make_uint4(0, 0, 0, 0)); }


// ---- the forward convolution, and the backward data pass, with the layer's
// weights staged in LDS once per workgroup of 8 waves (a wave re-reading
// them through the vector cache streams 9 KIN KOUT BF16 from L2 per pixel
// block otherwise).  out[p][o] = sum_t,i Wl[o][t][i] in[p + s d t][i] with
// KIN input and KOUT output channels: the forward pass (s = +1, Wl = W,
// bias, then ReLU to BF16 or linear to FP32) or the data gradient (s = -1,
// Wl = W transposed to [ci][dy][dx][co], masked by the ReLU whose output
// `post` is the layer's input).  LDS holds 16-byte B fragments in the order
// [tap][c][kq][o], so the 16 lanes of one kq read 256 contiguous bytes.
// Each wave takes PB pixel blocks at a time, grid-strided.
constexpr int kConvLdsWaves = 8;
enum ConvMode { kFwdRelu, kFwdLinear, kBwdData };
template <int KIN, int KOUT>
constexpr size_t conv_lds_bytes() { return (size_t)9 * KIN * KOUT * 2; }

template <int KIN, int KOUT, int MODE, int PB, int NW = kConvLdsWaves>
__global__ void __launch_bounds__(64 * NW) k_conv_lds(Geo g, const bf16_t* __restrict__ in,
                                                                 const bf16_t* __restrict__ w,
                                                                 const float* __restrict__ bias,
                                                                 const bf16_t* __restrict__ post, void* out) {
  extern __shared__ uint4 wl[];
  constexpr int NC = KIN / 32, NF = 9 * NC * 4 * KOUT, S = MODE == kBwdData ? -1 : 1;
  // global order (o, tap, 8-channel group), LDS order (tap, c, kq, o)
  for (int f = threadIdx.x; f < NF; f += blockDim.x) {
    const int o = f / (9 * NC * 4), rem = f % (9 * NC * 4);
    const int t = rem / (NC * 4), c8 = rem % (NC * 4);
    wl[((t * NC + c8 / 4) * 4 + c8 % 4) * KOUT + o] = reinterpret_cast<const uint4*>(w)[f];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int kq = lane >> 4, r = lane & 15;
  const long np = g.npix(), npb = np / 16;
  float bb[KOUT / 16];
#pragma unroll
  for (int nb = 0; nb < KOUT / 16; ++nb) bb[nb] = MODE == kBwdData ? 0.f : bias[nb * 16 + r];
  const long stride = (long)gridDim.x * NW * PB;
  for (long pb0 = ((long)xcd_block() * NW + wave) * PB; pb0 < npb; pb0 += stride) {
    int x[PB], y[PB], n[PB];
    bool live[PB];
#pragma unroll
    for (int b = 0; b < PB; ++b) {
      live[b] = pb0 + b < npb;
      const long pc = live[b] ? (pb0 + b) * 16 + r : 0;
      x[b] = (int)(pc % g.W);
      const long t = pc / g.W;
      y[b] = (int)(t % g.H);
      n[b] = (int)(t / g.H);
    }
    f4 acc[PB][KOUT / 16];
#pragma unroll
    for (int b = 0; b < PB; ++b)
#pragma unroll
      for (int nb = 0; nb < KOUT / 16; ++nb) acc[b][nb] = f4{0, 0, 0, 0};
    // one row of taps at a time (the registers of all nine would spill)
#pragma unroll 1
    for (int dy = 0; dy < 3; ++dy) {
#pragma unroll
      for (int dx = 0; dx < 3; ++dx) {
        const bf16_t* src[PB];
        bool ok[PB];
#pragma unroll
        for (int b = 0; b < PB; ++b) {
          const int yy = y[b] + S * (dy - 1) * g.dil;
          ok[b] = live[b] && yy >= 0 && yy < g.H;
          const int xx = g.wrap(x[b] + S * (dx - 1) * g.dil);
          src[b] = in + (((long)n[b] * g.H + (ok[b] ? yy : 0)) * g.W + xx) * KIN + kq * 8;
        }
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          bf8 a[PB];
#pragma unroll
          for (int b = 0; b < PB; ++b) a[b] = ok[b] ? load8(src[b] + c * 32) : zero8();
#pragma unroll
          for (int nb = 0; nb < KOUT / 16; ++nb) {
            const bf8 wb = __builtin_bit_cast(bf8, wl[(((dy * 3 + dx) * NC + c) * 4 + kq) * KOUT + nb * 16 + r]);
#pragma unroll
            for (int b = 0; b < PB; ++b)
              acc[b][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[b], wb, acc[b][nb], 0, 0, 0);
          }
        }
      }
    }
#pragma unroll
    for (int b = 0; b < PB; ++b) {
      if (!live[b]) continue;
#pragma unroll
      for (int nb = 0; nb < KOUT / 16; ++nb) {
        const int o = nb * 16 + r;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const long e = ((pb0 + b) * 16 + kq * 4 + i) * KOUT + o;
          const float v = acc[b][nb][i] + bb[nb];
          if (MODE == kFwdRelu) static_cast<bf16_t*>(out)[e] = f2bf(v > 0.f ? v : 0.f);
          if (MODE == kFwdLinear) static_cast<float*>(out)[e] = v;
          if (MODE == kBwdData) static_cast<bf16_t*>(out)[e] = bf2f(post[e]) > 0.f ? f2bf(v) : (bf16_t)0;
        }
      }
    }
  }
}


// ---- backward weights, one tap (dy, dx) per blockIdx.y and one pixel chunk
// per blockIdx.x: part[chunk][tap][co][ci] = sum_p dpre[p][co] in[p+d(dy-1,dx-1)][ci].
// GEMM with M = co, N = ci, K = pixels.  Tiles of 64 pixels of both
// operands are staged in LDS transposed ([channel][pixel], rows padded to
// 72 entries), so each matrix-core operand (8 consecutive pixels of one
// channel) is one 16-byte LDS read; a thread stages 4 pixels x 8 channels
// of one operand per tile (four 16-byte global loads, eight 8-byte LDS
// writes).  The MB x NB output tiles of 16 x 16 are dealt to the 4 waves.
constexpr int kWTile = 64, kWPad = 72;
template <int CIN, int COUT>
__global__ void __launch_bounds__(256) k_conv_bwd_w(Geo g, const bf16_t* __restrict__ in,
                                                    const bf16_t* __restrict__ dpre, float* part, long chunk) {
  __shared__ __align__(16) bf16_t sD[COUT][kWPad];
  __shared__ __align__(16) bf16_t sI[CIN][kWPad];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int kq = lane >> 4, r = lane & 15;
  const int tap = blockIdx.y, dy = tap / 3, dx = tap % 3;
  const long np = g.npix();
  const long pbeg = (long)blockIdx.x * chunk;
  const long pend = pbeg + chunk < np ? pbeg + chunk : np;
  constexpr int MB = COUT / 16, NB = CIN / 16, T = MB * NB, TW = (T + 3) / 4;
  f4 acc[TW];
#pragma unroll
  for (int j = 0; j < TW; ++j) acc[j] = f4{0, 0, 0, 0};
  // staging tasks: (4-pixel quad, 8-channel group) of dpre, then of the input
  constexpr int QD = (kWTile / 4) * (COUT / 8), QI = (kWTile / 4) * (CIN / 8);
  for (long ps = pbeg; ps < pend; ps += kWTile) {
    for (int task = threadIdx.x; task < QD + QI; task += 256) {
      const bool isd = task < QD;
      const int tt = isd ? task : task - QD;
      const int ng = isd ? COUT / 8 : CIN / 8;
      const int quad = tt / ng, cg = tt % ng;
      uint4 v[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const long p = ps + quad * 4 + e;
        v[e] = make_uint4(0, 0, 0, 0);
        if (p < pend) {
          if (isd) {
            v[e] = *reinterpret_cast<const uint4*>(dpre + p * COUT + cg * 8);
          } else {
            const int x = (int)(p % g.W);
            const long t = p / g.W;
            const int y = (int)(t % g.H), n = (int)(t / g.H);
            const int yy = y + (dy - 1) * g.dil;
            const int xx = g.wrap(x + (dx - 1) * g.dil);
            if (yy >= 0 && yy < g.H) v[e] = *reinterpret_cast<const uint4*>(in + (((long)n * g.H + yy) * g.W + xx) * CIN + cg * 8);
          }
        }
      }
      const unsigned short* h0 = reinterpret_cast<const unsigned short*>(&v[0]);
      const unsigned short* h1 = reinterpret_cast<const unsigned short*>(&v[1]);
      const unsigned short* h2 = reinterpret_cast<const unsigned short*>(&v[2]);
      const unsigned short* h3 = reinterpret_cast<const unsigned short*>(&v[3]);
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const unsigned long long q = (unsigned long long)h0[c] | ((unsigned long long)h1[c] << 16) |
                                     ((unsigned long long)h2[c] << 32) | ((unsigned long long)h3[c] << 48);
        bf16_t* dst = isd ? &sD[cg * 8 + c][quad * 4] : &sI[cg * 8 + c][quad * 4];
        *reinterpret_cast<unsigned long long*>(dst) = q;
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < TW; ++j) {
      const int t = wave * TW + j;
      if (t < T) {
        const int mb = t / NB, nb = t % NB;
#pragma unroll
        for (int kb = 0; kb < kWTile / 32; ++kb) {
          const bf8 A = __builtin_bit_cast(bf8, *reinterpret_cast<const uint4*>(&sD[mb * 16 + r][kb * 32 + kq * 8]));
          const bf8 B = __builtin_bit_cast(bf8, *reinterpret_cast<const uint4*>(&sI[nb * 16 + r][kb * 32 + kq * 8]));
          acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, acc[j], 0, 0, 0);
        }
      }
    }
    __syncthreads();
  }
  float* dst = part + ((long)blockIdx.x * 9 + tap) * COUT * CIN;
#pragma unroll
  for (int j = 0; j < TW; ++j) {
    const int t = wave * TW + j;
    if (t < T) {
      const int mb = t / NB, nb = t % NB;
#pragma unroll
      for (int i = 0; i < 4; ++i) dst[(mb * 16 + kq * 4 + i) * CIN + nb * 16 + r] = acc[j][i];
    }
  }
}

// sum the chunk partials in chunk order into gw[co][dy][dx][ci]; bias
// gradient partials likewise into gb[co]
__global__ void k_reduce_w(const float* part, int nchunk, int cout, int cin, float* gw) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long m = 9L * cout * cin;
  if (e >= m) return;
  const int ci = (int)(e % cin);
  const long t = e / cin;
  const int co = (int)(t % cout), dydx = (int)(t / cout);
  // four independent running sums (chunks c = 4 m + j), added at the end
  const long stride = 9L * cout * cin, off = (long)dydx * cout * cin + (long)co * cin + ci;
  float s4[4] = {0.f, 0.f, 0.f, 0.f};
  int c = 0;
  for (; c + 4 <= nchunk; c += 4)
#pragma unroll
    for (int j = 0; j < 4; ++j) s4[j] += part[(long)(c + j) * stride + off];
  for (; c < nchunk; ++c) s4[c & 3] += part[(long)c * stride + off];
  gw[((long)co * 9 + dydx) * cin + ci] = (s4[0] + s4[1]) + (s4[2] + s4[3]);
}

// bias gradient partials: part[block][co] = sum over the block's kBiasChunk
// pixels of dpre[p][co]; 256 / cout groups of cout threads take every
// (256 / cout)-th pixel (each group's loads are one contiguous row), folded
// in LDS in group order
constexpr int kBiasChunk = 256;
__global__ void __launch_bounds__(256) k_bias_part(Geo g, const bf16_t* __restrict__ dpre, int cout, float* part) {
  __shared__ float sh[256];
  const int groups = 256 / cout;
  const int co = threadIdx.x % cout, grp = threadIdx.x / cout;
  const long np = g.npix();
  const long pbeg = (long)blockIdx.x * kBiasChunk;
  float s = 0.f;
  if (grp < groups)
    for (long p = pbeg + grp; p < pbeg + kBiasChunk && p < np; p += groups) s += bf2f(dpre[p * cout + co]);
  sh[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x < cout) {
    float t = 0.f;
    for (int q = 0; q < groups; ++q) t += sh[q * cout + threadIdx.x];
    part[(long)blockIdx.x * cout + threadIdx.x] = t;
  }
}

// one block per output channel folds the partials: 256 strided lanes, then
// a tree in LDS
__global__ void __launch_bounds__(256) k_bias_reduce(const float* part, int nchunk, int cout, float* gb) {
  __shared__ float sh[256];
  const int co = blockIdx.x;
  float s = 0.f;
  for (int c = threadIdx.x; c < nchunk; c += 256) s += part[(long)c * cout + co];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) gb[co] = sh[0];
}

// the sums over nb per-block partials of NF fields (part[NF * b + f]) by
// one block of 256 threads: a strided pass per thread, then a tree in LDS
// (a fixed order, so the result does not vary from run to run)
template <int NF>
__device__ __forceinline__ void fold_parts(const double* part, int nb, double out[NF]) {
  __shared__ double sh[NF][256];
  double a[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) a[f] = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x)
#pragma unroll
    for (int f = 0; f < NF; ++f) a[f] += part[NF * b + f];
#pragma unroll
  for (int f = 0; f < NF; ++f) sh[f][threadIdx.x] = a[f];
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st)
#pragma unroll
      for (int f = 0; f < NF; ++f) sh[f][threadIdx.x] += sh[f][threadIdx.x + st];
    __syncthreads();
  }
#pragma unroll
  for (int f = 0; f < NF; ++f) out[f] = sh[f][0];
}

// masked mean-squared error over the 4 used outputs; channel c is masked by
// input channel 5 (T-point sea) for c < 2 and 6 (U-point sea) otherwise.
// dpre = d(loss)/d(out) in BF16; per-block partial sums of loss and count.
__global__ void k_loss(long np, const float* __restrict__ out, const float* __restrict__ tgt,
                       const bf16_t* __restrict__ x, float inv_count, bf16_t* dpre, double* part) {
  __shared__ double sh[256];
  const long p = (long)xcd_block() * blockDim.x + threadIdx.x;
  double l = 0.0;
  if (p < np) {
    const float mt = bf2f(x[p * kInC + 5]), mu = bf2f(x[p * kInC + 6]);
    for (int c = 0; c < kOutC; ++c) {
      float d = 0.f;
      if (c < kOutUsed) {
        const float m = c < 2 ? mt : mu;
        d = m * (out[p * kOutC + c] - tgt[p * kOutUsed + c]);
        l += (double)d * d;
      }
      dpre[p * kOutC + c] = f2bf(2.f * d * inv_count);
    }
  }
  sh[threadIdx.x] = l;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(256) k_sum_parts(const double* part, int n, double* out) {
  double s[1];
  fold_parts<1>(part, n, s);
  if (threadIdx.x == 0) *out = s[0];
}

// Adam on the FP32 master weights, then the BF16 copies: w[co][dy][dx][ci]
// and its transpose wt[ci][dy][dx][co] for the backward-data pass
__global__ void k_adam(long n, float* w32, float* m, float* v, const float* g, float lr, float b1, float b2,
                       float eps, float bc1, float bc2) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const float gg = g[e];
  if (!isfinite(gg)) return;  // a non-finite gradient leaves the weight and its moments
  const float mm = b1 * m[e] + (1.f - b1) * gg;
  const float vv = b2 * v[e] + (1.f - b2) * gg * gg;
  m[e] = mm;
  v[e] = vv;
  w32[e] -= lr * (mm / bc1) / (sqrtf(vv / bc2) + eps);
}

__global__ void k_w_bf16(int cout, int cin, const float* w32, bf16_t* w, bf16_t* wt) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 9L * cout * cin) return;
  const int ci = (int)(e % cin);
  const long t = e / cin;
  const int dydx = (int)(t % 9), co = (int)(t / 9);
  const bf16_t h = f2bf(w32[e]);
  w[e] = h;
  wt[((long)ci * 9 + dydx) * cout + co] = h;
}

// ---- the first substep of the plain cycle (mod_int.f90:184-273 at nt = 1,
// where vs - vs1 = 0), from the state the cycle starts from (chp, bup, bvp),
// on every point of the halo'd arrays; the operators of pcom_dev.h apply
// the reference's halo rule.  Predictor free surface first:
__global__ void k_sub1_pred(Grid<double> g, const double* __restrict__ chp, const double* __restrict__ bup,
                            const double* __restrict__ bvp, const double* __restrict__ chc, double bt,
                            double* chpred) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.n2) return;
  const int i = (int)(p % g.ni) + 1, j = (int)(p / g.ni) + 1;
  auto CHP = [&](int m, int jj) -> double { return chp[g.ix(m, jj)]; };
  auto SCH = [&](int m, int jj) -> double {
    return g.lev_u[g.ix(m, jj)] > 0 ? sqrt(ter_di<1>(g, CHP, m, jj, 0.0)) : 1.0;
  };
  auto FU = [&](int m, int jj) -> double { return g.prh_u[g.ix(m, jj)] * SCH(m, jj) * bup[g.ix(m, jj)]; };
  auto FV = [&](int m, int jj) -> double { return g.prh_u[g.ix(m, jj)] * SCH(m, jj) * bvp[g.ix(m, jj)]; };
  const double wk = chp[p] - divg<3, 3, 1>(g, FU, FV, i, j) * bt * GAMMA_B / g.prh_t[p];
  chpred[p] = g.lev_t[p] > 0 ? wk : chc[p];
}

struct Sub1In {
  const double *bup, *bvp, *but, *bvt, *ghx, *ghy, *cor;
  const double *bphi[4], *bgra[4];
};

// the momentum stage: pressure gradient of the predictor free surface,
// Coriolis predictor-corrector, new barotropic velocity (and its sqrt(ch)
// factor for the corrector's fluxes)
__global__ void k_sub1_mom(Grid<double> g, Sub1In a, const double* __restrict__ chpred, double bt, double* u1,
                           double* v1, double* sch1) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.n2) return;
  const int i = (int)(p % g.ni) + 1, j = (int)(p / g.ni) + 1;
  const int ic = g.canon(i);
  auto CP = [&](int m, int jj) -> double { return chpred[g.ix(m, jj)]; };
  const double su = g.lev_u[p] >= 1 ? 1.0 : 0.0;
  const double sch = g.lev_u[p] > 0 ? sqrt(ter_di<1>(g, CP, i, j, 0.0)) : 1.0;
  auto GX = [&](int m, int jj) -> double { return grax<1, 2>(g, CP, m, jj); };
  auto GY = [&](int m, int jj) -> double { return gray<1, 4>(g, CP, m, jj); };
  auto TE = [&](int m, int jj) -> double { return ter_ew<1>(g, CP, m, jj); };
  auto TN = [&](int m, int jj) -> double { return ter_ns<1>(g, CP, m, jj, 0.0); };
  double gi1 = 0.0, ci1 = 0.0;
  if (j <= g.njm) {
    gi1 = 0.5 * (GX(ic, j) * a.bphi[0][g.ix(ic, j)] + GX(ic, j + 1) * a.bphi[1][g.ix(ic, j + 1)]);
    ci1 = 0.5 * (TE(ic, j) * a.bgra[0][g.ix(ic, j)] + TE(ic, j + 1) * a.bgra[1][g.ix(ic, j + 1)]);
  }
  const double gi2 = 0.5 * (GY(ic, j) * a.bphi[2][g.ix(ic, j)] + GY(ic + 1, j) * a.bphi[3][g.ix(ic + 1, j)]);
  const double ci2 = 0.5 * (TN(ic, j) * a.bgra[2][g.ix(ic, j)] + TN(ic + 1, j) * a.bgra[3][g.ix(ic + 1, j)]);
  const double pg1 = (gi1 + ci1 + a.ghx[p]) * su * sch;
  const double pg2 = (gi2 + ci2 + a.ghy[p]) * su * sch;
  const double up = a.bup[p], vp = a.bvp[p], f = a.cor[p];
  double pu1 = su * (a.but[p] - pg1 + 0.5 * f * vp);
  double pu2 = su * (a.bvt[p] - pg2 - 0.5 * f * up);
  const double w1 = pu1 * bt + up, w2 = pu2 * bt + vp;
  const double x = 0.5 * f * bt;
  double uc = (w1 + 0.5 * f * bt * w2) / (1 + x * x);
  double vc = (w2 - 0.5 * f * bt * w1) / (1 + x * x);
  uc = 0.5 * (uc + up);
  vc = 0.5 * (vc + vp);
  pu1 = su * (0.0 - pg1 + a.but[p] + f * vc);
  pu2 = su * (0.0 - pg2 + a.bvt[p] - f * uc);
  u1[p] = pu1 * bt + up;
  v1[p] = pu2 * bt + vp;
  sch1[p] = sch;
}

// the network's raw inputs of one snapshot, kDyn channels [kDyn][HW]
struct RawIn {
  const double *chp, *bup, *bvp, *u1, *v1, *sch1, *am1;
  const double *bphi[4], *bgra[4];
};

// corrector free surface, then r = x_1 - x_0 and the other raw channels of
// the interior pixel q = y * W + x (Fortran (i, j) = (x + 2, y + 2)); and
// the snapshot's scale s = sea RMS of r_u, r_v (as the FP32 channels hold
// them): per-block partials of (sum r^2, count) in part, which one block
// folds in order into s (k_scale_final, or the live path's k_pack_live)
__global__ void __launch_bounds__(256) k_sub1_raw(Grid<double> g, int H, int W, RawIn a, double bt, float* raw,
                                                  const float* stat, double* part) {
  __shared__ double s1[256], s2[256];
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long hw = (long)H * W;
  double ssq = 0.0, cnt = 0.0;
  if (q < hw) {
    const int i = (int)(q % W) + 2, j = (int)(q / W) + 2;
    const long p = g.ix(i, j);
    auto FU = [&](int m, int jj) -> double {
      const long k = g.ix(m, jj);
      return g.prh_u[k] * a.sch1[k] * a.u1[k];
    };
    auto FV = [&](int m, int jj) -> double {
      const long k = g.ix(m, jj);
      return g.prh_u[k] * a.sch1[k] * a.v1[k];
    };
    const double ch1 =
        a.chp[p] - (double)(g.lev_t[p] >= 1 ? 1 : 0) * divg<3, 3, 1>(g, FU, FV, i, j) * bt / g.prh_t[p];
    const float ru = (float)(a.u1[p] - a.bup[p]), rv = (float)(a.v1[p] - a.bvp[p]);
    raw[q] = (float)(ch1 - a.chp[p]);
    raw[hw + q] = ru;
    raw[2 * hw + q] = rv;
    raw[3 * hw + q] = (float)a.am1[p];
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      raw[(4 + f) * hw + q] = (float)a.bphi[f][p];
      raw[(8 + f) * hw + q] = (float)a.bgra[f][p];
    }
    const double mu = (double)stat[hw + q];  // sea U mask
    const double u = ru, v = rv;
    ssq = mu * (u * u + v * v);
    cnt = 2.0 * mu;
  }
  s1[threadIdx.x] = ssq;
  s2[threadIdx.x] = cnt;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
      s1[threadIdx.x] += s1[threadIdx.x + st];
      s2[threadIdx.x] += s2[threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s1[0];
    part[2 * blockIdx.x + 1] = s2[0];
  }
}

// the snapshot scale s = sea RMS of (r_u, r_v) from k_sub1_raw's partials
__device__ __forceinline__ double fold_scale(const double* part, int nb) {
  double a[2];
  fold_parts<2>(part, nb, a);
  return a[1] > 0.0 && a[0] > 0.0 ? sqrt(a[0] / a[1]) : 1e-30;
}

__global__ void __launch_bounds__(256) k_scale_final(const double* part, int nb, double* scale, float* keep) {
  const double sc = fold_scale(part, nb);
  if (threadIdx.x == 0) {
    *scale = sc;
    if (keep) *keep = (float)sc;
  }
}

// ---- perturbed snapshots.  The network must also map states it makes
// itself, which carry unbalanced errors the plain model never shows; the
// plain cycle damps them (a gravity wave averages out over the cycle), and
// snapshots of perturbed states teach the network that response.  Noise:
// a hashed uniform value per (field, pixel, seed), smoothed by `npass`
// passes of a 5-point average (zonally periodic, rows clamped at the walls),
// scaled to a chosen sea RMS.
__device__ __forceinline__ float hash_uniform(unsigned long long seed, unsigned f, long q) {
  unsigned long long z = seed + 0x9E3779B97F4A7C15ull * ((unsigned long long)q * 4 + f + 1);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (float)((double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0);
}

__global__ void k_noise(int H, int W, unsigned long long seed, float* out) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long hw = (long)H * W;
  if (q >= 3 * hw) return;
  out[q] = hash_uniform(seed, (unsigned)(q / hw), q % hw);
}

__global__ void k_smooth(int H, int W, const float* in, float* out) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long hw = (long)H * W;
  if (e >= 3 * hw) return;
  const long base = (e / hw) * hw, q = e % hw;
  const int x = (int)(q % W), y = (int)(q / W);
  const int xw = x == 0 ? W - 1 : x - 1, xe = x == W - 1 ? 0 : x + 1;
  const int ys = y == 0 ? 0 : y - 1, yn = y == H - 1 ? H - 1 : y + 1;
  out[e] = 0.2f * (in[base + q] + in[base + (long)y * W + xw] + in[base + (long)y * W + xe] +
                   in[base + (long)ys * W + x] + in[base + (long)yn * W + x]);
}

// per-block sums of noise^2 over sea (field 0 on T points, 1 and 2 on U)
__global__ void k_noise_part(int H, int W, const float* nz, const float* stat, double* part) {
  __shared__ double s[4][256];
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long hw = (long)H * W;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  if (q < hw) {
    const double mt = stat[q], mu = stat[hw + q];
    v[0] = mt * (double)nz[q] * nz[q];
    v[1] = mu * ((double)nz[hw + q] * nz[hw + q] + (double)nz[2 * hw + q] * nz[2 * hw + q]);
    v[2] = mt;
    v[3] = 2.0 * mu;
  }
  for (int c = 0; c < 4; ++c) s[c][threadIdx.x] = v[c];
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st)
      for (int c = 0; c < 4; ++c) s[c][threadIdx.x] += s[c][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int c = 0; c < 4; ++c) part[4 * blockIdx.x + c] = s[c][0];
}

// the factors that scale the noise of ch and of (u, v) to unit sea RMS
__global__ void __launch_bounds__(256) k_noise_final(const double* part, int nb, double* inv_rms) {
  double a[4];
  fold_parts<4>(part, nb, a);
  if (threadIdx.x == 0) {
    inv_rms[0] = a[0] > 0.0 ? sqrt(a[2] / a[0]) : 0.0;
    inv_rms[1] = a[1] > 0.0 ? sqrt(a[3] / a[1]) : 0.0;
  }
}

// add amp_ch (ch) and amp_u (u, v) times the unit-RMS noise to the state
// the cycle starts from, on sea points of the interior and their periodic
// halo copies
__global__ void k_perturb(int ni, int H, int W, const float* nz, const double* inv_rms, double amp_ch,
                          double amp_u, const float* stat, double* chc, double* chp, double* buc, double* bvc,
                          double* bup, double* bvp) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long hw = (long)H * W;
  if (q >= hw) return;
  const double rc = inv_rms[0], ru = inv_rms[1];
  const int x = (int)(q % W), y = (int)(q / W);
  const double dc = stat[q] * amp_ch * rc * nz[q];
  const double du = stat[hw + q] * amp_u * ru * nz[hw + q];
  const double dv = stat[hw + q] * amp_u * ru * nz[2 * hw + q];
  const int i = x + 2;
  const int cols[2] = {i, i == W + 1 ? 1 : (i == 2 ? W + 2 : 0)};
  for (int k = 0; k < 2; ++k) {
    if (cols[k] == 0) continue;
    const long o = (long)(cols[k] - 1) + (long)ni * (y + 1);
    chc[o] += dc;
    chp[o] = chc[o];
    buc[o] += du;
    bup[o] = buc[o];
    bvc[o] += dv;
    bvp[o] = bvc[o];
  }
}

// ---- data movement between the snapshots and the network tensors.
// Static per-pixel channels (stat, [kStat][HW]): T-point sea, U-point sea,
// Coriolis / its largest magnitude, normalised bottom pressure.
struct NnNorm {
  float in_mean[kDyn], in_std[kDyn], out_std[kOutUsed];
  float cor_scale, prh_scale;
  // least snapshot scale s: a fraction of the recorded snapshots' median,
  // so that a snapshot of unusually small first-substep velocity increment
  // cannot blow its divided values up
  float s_floor;
};
constexpr int kNormValues = 2 * kDyn + kOutUsed + 3;
constexpr float kFloorOfMedian = 0.2f;
// normalised values beyond this are clamped: far above any real snapshot's,
// only a guard against infinities
constexpr float kNormClip = 1024.f;
__device__ __forceinline__ float floored(float s, const NnNorm& nm) { return s > nm.s_floor ? s : nm.s_floor; }
__device__ __forceinline__ float nclip(float v) { return fminf(fmaxf(v, -kNormClip), kNormClip); }

__device__ __forceinline__ long nn_ix(int ni, int x, int y) { return (long)(x + 1) + (long)ni * (y + 1); }

// the outputs of one plain subcycle as increments over its initial state
// (base: chc, chc, buc, bvc of the interior pixels)
__global__ void k_rec_out(int ni, int H, int W, const double* mch2, const double* mch, const double* buc,
                          const double* bvc, const double* base, float* tgt) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long hw = (long)H * W;
  if (q >= hw) return;
  const long k = nn_ix(ni, (int)(q % W), (int)(q / W));
  tgt[q] = (float)(mch2[k] - base[q]);
  tgt[hw + q] = (float)(mch[k] - base[hw + q]);
  tgt[2 * hw + q] = (float)(buc[k] - base[2 * hw + q]);
  tgt[3 * hw + q] = (float)(bvc[k] - base[3 * hw + q]);
}

// channel moments of the store (one block per channel, FP64, fixed order);
// the first `nscaled` channels are divided by their snapshot's scale
__global__ void k_moments(const float* store, long per_snap, long hw, int nsnap, const float* scale, int nscaled,
                          float s_floor, double* out) {
  __shared__ double s1[256], s2[256];
  const int c = blockIdx.x;
  double a = 0.0, b = 0.0;
  for (int sn = 0; sn < nsnap; ++sn) {
    const double inv = c < nscaled ? 1.0 / (double)(scale[sn] > s_floor ? scale[sn] : s_floor) : 1.0;
    for (long q = threadIdx.x; q < hw; q += blockDim.x) {
      const double v = store[(long)sn * per_snap + (long)c * hw + q] * inv;
      a += v;
      b += v * v;
    }
  }
  s1[threadIdx.x] = a;
  s2[threadIdx.x] = b;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      s1[threadIdx.x] += s1[threadIdx.x + s];
      s2[threadIdx.x] += s2[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[2 * c] = s1[0];
    out[2 * c + 1] = s2[0];
  }
}

// one pixel's network input (BF16, 32 channels) from its raw channels and
// the snapshot's scale
__device__ __forceinline__ void pack_pixel(const float* raw, long hw, long q, float sc, const NnNorm& nm,
                                           const float* stat, bf16_t* xo) {
  __attribute__((aligned(16))) bf16_t v[kInC];
#pragma unroll
  for (int d = 0; d < kDyn; ++d) {
    const float r = raw[d * hw + q];
    v[dyn_channel(d)] = f2bf(nclip(d < kScaled ? r / (sc * nm.in_std[d]) : (r - nm.in_mean[d]) / nm.in_std[d]));
  }
  v[4] = f2bf(stat[3 * hw + q]);
  v[5] = f2bf(stat[q]);
  v[6] = f2bf(stat[hw + q]);
  v[7] = f2bf(stat[2 * hw + q]);
#pragma unroll
  for (int c = kInUsed; c < kInC; ++c) v[c] = 0;
  // one pixel's 64 bytes as four 16-byte stores
#pragma unroll
  for (int k = 0; k < kInC / 8; ++k) reinterpret_cast<uint4*>(xo)[k] = reinterpret_cast<const uint4*>(v)[k];
}

// a minibatch of recorded snapshots into the network input and targets
// (FP32, 4 channels, divided by the scale and the output normalisation)
__global__ void k_pack(int H, int W, const float* sin, const float* stg, const float* sscale, const int* idx,
                       int batch, NnNorm nm, const float* stat, bf16_t* x, float* t) {
  const long hw = (long)H * W;
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= hw * batch) return;
  const int b = (int)(e / hw);
  const long q = e % hw;
  const long s = idx[b];
  const float sc = floored(sscale[s], nm);
  pack_pixel(sin + s * kDyn * hw, hw, q, sc, nm, stat, x + e * kInC);
  const float* tg = stg + s * kOutUsed * hw;
#pragma unroll
  for (int c = 0; c < kOutUsed; ++c) t[e * kOutUsed + c] = nclip(tg[c * hw + q] / (sc * nm.out_std[c]));
}

// the live snapshot into the network input (one image)
// the live state's network input; every block folds the scale partials
// itself (the same fixed order, so the same s), block 0 keeps s in *scale
__global__ void __launch_bounds__(256) k_pack_live(int H, int W, const float* raw, const double* part, int nb,
                                                   double* scale, NnNorm nm, const float* stat, bf16_t* x) {
  const double sc = floored((float)fold_scale(part, nb), nm);
  if (blockIdx.x == 0 && threadIdx.x == 0) *scale = sc;
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long hw = (long)H * W;
  if (q >= hw) return;
  pack_pixel(raw, hw, q, (float)sc, nm, stat, x + q * kInC);
}

// prediction -> increments (FP64) with per-block sums over sea of
// w*d(mch2), w*d(mch) and w, w = area * prh of the T points: the barotropic
// continuity equation (mod_int.f90:189-193, 264-267) conserves sum(w chc)
struct NnLimit {
  double v[kOutUsed];  // largest increment magnitude per output (0: no bound)
};

__global__ void k_pred(int ni, int H, int W, const float* out, NnNorm nm, const double* scale, const float* stat,
                       const double* area, NnLimit lim, double* inc, double* part) {
  __shared__ double s[3][256];
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long hw = (long)H * W;
  double v[3] = {0.0, 0.0, 0.0};
  if (q < hw) {
    const double sc = *scale;
    const float mt = stat[q], mu = stat[hw + q];
#pragma unroll
    for (int c = 0; c < kOutUsed; ++c) {
      double d = (double)(out[q * kOutC + c] * nm.out_std[c]) * sc * (double)(c < 2 ? mt : mu);
      // a non-finite prediction is no change; one beyond the limit is cut
      if (!(fabs(d) < HUGE_VAL)) d = 0.0;
      if (lim.v[c] > 0.0) d = fmin(fmax(d, -lim.v[c]), lim.v[c]);
      inc[c * hw + q] = d;
    }
    const double ar = area[nn_ix(ni, (int)(q % W), (int)(q / W))] * (double)mt;
    v[0] = ar * inc[q];
    v[1] = ar * inc[hw + q];
    v[2] = ar;
  }
  for (int c = 0; c < 3; ++c) s[c][threadIdx.x] = v[c];
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st)
      for (int c = 0; c < 3; ++c) s[c][threadIdx.x] += s[c][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int c = 0; c < 3; ++c) part[3 * blockIdx.x + c] = s[c][0];
}

// the weighted sea means of the two free-surface increments (every block
// folds k_pred's partials itself, in the same order), removed from them
__global__ void __launch_bounds__(256) k_pred_fix(long hw, const double* part, int nb, const float* stat,
                                                  double* inc) {
  double a[3];
  fold_parts<3>(part, nb, a);
  const double m0 = a[2] > 0.0 ? a[0] / a[2] : 0.0, m1 = a[2] > 0.0 ? a[1] / a[2] : 0.0;
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= hw) return;
  const double sea = (double)stat[q];
  inc[q] = inc[q] - m0 * sea;
  inc[hw + q] = inc[hw + q] - m1 * sea;
}

// write the end-of-cycle fields (mod_int.f90:276-287) from the increments,
// interior points and their periodic halo copies, after removing the two
// free-surface increments' weighted sea means as k_pred_fix does (the same
// fold of k_pred's partials and the same operations; inc is left as
// k_pred_fix leaves it)
__global__ void __launch_bounds__(256) k_apply(int ni, int H, int W, const double* part, int nb, const float* stat,
                                               double* inc, double* chc, double* chp, double* mch, double* mch2,
                                               double* acch, double* buc, double* bvc, double* bup, double* bvp) {
  double fm[3];
  fold_parts<3>(part, nb, fm);
  const double f0 = fm[2] > 0.0 ? fm[0] / fm[2] : 0.0, f1 = fm[2] > 0.0 ? fm[1] / fm[2] : 0.0;
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long hw = (long)H * W;
  if (q >= hw) return;
  const double sea = (double)stat[q];
  const double i0 = inc[q] - f0 * sea, i1 = inc[hw + q] - f1 * sea;
  inc[q] = i0;
  inc[hw + q] = i1;
  const int x = (int)(q % W), y = (int)(q / W);
  const long k = nn_ix(ni, x, y);
  const double c0 = chc[k];
  const double m2 = c0 + i0;
  const double m1 = c0 + i1;
  const double u = buc[k] + inc[2 * hw + q];
  const double v = bvc[k] + inc[3 * hw + q];
  const int i = x + 2;
  const int cols[2] = {i, i == W + 1 ? 1 : (i == 2 ? W + 2 : 0)};
  for (int cidx = 0; cidx < 2; ++cidx) {
    if (cols[cidx] == 0) continue;
    const long o = (long)(cols[cidx] - 1) + (long)ni * (y + 1);
    buc[o] = u;
    bvc[o] = v;
    bup[o] = u;
    bvp[o] = v;
    mch[o] = m1;
    mch2[o] = m2;
    chc[o] = m2;
    chp[o] = m2;
    acch[o] = acch[o] + m2;
  }
}

// root-mean-square error over sea of the prediction against the plain
// subcycle's result (err[c]), and of persistence, i.e. no change over the
// step (err[4 + c]); one block per channel, fixed order
// The check's RMS of the prediction's and of persistence's error against
// the plain cycle, per output over sea pixels: partial sums of 512 pixels
// (two per thread) in part[(3 * c + v) * nb + block], then one pass over
// the partials in block order, so the result does not depend on timing.
constexpr int kCheckPix = 512;
__global__ void k_check_part(int ni, int H, int W, const double* inc, const double* c0, const double* mch2,
                             const double* mch, const double* buc, const double* bvc, const float* stat,
                             double* part) {
  __shared__ double s[3][256];
  const int c = blockIdx.y;
  const long hw = (long)H * W;
  const int nb = gridDim.x;
  const double* truth = c == 0 ? mch2 : c == 1 ? mch : c == 2 ? buc : bvc;
  double a = 0.0, a0 = 0.0, n = 0.0;
  for (int h = 0; h < 2; ++h) {
    const long q = (long)blockIdx.x * kCheckPix + h * 256 + threadIdx.x;
    if (q >= hw) break;
    const int y = (int)(q / W), x = (int)(q - (long)y * W);
    const long k = nn_ix(ni, x, y);
    const double m = (double)stat[(c < 2 ? 0 : hw) + q];
    const double base = c0[c * hw + q];
    const double d = m * (base + inc[c * hw + q] - truth[k]);
    const double d0 = m * (base - truth[k]);
    a += d * d;
    a0 += d0 * d0;
    n += m;
  }
  s[0][threadIdx.x] = a;
  s[1][threadIdx.x] = a0;
  s[2][threadIdx.x] = n;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st)
      for (int v = 0; v < 3; ++v) s[v][threadIdx.x] += s[v][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x < 3) part[(3 * c + threadIdx.x) * nb + blockIdx.x] = s[threadIdx.x][0];
}
__global__ void k_check_final(int nb, const double* part, double* err) {
  __shared__ double s[3][256];
  const int c = blockIdx.x;
  for (int v = 0; v < 3; ++v) {
    double t = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) t += part[(3 * c + v) * nb + b];
    s[v][threadIdx.x] = t;
  }
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st)
      for (int v = 0; v < 3; ++v) s[v][threadIdx.x] += s[v][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    err[c] = s[2][0] > 0 ? sqrt(s[0][0] / s[2][0]) : 0.0;
    err[4 + c] = s[2][0] > 0 ? sqrt(s[1][0] / s[2][0]) : 0.0;
  }
}

// squared prediction error and squared target (persistence error) per used
// output channel, in physical units, summed over sea points of a batch
__global__ void k_eval(long np, long hw, const float* __restrict__ out, const float* __restrict__ tgt, NnNorm nm,
                       const float* __restrict__ stat, const float* __restrict__ sscale, const int* __restrict__ idx,
                       double* part) {
  __shared__ double s[9][256];
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (p < np) {
    const long q = p % hw;
    const double sc = floored(sscale[idx[p / hw]], nm);
    const float mt = stat[q], mu = stat[hw + q];
    for (int c = 0; c < kOutUsed; ++c) {
      const double m = c < 2 ? mt : mu;
      const double d = m * ((double)out[p * kOutC + c] - (double)tgt[p * kOutUsed + c]) * nm.out_std[c] * sc;
      const double t = m * (double)tgt[p * kOutUsed + c] * nm.out_std[c] * sc;
      v[c] = d * d;
      v[4 + c] = t * t;
    }
    v[8] = mt;
  }
  for (int c = 0; c < 9; ++c) s[c][threadIdx.x] = v[c];
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st)
      for (int c = 0; c < 9; ++c) s[c][threadIdx.x] += s[c][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int c = 0; c < 9; ++c) part[9 * blockIdx.x + c] = s[c][0];
}

__global__ void k_save_base(int ni, int H, int W, const double* chc, const double* buc, const double* bvc,
                            double* base) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long hw = (long)H * W;
  if (q >= hw) return;
  const long k = nn_ix(ni, (int)(q % W), (int)(q / W));
  base[q] = chc[k];
  base[hw + q] = chc[k];
  base[2 * hw + q] = buc[k];
  base[3 * hw + q] = bvc[k];
}

inline unsigned nblk(long n, int b = 256) { return (unsigned)((n + b - 1) / b); }

#define NN_LAUNCH(kern, grid, block, ...)                                        \
  do {                                                                           \
    hipLaunchKernelGGL(kern, grid, block, 0, st, __VA_ARGS__);                   \
    hipError_t e_ = hipGetLastError();                                           \
    if (e_ != hipSuccess) return std::string(#kern ": ") + hipGetErrorString(e_); \
  } while (0)

}  // namespace

// ---------------------------------------------------------------------------
struct NnLayer {
  int cin = 0, cout = 0;
  bool relu = true;
  float *w32 = nullptr, *b = nullptr;     // master weights [co][9][ci], bias [co]
  float *mw = nullptr, *vw = nullptr, *mb = nullptr, *vb = nullptr;  // Adam moments
  float *gw = nullptr, *gb = nullptr;     // gradients
  bf16_t *w = nullptr, *wt = nullptr;     // BF16 weights and transpose
  long nw() const { return 9L * cin * cout; }
};

// pointwise bound of a predicted increment, in units of the sea RMS of one
// step's change when the network took over
constexpr double kLimitAnchors = 50.0;

struct Nn {
  int H = 0, W = 0, ni = 0;
  int cap = 0, nrec = 0, cursor = 0;
  bool recording = false;
  int use_every = 0;  // surrogate in use: plain subcycle every use_every-th step
  double tol = 0.0;
  // after a failed check: plain steps before the network resumes (0: it
  // does not), and the last step of the current hold
  int resume_after = 0;
  long long hold_until = 0;
  // the network's first step after a hold is a check: a network that failed
  // is never trusted again unchecked
  bool check_next = false;
  long long step_count = 0, surrogate_steps = 0, plain_steps = 0, checks = 0, fallbacks = 0;
  double last_err[4] = {0, 0, 0, 0}, max_err[4] = {0, 0, 0, 0}, last_persist[4] = {0, 0, 0, 0};
  // persistence RMSE at the network's first check: the size of one step's
  // change when the network first took over, which the relative bound and
  // the pointwise limit of increments use.  It is kept across pcom_nn_use
  // calls: anchored afresh where the network has just gone wrong, it would
  // measure the damage instead of the flow.
  double anchor[4] = {0, 0, 0, 0};
  bool anchored = false;
  bool normalised = false;
  int adam_t = 0;              // Adam steps taken
  double* misc = nullptr;      // [16]: sea means, check errors
  NnNorm norm{};
  NnLayer L[kLayers];
  float* store_in = nullptr;   // [cap][kDyn][HW] raw inputs
  float* store_tg = nullptr;   // [cap][4][HW] increments
  float* store_sc = nullptr;   // [cap] scale of each snapshot
  float* stat = nullptr;       // [kStat][HW] sea T, sea U, Coriolis / scale, prh / scale
  float* live = nullptr;       // [kDyn][HW] raw inputs of the live state
  double *chpred = nullptr, *u1 = nullptr, *v1 = nullptr, *sch1 = nullptr;  // [n2] first substep
  // perturbed snapshots: fraction of the recorded ones, noise amplitudes
  // (ch, m/s), seed; state backup [9][n2] and noise [2][3][HW]
  double aug_fraction = 0.0, aug_amp_ch = 0.0, aug_amp_u = 0.0;
  unsigned long long aug_seed = 0, aug_count = 0;
  double* bak = nullptr;
  float* noise = nullptr;
  int batch_max = 0;
  bf16_t* act[kLayers + 1] = {};  // layer inputs / outputs (BF16), act[0] = x
  float* out = nullptr;        // last layer output (FP32) [batch][HW][32]
  float* tgt = nullptr;        // [batch][HW][4]
  bf16_t *g0 = nullptr, *g1 = nullptr;  // gradient ping-pong (BF16, 64 channels)
  float* wpart = nullptr;      // weight-gradient partials
  long wchunk = 0;
  int wnchunk = 0;
  double* dpart = nullptr;     // loss partials, moments, prediction sums
  double* inc = nullptr;       // [4][HW] predicted increments
  double* base = nullptr;      // [4][HW] state before a checked subcycle
  double* massw = nullptr;     // [n2] area * prh of the T points (mass weight)
  int* idx = nullptr;          // minibatch indices
  std::vector<void*> allocs;
};

namespace {

std::string nn_alloc(Nn& n, void** p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes ? bytes : 8);
  if (e != hipSuccess) return std::string("hipMalloc (surrogate): ") + hipGetErrorString(e);
  n.allocs.push_back(*p);
  return "";
}

template <typename T>
std::string nn_make(Nn& n, T** p, size_t count) {
  void* v = nullptr;
  std::string e = nn_alloc(n, &v, count * sizeof(T));
  if (!e.empty()) return e;
  hipError_t h = hipMemset(v, 0, count * sizeof(T));
  if (h != hipSuccess) return std::string("hipMemset (surrogate): ") + hipGetErrorString(h);
  *p = static_cast<T*>(v);
  return "";
}

std::string nn_batch_buffers(Nn& n, int batch) {
  if (batch <= n.batch_max) return "";
  const long np = (long)batch * n.H * n.W;
  std::string e;
  for (int l = 0; l <= kLayers; ++l) {
    const int c = l == 0 ? kInC : (l == kLayers ? kOutC : kHid);
    if (!(e = nn_make(n, &n.act[l], (size_t)np * c)).empty()) return e;
  }
  if (!(e = nn_make(n, &n.out, (size_t)np * kOutC)).empty()) return e;
  if (!(e = nn_make(n, &n.tgt, (size_t)np * kOutUsed)).empty()) return e;
  if (!(e = nn_make(n, &n.g0, (size_t)np * kHid)).empty()) return e;
  if (!(e = nn_make(n, &n.g1, (size_t)np * kHid)).empty()) return e;
  n.wchunk = 4096;
  n.wnchunk = (int)((np + n.wchunk - 1) / n.wchunk);
  if (!(e = nn_make(n, &n.wpart, (size_t)n.wnchunk * 9 * kHid * kHid)).empty()) return e;
  // block partials: up to 9 per block of 256 pixels (k_eval), with room
  if (!(e = nn_make(n, &n.dpart, (size_t)(9 * nblk(np) + 64))).empty()) return e;
  if (!(e = nn_make(n, &n.idx, (size_t)batch)).empty()) return e;
  n.batch_max = batch;
  return "";
}

// forward pass over `batch` images already packed into act[0]: the layers'
// weights staged in LDS, workgroups of NW waves taking PB pixel blocks per
// wave at a time, at most kConvGroups workgroups
constexpr long kConvGroups = 512;
template <int KIN, int KOUT, int MODE, int PB = 2, int NW = kConvLdsWaves>
std::string conv_lds(hipStream_t st, const Geo& g, const bf16_t* in, const bf16_t* w, const float* b,
                     const bf16_t* post, void* out) {
  const auto kern = k_conv_lds<KIN, KOUT, MODE, PB, NW>;
  static bool attr = false;  // once per process: the attribute holds for every device
  if (!attr) {
    hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)conv_lds_bytes<KIN, KOUT>());
    if (e != hipSuccess) return std::string("hipFuncSetAttribute k_conv_lds: ") + hipGetErrorString(e);
    attr = true;
  }
  const long pblocks = g.npix() / 16;
  const unsigned blocks = (unsigned)std::min(kConvGroups, (pblocks + NW * PB - 1) / (NW * PB));
  const size_t lds = conv_lds_bytes<KIN, KOUT>();
  hipLaunchKernelGGL(kern, dim3(blocks), dim3(64 * NW), lds, st, g, in, w, b, post, out);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return std::string("k_conv_lds: ") + hipGetErrorString(e);
  return "";
}

template <int PB, int NW = kConvLdsWaves>
std::string nn_forward_pb(Nn& n, hipStream_t st, int batch) {
  Geo g{batch, n.H, n.W, layer_dil(0)};
  std::string e =
      conv_lds<kInC, kHid, kFwdRelu, PB, NW>(st, g, n.act[0], n.L[0].w, n.L[0].b, nullptr, (void*)n.act[1]);
  for (int l = 1; l < kLayers - 1 && e.empty(); ++l) {
    g.dil = layer_dil(l);
    e = conv_lds<kHid, kHid, kFwdRelu, PB, NW>(st, g, n.act[l], n.L[l].w, n.L[l].b, nullptr, (void*)n.act[l + 1]);
  }
  if (!e.empty()) return e;
  g.dil = layer_dil(kLayers - 1);
  return conv_lds<kHid, kOutC, kFwdLinear, PB, NW>(st, g, n.act[kLayers - 1], n.L[kLayers - 1].w,
                                                   n.L[kLayers - 1].b, nullptr, (void*)n.out);
}

// one image: a pixel block per wave in workgroups of 16 waves (one per CU
// at 1 degree: half as many stagings of the layer's weights, each thread
// loading half as many of them); minibatches: two blocks per wave, 8 waves
constexpr int kConvOneWaves = 16;
std::string nn_forward(Nn& n, hipStream_t st, int batch) {
  return batch == 1 ? nn_forward_pb<1, kConvOneWaves>(n, st, batch) : nn_forward_pb<2>(n, st, batch);
}

// backward pass from dpre of the last layer (in act[kLayers], BF16)
std::string nn_backward(Nn& n, hipStream_t st, int batch) {
  Geo g{batch, n.H, n.W, 1};
  const long np = g.npix();
  const int nch = (int)((np + n.wchunk - 1) / n.wchunk);
  const int nbias = (int)((np + kBiasChunk - 1) / kBiasChunk);
  bf16_t* dpre = n.act[kLayers];  // d loss / d (last pre-activation), 32 channels
  for (int l = kLayers - 1; l >= 0; --l) {
    NnLayer& L = n.L[l];
    g.dil = layer_dil(l);
    // weight and bias gradients
    if (L.cin == kInC && L.cout == kHid)
      NN_LAUNCH((k_conv_bwd_w<kInC, kHid>), dim3(nch, 9), dim3(256), g, n.act[l], dpre, n.wpart, n.wchunk);
    else if (L.cin == kHid && L.cout == kHid)
      NN_LAUNCH((k_conv_bwd_w<kHid, kHid>), dim3(nch, 9), dim3(256), g, n.act[l], dpre, n.wpart, n.wchunk);
    else
      NN_LAUNCH((k_conv_bwd_w<kHid, kOutC>), dim3(nch, 9), dim3(256), g, n.act[l], dpre, n.wpart, n.wchunk);
    NN_LAUNCH(k_reduce_w, dim3(nblk(L.nw())), dim3(256), n.wpart, nch, L.cout, L.cin, L.gw);
    NN_LAUNCH(k_bias_part, dim3(nbias), dim3(256), g, dpre, L.cout, (float*)n.wpart);
    NN_LAUNCH(k_bias_reduce, dim3(L.cout), dim3(256), (const float*)n.wpart, nbias, L.cout, L.gb);
    if (l == 0) break;
    // gradient of this layer's input = previous layer's post-activation
    bf16_t* dpost = (dpre == n.g0) ? n.g1 : n.g0;
    // through the ReLU of layer l-1, whose output is act[l]
    std::string e = L.cout == kOutC
                        ? conv_lds<kOutC, kHid, kBwdData>(st, g, dpre, L.wt, nullptr, n.act[l], (void*)dpost)
                        : conv_lds<kHid, kHid, kBwdData>(st, g, dpre, L.wt, nullptr, n.act[l], (void*)dpost);
    if (!e.empty()) return e;
    dpre = dpost;
  }
  return "";
}

std::string nn_loss(Nn& n, hipStream_t st, int batch, double* loss_dev) {
  const long np = (long)batch * n.H * n.W;
  // masked count: sum of masks over used channels (computed from the input)
  const float inv = 1.0f / (float)(np * kOutUsed);
  NN_LAUNCH(k_loss, dim3(nblk(np)), dim3(256), np, n.out, n.tgt, n.act[0], inv, n.act[kLayers], n.dpart);
  NN_LAUNCH(k_sum_parts, dim3(1), dim3(256), n.dpart, (int)nblk(np), loss_dev);
  return "";
}

std::string nn_adam(Nn& n, hipStream_t st, float lr, int t) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float bc1 = 1.f - std::pow(b1, (float)t), bc2 = 1.f - std::pow(b2, (float)t);
  for (int l = 0; l < kLayers; ++l) {
    NnLayer& L = n.L[l];
    NN_LAUNCH(k_adam, dim3(nblk(L.nw())), dim3(256), L.nw(), L.w32, L.mw, L.vw, L.gw, lr, b1, b2, eps, bc1, bc2);
    NN_LAUNCH(k_adam, dim3(1), dim3(256), (long)L.cout, L.b, L.mb, L.vb, L.gb, lr, b1, b2, eps, bc1, bc2);
    NN_LAUNCH(k_w_bf16, dim3(nblk(L.nw())), dim3(256), L.cout, L.cin, L.w32, L.w, L.wt);
  }
  return "";
}

}  // namespace

// ---------------------------------------------------------------------------
// entry points used by api.hip

std::string nn_create(Ctx& c, int capacity, unsigned long long seed) {
  if (c.nn) return "a surrogate is already attached to this context";
  const int H = c.nj - 2, W = c.ni - 2;
  if (((long)H * W) % 16 != 0) return "the surrogate needs glo_ni*glo_nj to be a multiple of 16";
  Nn* n = new Nn();
  c.nn = n;
  n->H = H;
  n->W = W;
  n->ni = c.ni;
  n->cap = capacity;
  const long hw = (long)H * W;
  std::string e;
  const int dims[kLayers][2] = {{kInC, kHid}, {kHid, kHid}, {kHid, kHid}, {kHid, kHid}, {kHid, kHid}, {kHid, kOutC}};
  std::mt19937_64 rng(seed);
  for (int l = 0; l < kLayers; ++l) {
    NnLayer& L = n->L[l];
    L.cin = dims[l][0];
    L.cout = dims[l][1];
    L.relu = l < kLayers - 1;
    const long nw = L.nw();
    float** f32[] = {&L.w32, &L.mw, &L.vw, &L.gw};
    for (float** p : f32)
      if (!(e = nn_make(*n, p, (size_t)nw)).empty()) return e;
    float** fb[] = {&L.b, &L.mb, &L.vb, &L.gb};
    for (float** p : fb)
      if (!(e = nn_make(*n, p, (size_t)L.cout)).empty()) return e;
    if (!(e = nn_make(*n, &L.w, (size_t)nw)).empty()) return e;
    if (!(e = nn_make(*n, &L.wt, (size_t)nw)).empty()) return e;
    // He initialisation over the used input channels; padded channels zero
    const int cin_used = l == 0 ? kInUsed : L.cin;
    const int cout_used = l == kLayers - 1 ? kOutUsed : L.cout;
    std::normal_distribution<float> nd(0.f, std::sqrt(2.f / (9.f * cin_used)) * (l == kLayers - 1 ? 0.1f : 1.f));
    std::vector<float> hw32((size_t)nw, 0.f);
    for (int co = 0; co < cout_used; ++co)
      for (int k = 0; k < 9; ++k)
        for (int ci = 0; ci < cin_used; ++ci) hw32[((size_t)co * 9 + k) * L.cin + ci] = nd(rng);
    if (hipMemcpy(L.w32, hw32.data(), nw * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
      return "upload surrogate weights";
  }
  // static channels: sea masks, Coriolis and bottom pressure of the interior
  std::vector<float> stat((size_t)kStat * hw);
  double cmax = 0.0, pmax = 0.0;
  for (long q = 0; q < hw; ++q) {
    const size_t k = (size_t)(q % W) + 1 + (size_t)c.ni * (q / W + 1);
    cmax = std::max(cmax, std::fabs(c.hs.cor[k]));
    if (c.hs.lev_t[k] > 0) pmax = std::max(pmax, c.hs.prh_t[k]);
  }
  n->norm.cor_scale = (float)(cmax > 0 ? cmax : 1.0);
  n->norm.prh_scale = (float)(pmax > 0 ? pmax : 1.0);
  for (long q = 0; q < hw; ++q) {
    const size_t k = (size_t)(q % W) + 1 + (size_t)c.ni * (q / W + 1);
    stat[q] = c.hs.lev_t[k] > 0 ? 1.f : 0.f;
    stat[hw + q] = c.hs.lev_u[k] > 0 ? 1.f : 0.f;
    stat[2 * hw + q] = (float)(c.hs.cor[k] / n->norm.cor_scale);
    stat[3 * hw + q] = c.hs.lev_t[k] > 0 ? (float)(c.hs.prh_t[k] / n->norm.prh_scale) : 0.f;
  }
  {
    std::vector<double> w(c.n2);
    for (size_t k = 0; k < c.n2; ++k) w[k] = c.hs.rh[1][k] * c.hs.dx1[1][k] * c.hs.dx2[1][k] * c.hs.prh_t[k];
    if (!(e = nn_make(*n, &n->massw, c.n2)).empty()) return e;
    if (hipMemcpy(n->massw, w.data(), c.n2 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
      return "upload surrogate mass weights";
  }
  if (!(e = nn_make(*n, &n->stat, (size_t)kStat * hw)).empty()) return e;
  if (hipMemcpy(n->stat, stat.data(), stat.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
    return "upload surrogate statics";
  double** sub1[] = {&n->chpred, &n->u1, &n->v1, &n->sch1};
  for (double** q : sub1)
    if (!(e = nn_make(*n, q, c.n2)).empty()) return e;
  if (!(e = nn_make(*n, &n->live, (size_t)kDyn * hw)).empty()) return e;
  if (!(e = nn_make(*n, &n->bak, (size_t)9 * c.n2)).empty()) return e;
  if (!(e = nn_make(*n, &n->noise, (size_t)6 * hw)).empty()) return e;
  if (capacity > 0) {
    if (!(e = nn_make(*n, &n->store_in, (size_t)capacity * kDyn * hw)).empty()) return e;
    if (!(e = nn_make(*n, &n->store_tg, (size_t)capacity * kOutUsed * hw)).empty()) return e;
    if (!(e = nn_make(*n, &n->store_sc, (size_t)capacity)).empty()) return e;
  }
  if (!(e = nn_make(*n, &n->inc, (size_t)4 * hw)).empty()) return e;
  if (!(e = nn_make(*n, &n->misc, (size_t)24)).empty()) return e;
  if (!(e = nn_make(*n, &n->base, (size_t)4 * hw)).empty()) return e;
  if (!(e = nn_batch_buffers(*n, 1)).empty()) return e;
  hipStream_t st = c.stream;
  for (int l = 0; l < kLayers; ++l)
    NN_LAUNCH(k_w_bf16, dim3(nblk(n->L[l].nw())), dim3(256), n->L[l].cout, n->L[l].cin, n->L[l].w32, n->L[l].w,
              n->L[l].wt);
  if (hipStreamSynchronize(st) != hipSuccess) return "surrogate set-up sync";
  return "";
}

void nn_destroy(Ctx& c) {
  if (!c.nn) return;
  for (void* p : c.nn->allocs) (void)hipFree(p);
  delete c.nn;
  c.nn = nullptr;
}

void nn_set_record(Ctx& c, bool on) {
  if (c.nn) c.nn->recording = on && c.nn->cap > 0;
}

std::string nn_set_augment(Ctx& c, double fraction, double amp_ch, double amp_u, unsigned long long seed) {
  if (!c.nn) return "no surrogate attached";
  if (!(fraction >= 0.0 && fraction <= 1.0) || !(amp_ch >= 0.0) || !(amp_u >= 0.0)) return "bad augmentation";
  c.nn->aug_fraction = fraction;
  c.nn->aug_amp_ch = amp_ch;
  c.nn->aug_amp_u = amp_u;
  c.nn->aug_seed = seed;
  c.nn->aug_count = 0;
  return "";
}

std::string nn_set_use(Ctx& c, int every, double tol) {
  if (!c.nn) return "no surrogate attached";
  if (every > 0 && !c.nn->normalised) return "the surrogate is untrained (no normalisation statistics)";
  c.nn->use_every = every;
  c.nn->tol = tol;
  c.nn->hold_until = 0;
  // a network not yet anchored starts with a check, so that no step of it
  // runs without the pointwise limit
  c.nn->check_next = every > 0 && !c.nn->anchored;
  return "";
}

std::string nn_set_resume(Ctx& c, int plain_steps) {
  if (!c.nn) return "no surrogate attached";
  c.nn->resume_after = plain_steps;
  return "";
}

namespace {
// the raw inputs of the live state (the first substep's increment and the
// fields the subcycle's operator depends on) into `raw`, and its scale into
// misc[20] (and *keep when recording)
std::string nn_raw_inputs(Ctx& c, float* raw, float* keep, bool final = true) {
  Nn* n = c.nn;
  hipStream_t st = c.stream;
  const Fields2& h = c.f2;
  const long hw = (long)n->H * n->W;
  const double bt = (double)c.dtbt;
  NN_LAUNCH(k_sub1_pred, dim3(nblk((long)c.n2)), dim3(256), c.gd, h.chp, h.bup, h.bvp, h.chc, bt, n->chpred);
  Sub1In a;
  a.bup = h.bup; a.bvp = h.bvp; a.but = h.but; a.bvt = h.bvt;
  a.ghx = h.graphihx; a.ghy = h.graphihy; a.cor = c.gd.cor;
  for (int q = 0; q < 4; ++q) {
    a.bphi[q] = h.bphi[q];
    a.bgra[q] = h.bgra[q];
  }
  NN_LAUNCH(k_sub1_mom, dim3(nblk((long)c.n2)), dim3(256), c.gd, a, n->chpred, bt, n->u1, n->v1, n->sch1);
  RawIn r;
  r.chp = h.chp; r.bup = h.bup; r.bvp = h.bvp; r.u1 = n->u1; r.v1 = n->v1; r.sch1 = n->sch1; r.am1 = h.am1;
  for (int q = 0; q < 4; ++q) {
    r.bphi[q] = h.bphi[q];
    r.bgra[q] = h.bgra[q];
  }
  NN_LAUNCH(k_sub1_raw, dim3(nblk(hw)), dim3(256), c.gd, n->H, n->W, r, bt, raw, n->stat, n->dpart);
  // (the live path folds the partials in k_pack_live instead)
  if (final) NN_LAUNCH(k_scale_final, dim3(1), dim3(256), n->dpart, (int)nblk(hw), n->misc + 20, keep);
  return "";
}

// the network's increments for the live state into n->inc, mass-corrected
// (with fix false the correction is left to k_apply, which makes it itself)
std::string nn_predict(Ctx& c, bool fix = true) {
  Nn* n = c.nn;
  hipStream_t st = c.stream;
  const long hw = (long)n->H * n->W;
  std::string e = nn_raw_inputs(c, n->live, nullptr, false);
  if (!e.empty()) return e;
  NN_LAUNCH(k_pack_live, dim3(nblk(hw)), dim3(256), n->H, n->W, n->live, n->dpart, (int)nblk(hw), n->misc + 20, n->norm,
            n->stat, n->act[0]);
  e = nn_forward(*n, st, 1);
  if (!e.empty()) return e;
  // once anchored, increments are bounded pointwise by kLimitAnchors times
  // the sea RMS of one step's change at the first check
  NnLimit lim{};
  if (n->anchored)
    for (int q = 0; q < kOutUsed; ++q) lim.v[q] = kLimitAnchors * n->anchor[q];
  NN_LAUNCH(k_pred, dim3(nblk(hw)), dim3(256), n->ni, n->H, n->W, n->out, n->norm, n->misc + 20, n->stat, n->massw,
            lim, n->inc, n->dpart);
  if (fix) NN_LAUNCH(k_pred_fix, dim3(nblk(hw)), dim3(256), hw, n->dpart, (int)nblk(hw), n->stat, n->inc);
  return "";
}
}  // namespace

namespace {
// with probability aug_fraction, two extra recorded snapshots: the state
// perturbed by seeded noise (random smoothness, amplitudes log-uniform in
// [0.02, 1] times aug_amp_*, on ch, on (u, v) or on both) and the state its
// plain cycle leaves, each with its plain cycle; then the state is
// restored and the step itself follows unperturbed
std::string nn_record_perturbed(Ctx& c, std::string (*plain)(Ctx&)) {
  Nn* n = c.nn;
  hipStream_t st = c.stream;
  Fields2& h = c.f2;
  const long hw = (long)n->H * n->W;
  std::mt19937_64 rng(n->aug_seed + 7919ull * (n->aug_count++));
  std::uniform_real_distribution<double> u01(0.0, 1.0);
  if (!(u01(rng) < n->aug_fraction)) return "";
  static const int passes[5] = {0, 1, 3, 8, 20};
  const int npass = passes[(int)(u01(rng) * 5.0) % 5];
  const double fc = std::exp(std::log(0.02) + u01(rng) * std::log(50.0));
  const double fu = std::exp(std::log(0.02) + u01(rng) * std::log(50.0));
  const int which = (int)(u01(rng) * 3.0) % 3;  // 0 both, 1 ch only, 2 (u, v) only
  const double amp_ch = which == 2 ? 0.0 : fc * n->aug_amp_ch, amp_u = which == 1 ? 0.0 : fu * n->aug_amp_u;
  double* fields[9] = {h.chc, h.chp, h.mch, h.mch2, h.acch, h.buc, h.bvc, h.bup, h.bvp};
  const size_t n2 = c.n2;
  for (int f = 0; f < 9; ++f)
    if (hipMemcpyAsync(n->bak + f * n2, fields[f], n2 * sizeof(double), hipMemcpyDeviceToDevice, st) != hipSuccess)
      return "surrogate state backup";
  float* nz = n->noise;
  float* tmp = n->noise + 3 * hw;
  NN_LAUNCH(k_noise, dim3(nblk(3 * hw)), dim3(256), n->H, n->W, rng(), nz);
  for (int k = 0; k < npass; ++k) {
    NN_LAUNCH(k_smooth, dim3(nblk(3 * hw)), dim3(256), n->H, n->W, nz, tmp);
    std::swap(nz, tmp);
  }
  NN_LAUNCH(k_noise_part, dim3(nblk(hw)), dim3(256), n->H, n->W, nz, n->stat, n->dpart);
  NN_LAUNCH(k_noise_final, dim3(1), dim3(256), n->dpart, (int)nblk(hw), n->misc + 22);
  NN_LAUNCH(k_perturb, dim3(nblk(hw)), dim3(256), n->ni, n->H, n->W, nz, n->misc + 22, amp_ch, amp_u, n->stat, h.chc,
            h.chp, h.buc, h.bvc, h.bup, h.bvp);
  // the perturbed state's cycle, then the cycle from what it left: its
  // averaging keeps the slow, nearly balanced part of the perturbation,
  // whose evolution the network must follow as closely as it damps the rest
  for (int pass = 0; pass < 2; ++pass) {
    const int slot = n->cursor;
    NN_LAUNCH(k_save_base, dim3(nblk(hw)), dim3(256), n->ni, n->H, n->W, h.chc, h.buc, h.bvc, n->base);
    std::string e = nn_raw_inputs(c, n->store_in + (long)slot * kDyn * hw, n->store_sc + slot);
    if (!e.empty()) return e;
    if (!(e = plain(c)).empty()) return e;
    NN_LAUNCH(k_rec_out, dim3(nblk(hw)), dim3(256), n->ni, n->H, n->W, h.mch2, h.mch, h.buc, h.bvc, n->base,
              n->store_tg + (long)slot * kOutUsed * hw);
    n->cursor = (n->cursor + 1) % n->cap;
    n->nrec = std::min(n->nrec + 1, n->cap);
  }
  for (int f = 0; f < 9; ++f)
    if (hipMemcpyAsync(fields[f], n->bak + f * n2, n2 * sizeof(double), hipMemcpyDeviceToDevice, st) != hipSuccess)
      return "surrogate state restore";
  return "";
}
}  // namespace

namespace {
// one recorded snapshot of the state the network is about to step: its
// inputs and what the plain cycle makes of them, after which the state is
// restored.  Snapshots of the states the network itself leads to teach it
// to correct the errors it made (its own inputs drift from those of plain
// runs as soon as it drives the model).
std::string nn_record_beside(Ctx& c, std::string (*plain)(Ctx&)) {
  Nn* n = c.nn;
  hipStream_t st = c.stream;
  Fields2& h = c.f2;
  const long hw = (long)n->H * n->W;
  double* fields[9] = {h.chc, h.chp, h.mch, h.mch2, h.acch, h.buc, h.bvc, h.bup, h.bvp};
  const size_t n2 = c.n2;
  for (int f = 0; f < 9; ++f)
    if (hipMemcpyAsync(n->bak + f * n2, fields[f], n2 * sizeof(double), hipMemcpyDeviceToDevice, st) != hipSuccess)
      return "surrogate state backup";
  const int slot = n->cursor;
  NN_LAUNCH(k_save_base, dim3(nblk(hw)), dim3(256), n->ni, n->H, n->W, h.chc, h.buc, h.bvc, n->base);
  std::string e = nn_raw_inputs(c, n->store_in + (long)slot * kDyn * hw, n->store_sc + slot);
  if (!e.empty()) return e;
  if (!(e = plain(c)).empty()) return e;
  NN_LAUNCH(k_rec_out, dim3(nblk(hw)), dim3(256), n->ni, n->H, n->W, h.mch2, h.mch, h.buc, h.bvc, n->base,
            n->store_tg + (long)slot * kOutUsed * hw);
  n->cursor = (n->cursor + 1) % n->cap;
  n->nrec = std::min(n->nrec + 1, n->cap);
  for (int f = 0; f < 9; ++f)
    if (hipMemcpyAsync(fields[f], n->bak + f * n2, n2 * sizeof(double), hipMemcpyDeviceToDevice, st) != hipSuccess)
      return "surrogate state restore";
  return "";
}
}  // namespace

// the barotropic part of one step: plain subcycle, or the network's
// prediction; records snapshots of plain subcycles when recording, and of
// the plain cycle beside the network's steps when recording in use.
std::string nn_trop_step(Ctx& c, std::string (*plain)(Ctx&)) {
  Nn* n = c.nn;
  hipStream_t st = c.stream;
  Fields2& h = c.f2;
  const long hw = (long)n->H * n->W;
  n->step_count += 1;
  // held: plain steps after a failed check, before the network resumes
  const bool use = n->use_every > 0 && n->step_count > n->hold_until;
  const bool check = use && (n->check_next || n->step_count % n->use_every == 0);
  if (c.timing) (void)hipEventRecord(c.ev[1], st);
  if (use && !check) {
    if (n->recording) {
      std::string e = nn_record_beside(c, plain);
      if (!e.empty()) return e;
    }
    std::string e = nn_predict(c, false);
    if (!e.empty()) return e;
    NN_LAUNCH(k_apply, dim3(nblk(hw)), dim3(256), n->ni, n->H, n->W, n->dpart, (int)nblk(hw), n->stat, n->inc, h.chc,
              h.chp, h.mch, h.mch2, h.acch, h.buc, h.bvc, h.bup, h.bvp);
    n->surrogate_steps += 1;
    return "";
  }
  if (n->recording && n->aug_fraction > 0.0) {
    std::string e = nn_record_perturbed(c, plain);
    if (!e.empty()) return e;
  }
  const int slot = n->cursor;
  if (n->recording || check)
    NN_LAUNCH(k_save_base, dim3(nblk(hw)), dim3(256), n->ni, n->H, n->W, h.chc, h.buc, h.bvc, n->base);
  if (n->recording) {
    std::string e = nn_raw_inputs(c, n->store_in + (long)slot * kDyn * hw, n->store_sc + slot);
    if (!e.empty()) return e;
  }
  if (check) {
    // the prediction the network would have made, checked by the plain cycle
    std::string e = nn_predict(c);
    if (!e.empty()) return e;
  }
  std::string e = plain(c);
  if (!e.empty()) return e;
  n->plain_steps += 1;
  if (n->recording) {
    NN_LAUNCH(k_rec_out, dim3(nblk(hw)), dim3(256), n->ni, n->H, n->W, h.mch2, h.mch, h.buc, h.bvc, n->base,
              n->store_tg + (long)slot * kOutUsed * hw);
    n->cursor = (n->cursor + 1) % n->cap;
    n->nrec = std::min(n->nrec + 1, n->cap);
  }
  if (check) {
    double* err = n->misc + 12;  // 8 values, read back below
    // the partials fit the prediction's (12 per 512 pixels against 9 per
    // 256), which k_pred_fix has consumed
    const int nb = (int)nblk(hw, kCheckPix);
    NN_LAUNCH(k_check_part, dim3(nb, 4), dim3(256), n->ni, n->H, n->W, n->inc, n->base, h.mch2, h.mch, h.buc, h.bvc,
              n->stat, n->dpart);
    NN_LAUNCH(k_check_final, dim3(4), dim3(256), nb, n->dpart, err);
    double he[8];
    if (hipMemcpyAsync(he, err, sizeof he, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return "surrogate check read-back";
    n->checks += 1;
    n->check_next = false;
    if (!n->anchored) {
      for (int q = 0; q < 4; ++q) n->anchor[q] = he[4 + q];
      n->anchored = true;
    }
    bool bad = false;
    for (int q = 0; q < 4; ++q) {
      n->last_persist[q] = he[4 + q];
      n->last_err[q] = he[q];
      n->max_err[q] = std::max(n->max_err[q], he[q]);
      // tol > 0 bounds the error, tol < 0 bounds it relative to persistence,
      // now and at the first check (a run drifting into noise, where both
      // grow together, fails the second); NaN counts as a failed check
      const double bound = n->tol > 0.0 ? n->tol : -n->tol * std::min(he[4 + q], n->anchor[q]);
      if (!(he[q] <= bound)) bad = true;
    }
    if (bad && n->tol != 0.0) {
      n->fallbacks += 1;
      if (n->resume_after > 0) {
        n->hold_until = n->step_count + n->resume_after;  // the plain subcycle takes over for a while
        n->check_next = true;
      } else
        n->use_every = 0;  // ... or for good
    }
  }
  return "";
}

// normalisation statistics from the recorded snapshots
std::string nn_normalise(Ctx& c) {
  Nn* n = c.nn;
  hipStream_t st = c.stream;
  if (n->nrec < 1) return "no recorded snapshots";
  const long hw = (long)n->H * n->W;
  double mom[2 * (kDyn + kOutUsed)];
  // the scale floor: a fraction of the recorded snapshots' median scale
  {
    std::vector<float> sc(n->nrec);
    if (hipStreamSynchronize(st) != hipSuccess ||
        hipMemcpy(sc.data(), n->store_sc, sc.size() * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
      return "snapshot scales read-back";
    std::nth_element(sc.begin(), sc.begin() + sc.size() / 2, sc.end());
    const float med = sc[sc.size() / 2];
    n->norm.s_floor = med > 0.f && std::isfinite(med) ? kFloorOfMedian * med : 0.f;
  }
  NN_LAUNCH(k_moments, dim3(kDyn), dim3(256), n->store_in, (long)kDyn * hw, hw, n->nrec, n->store_sc, kScaled,
            n->norm.s_floor, n->dpart);
  NN_LAUNCH(k_moments, dim3(kOutUsed), dim3(256), n->store_tg, (long)kOutUsed * hw, hw, n->nrec, n->store_sc,
            kOutUsed, n->norm.s_floor, n->dpart + 2 * kDyn);
  // the null-stream copy does not wait for the context's (non-blocking) stream
  if (hipStreamSynchronize(st) != hipSuccess ||
      hipMemcpy(mom, n->dpart, sizeof mom, hipMemcpyDeviceToHost) != hipSuccess)
    return "moments read-back";
  const double cnt = (double)n->nrec * hw;
  for (int q = 0; q < kDyn + kOutUsed; ++q) {
    const double mean = mom[2 * q] / cnt;
    double var = mom[2 * q + 1] / cnt - mean * mean;
    double sd = var > 0 ? std::sqrt(var) : 0.0;
    if (!(sd > 1e-12)) sd = 1.0;
    if (q < kScaled) {  // r / s: a root-mean-square scale, no offset (the map is linear in r)
      const double ms = mom[2 * q + 1] / cnt;
      n->norm.in_mean[q] = 0.f;
      n->norm.in_std[q] = (float)(ms > 0 ? std::sqrt(ms) : 1.0);
    } else if (q < kDyn) {
      n->norm.in_mean[q] = (float)mean;
      n->norm.in_std[q] = (float)sd;
    } else {
      n->norm.out_std[q - kDyn] = (float)std::sqrt(mom[2 * q + 1] / cnt > 0 ? mom[2 * q + 1] / cnt : 1.0);
    }
  }
  n->normalised = true;
  return "";
}

std::string nn_train(Ctx& c, int iters, int batch, double lr, unsigned long long seed, double* loss) {
  Nn* n = c.nn;
  hipStream_t st = c.stream;
  if (n->nrec < 1) return "no recorded snapshots to train on";
  std::string e;
  if (!n->normalised && !(e = nn_normalise(c)).empty()) return e;
  if (!(e = nn_batch_buffers(*n, batch)).empty()) return e;
  const long hw = (long)n->H * n->W;
  std::mt19937_64 rng(seed);
  std::uniform_int_distribution<int> pick(0, n->nrec - 1);
  std::vector<int> idx(batch);
  double* ldev = n->misc + 8;
  for (int it = 0; it < iters; ++it) {
    for (int b = 0; b < batch; ++b) idx[b] = pick(rng);
    // synchronous: the host vector is refilled next iteration
    if (hipStreamSynchronize(st) != hipSuccess ||
        hipMemcpy(n->idx, idx.data(), batch * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
      return "minibatch upload";
    NN_LAUNCH(k_pack, dim3(nblk(hw * batch)), dim3(256), n->H, n->W, n->store_in, n->store_tg, n->store_sc, n->idx,
              batch, n->norm, n->stat, n->act[0], n->tgt);
    if (!(e = nn_forward(*n, st, batch)).empty()) return e;
    if (!(e = nn_loss(*n, st, batch, ldev)).empty()) return e;
    if (!(e = nn_backward(*n, st, batch)).empty()) return e;
    // cosine decay of the learning rate over this call's iterations
    const double lr_it = lr * 0.5 * (1.0 + std::cos(3.14159265358979 * it / (double)iters));
    if (!(e = nn_adam(*n, st, (float)lr_it, ++n->adam_t)).empty()) return e;
    if (loss) {
      double l;
      if (hipMemcpyAsync(&l, ldev, sizeof l, hipMemcpyDeviceToHost, st) != hipSuccess ||
          hipStreamSynchronize(st) != hipSuccess)
        return "loss read-back";
      loss[it] = l / (double)((long)batch * hw * kOutUsed);
    }
  }
  if (hipStreamSynchronize(st) != hipSuccess) return "training sync";
  // a diverged training is an error, not a network that predicts no change
  for (int l = 0; l < kLayers; ++l) {
    std::vector<float> w(n->L[l].nw());
    if (hipMemcpy(w.data(), n->L[l].w32, w.size() * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
      return "weights read-back";
    for (float v : w)
      if (!std::isfinite(v)) return "training diverged: non-finite weights";
  }
  return "";
}

// root-mean-square error over sea of the network and of persistence on
// recorded snapshots [first, first+count) of the ring, in physical units:
// out[0..3] network (mch2, mch, ubt, vbt increments), out[4..7] persistence
std::string nn_eval(Ctx& c, int first, int count, double* res) {
  Nn* n = c.nn;
  hipStream_t st = c.stream;
  if (!n->normalised) return "the surrogate is untrained";
  if (first < 0 || count < 1 || first + count > n->nrec) return "snapshot range outside the recorded ones";
  std::string e;
  const int batch = 4;
  if (!(e = nn_batch_buffers(*n, batch)).empty()) return e;
  const long hw = (long)n->H * n->W;
  std::vector<double> acc(9, 0.0), part;
  for (int s0 = first; s0 < first + count; s0 += batch) {
    const int b = std::min(batch, first + count - s0);
    std::vector<int> idx(b);
    for (int q = 0; q < b; ++q) idx[q] = s0 + q;
    if (hipStreamSynchronize(st) != hipSuccess ||
        hipMemcpy(n->idx, idx.data(), b * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
      return "index upload";
    NN_LAUNCH(k_pack, dim3(nblk(hw * b)), dim3(256), n->H, n->W, n->store_in, n->store_tg, n->store_sc, n->idx, b,
              n->norm, n->stat, n->act[0], n->tgt);
    if (!(e = nn_forward(*n, st, b)).empty()) return e;
    const long np = hw * b;
    const int nb = (int)nblk(np);
    NN_LAUNCH(k_eval, dim3(nb), dim3(256), np, hw, n->out, n->tgt, n->norm, n->stat, n->store_sc, n->idx, n->dpart);
    part.resize((size_t)9 * nb);
    if (hipMemcpyAsync(part.data(), n->dpart, part.size() * sizeof(double), hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return "evaluation read-back";
    for (int i = 0; i < nb; ++i)
      for (int q = 0; q < 9; ++q) acc[q] += part[(size_t)9 * i + q];
  }
  // U-point counts are not kept separately: both use the T-point sea count
  for (int q = 0; q < 8; ++q) res[q] = acc[8] > 0 ? std::sqrt(acc[q] / acc[8]) : 0.0;
  return "";
}

size_t nn_param_count(const Ctx& c) {
  if (!c.nn) return 0;
  size_t s = 0;
  for (int l = 0; l < kLayers; ++l) s += (size_t)c.nn->L[l].nw() + c.nn->L[l].cout;
  return s + kNormValues;
}

// parameters as [layer][w (co, 9, ci) then b (co)], FP32 master values
std::string nn_params(Ctx& c, float* host, bool set) {
  Nn* n = c.nn;
  if (hipStreamSynchronize(c.stream) != hipSuccess) return "sync";
  size_t off = 0;
  for (int l = 0; l < kLayers; ++l) {
    NnLayer& L = n->L[l];
    hipError_t e1, e2;
    if (set) {
      e1 = hipMemcpy(L.w32, host + off, L.nw() * sizeof(float), hipMemcpyHostToDevice);
      e2 = hipMemcpy(L.b, host + off + L.nw(), L.cout * sizeof(float), hipMemcpyHostToDevice);
    } else {
      e1 = hipMemcpy(host + off, L.w32, L.nw() * sizeof(float), hipMemcpyDeviceToHost);
      e2 = hipMemcpy(host + off + L.nw(), L.b, L.cout * sizeof(float), hipMemcpyDeviceToHost);
    }
    if (e1 != hipSuccess || e2 != hipSuccess) return "parameter copy";
    off += (size_t)L.nw() + L.cout;
  }
  float* nv[] = {n->norm.in_mean, n->norm.in_std, n->norm.out_std};
  const int nc[] = {kDyn, kDyn, kOutUsed};
  for (int a = 0; a < 3; ++a)
    for (int q = 0; q < nc[a]; ++q, ++off) {
      if (set) nv[a][q] = host[off];
      else host[off] = nv[a][q];
    }
  if (set) {
    n->norm.cor_scale = host[off];
    n->norm.prh_scale = host[off + 1];
    n->norm.s_floor = host[off + 2];
  } else {
    host[off] = n->norm.cor_scale;
    host[off + 1] = n->norm.prh_scale;
    host[off + 2] = n->norm.s_floor;
  }
  if (set) {
    n->normalised = true;
    hipStream_t st = c.stream;
    for (int l = 0; l < kLayers; ++l)
      NN_LAUNCH(k_w_bf16, dim3(nblk(n->L[l].nw())), dim3(256), n->L[l].cout, n->L[l].cin, n->L[l].w32, n->L[l].w,
                n->L[l].wt);
    if (hipStreamSynchronize(st) != hipSuccess) return "parameter sync";
  }
  return "";
}

// one forward + backward pass on caller-provided normalised tensors (host
// FP32): x [batch][H][W][8], target [batch][H][W][4]; returns the network
// output [batch][H][W][4], the loss and the gradients in parameter order
std::string nn_forward_backward(Ctx& c, int batch, const float* x, const float* target, float* out, float* grads,
                                double* loss) {
  Nn* n = c.nn;
  hipStream_t st = c.stream;
  std::string e;
  if (!(e = nn_batch_buffers(*n, batch)).empty()) return e;
  const long np = (long)batch * n->H * n->W;
  if (hipStreamSynchronize(st) != hipSuccess) return "sync";
  std::vector<bf16_t> hx((size_t)np * kInC, 0);
  for (long p = 0; p < np; ++p)
    for (int ch = 0; ch < kInUsed; ++ch) hx[(size_t)p * kInC + ch] = f2bf(x[p * kInUsed + ch]);
  if (hipMemcpy(n->act[0], hx.data(), hx.size() * sizeof(bf16_t), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(n->tgt, target, (size_t)np * kOutUsed * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
    return "input upload";
  if (!(e = nn_forward(*n, st, batch)).empty()) return e;
  double* ldev = n->misc + 8;
  if (!(e = nn_loss(*n, st, batch, ldev)).empty()) return e;
  if (!(e = nn_backward(*n, st, batch)).empty()) return e;
  if (hipStreamSynchronize(st) != hipSuccess) return "sync";
  std::vector<float> ho((size_t)np * kOutC);
  if (hipMemcpy(ho.data(), n->out, ho.size() * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) return "output";
  for (long p = 0; p < np; ++p)
    for (int ch = 0; ch < kOutUsed; ++ch) out[p * kOutUsed + ch] = ho[(size_t)p * kOutC + ch];
  double l;
  if (hipMemcpy(&l, ldev, sizeof l, hipMemcpyDeviceToHost) != hipSuccess) return "loss";
  *loss = l / (double)(np * kOutUsed);
  size_t off = 0;
  for (int ll = 0; ll < kLayers; ++ll) {
    NnLayer& L = n->L[ll];
    if (hipMemcpy(grads + off, L.gw, L.nw() * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(grads + off + L.nw(), L.gb, L.cout * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
      return "gradient read-back";
    off += (size_t)L.nw() + L.cout;
  }
  for (int q = 0; q < kNormValues; ++q) grads[off + q] = 0.f;
  return "";
}

void nn_stats(const Ctx& c, pcom_nn_stats* s) {
  std::memset(s, 0, sizeof *s);
  if (!c.nn) return;
  const Nn* n = c.nn;
  s->recorded = n->nrec;
  s->surrogate_steps = n->surrogate_steps;
  s->plain_steps = n->plain_steps;
  s->checks = n->checks;
  s->fallbacks = n->fallbacks;
  s->in_use = n->use_every > 0 && n->step_count >= n->hold_until ? 1 : 0;
  for (int q = 0; q < 4; ++q) {
    s->last_check_rmse[q] = n->last_err[q];
    s->last_persist_rmse[q] = n->last_persist[q];
    s->max_check_rmse[q] = n->max_err[q];
  }
}

}  // namespace pcom
