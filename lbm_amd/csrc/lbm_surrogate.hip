// Learned warm start (include/lbm_surrogate.h): a U-Net of 3x3
// convolutions on six resolution levels, trained and evaluated with
// hand-written gfx950 kernels.
//
// Tensors are NHWC.  Activations are bf16 (uint16 bit patterns), gradients
// of activations FP32, parameters FP32 masters with bf16 working copies.
// A convolution is an implicit GEMM on v_mfma_f32_32x32x16_bf16: one wave
// owns a 32-pixel x 32-output-channel tile and walks K = 9 taps x input
// channels in steps of 16 (lane l supplies pixel / output channel l % 32 and
// the 8 channels 8 (l / 32) .. +7 of the step, one 16-byte load each).  The
// decoder's concatenation of an upsampled deep tensor with a skip tensor is
// read in place (nearest upsampling in the address), never materialised.
//   forward:  Y = act(conv(X, W) + b)                       k_conv<0 or 2>
//   dgrad:    dX = conv(dY, Wd), Wd[ci][8-tap][co] = W[co][tap][ci]   k_conv<1>
//   wgrad:    dW[co][tap][ci] = sum_p dY[p][co] X[p + off(tap)][ci]  k_wgrad
// wgrad splits the pixel sum into chunks reduced in a fixed order, so a
// gradient is reproducible bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/lbm_hip.h"
#include "../../include/lbm_surrogate.h"

namespace sur {

typedef unsigned short u16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

constexpr int WCHUNK = 4096;   // pixels per bias-gradient partial sum
constexpr int WGCHUNK = 1024;  // pixels per weight-gradient partial sum (one wave each)

__device__ __forceinline__ float bf2f(u16 b) { return __uint_as_float(((unsigned)b) << 16); }
__device__ __forceinline__ u16 f2bf(float v) {
  unsigned u = __float_as_uint(v);
  if ((u & 0x7f800000u) == 0x7f800000u) return (u16)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (u16)(u >> 16);
}
__device__ __forceinline__ bf16x8 ld8(const u16 *p) {
  const u16x8 v = *reinterpret_cast<const u16x8 *>(p);
  return __builtin_bit_cast(bf16x8, v);
}

struct ConvArgs {
  const u16 *a;  // upsampled input (stored at half resolution), ca channels, may be null
  const u16 *s;  // same-resolution input, cs channels
  int ca, cs;
  const u16 *w;  // [9][(ca + cs) / 16][co][16]: K step k of output channel o at (k * co + o) * 16
  const float *bias;
  int nb, h, wd, co;
  // MODE 0: y = bf16(act(acc + bias)), [p][co]
  u16 *y;
  int relu;
  // MODE 1: channels < c0 are stored into g0 [p][c0], the rest into g1
  // [p][co - c0], added to what g1 holds when add1 is set
  float *g0, *g1;
  int c0, add1;
  // MODE 2: yf = acc + bias for the first cf channels, [p][cf]
  float *yf;
  int cf;
};

// One 32-pixel x 32-output-channel tile per `ksplit` waves of a block (1, 2
// or 4): the K steps (tap, 16 channels) are split into ksplit contiguous
// ranges, each wave issues the next step's two loads before this step's
// product, and the partial tiles are added in wave order through LDS by the
// tile's first wave, which applies the epilogue.
template <int MODE>
__global__ void __launch_bounds__(256) k_conv(const ConvArgs a, const int ksplit) {
  __shared__ float red[4][16][64];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long long npix = (long long)a.nb * a.h * a.wd;
  const int ntc = a.co >> 5;
  const int kp = wave % ksplit;  // this wave's part of K
  const long long tile = (long long)blockIdx.x * (4 / ksplit) + wave / ksplit;
  const long long p0 = (tile / ntc) * 32;
  const bool active = p0 < npix;  // uniform over the waves of a tile
  const int co0 = (int)(tile % ntc) * 32;
  const int ci = a.ca + a.cs;
  const int m = lane & 31, g = lane >> 5;
  const long long p = p0 + m;
  const bool pin = active && p < npix;
  int b = 0, y = 0, x = 0;
  if (pin) {
    x = (int)(p % a.wd);
    y = (int)((p / a.wd) % a.h);
    b = (int)(p / ((long long)a.wd * a.h));
  }
  const u16 *wrow = a.w + (long long)(co0 + m) * 16 + 8 * g;
  const long long wstep = (long long)a.co * 16;  // elements per K step
  f32x16 acc;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.f;
  const int h2 = a.h >> 1, w2 = a.wd >> 1;
  const int cpt = ci >> 4;  // K steps per tap
  const int nks = 9 * cpt;
  const int k0 = kp * nks / ksplit, k1 = (kp + 1) * nks / ksplit;
  const u16x8 zero = {0, 0, 0, 0, 0, 0, 0, 0};
  // operands of K step k: 8 input channels of this lane's pixel (shifted by
  // the tap) and 8 weights of its output channel
  auto load = [&](int k, u16x8 &av, u16x8 &bv) {
    const int tap = k / cpt, c = (k - tap * cpt) << 4;
    const int ys = y + tap / 3 - 1, xs = x + tap % 3 - 1;
    bv = *reinterpret_cast<const u16x8 *>(wrow + k * wstep);
    av = zero;
    if (pin && ys >= 0 && ys < a.h && xs >= 0 && xs < a.wd) {
      if (c < a.ca)
        av = *reinterpret_cast<const u16x8 *>(a.a + (((long long)b * h2 + (ys >> 1)) * w2 + (xs >> 1)) * a.ca + c + 8 * g);
      else
        av = *reinterpret_cast<const u16x8 *>(a.s + (((long long)b * a.h + ys) * a.wd + xs) * a.cs + (c - a.ca) + 8 * g);
    }
  };
  if (active && k0 < k1) {
    u16x8 av, bv;
    load(k0, av, bv);
    for (int k = k0; k < k1; ++k) {
      u16x8 av1 = zero, bv1 = zero;
      if (k + 1 < k1) load(k + 1, av1, bv1);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, av), __builtin_bit_cast(bf16x8, bv), acc,
                                                    0, 0, 0);
      av = av1;
      bv = bv1;
    }
  }
  if (ksplit > 1) {
    if (kp > 0)
#pragma unroll
      for (int r = 0; r < 16; ++r) red[wave][r][lane] = acc[r];
    __syncthreads();
    if (kp > 0) return;
    for (int q = 1; q < ksplit; ++q)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[r] += red[wave + q][r][lane];
  }
  if (!active) return;
  // lane holds C[mm][n], n = lane % 32, mm = 8 (r / 4) + 4 (lane / 32) + r % 4
  const int co_ = co0 + m;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const long long pp = p0 + ((r >> 2) << 3) + 4 * g + (r & 3);
    if (pp >= npix) continue;
    float v = acc[r];
    if (MODE == 0) {
      v += a.bias[co_];
      if (a.relu) v = fmaxf(v, 0.f);
      a.y[pp * a.co + co_] = f2bf(v);
    } else if (MODE == 1) {
      if (co_ < a.c0) {
        a.g0[pp * a.c0 + co_] = v;
      } else {
        float *d = a.g1 + pp * (a.co - a.c0) + (co_ - a.c0);
        *d = a.add1 ? *d + v : v;
      }
    } else {
      if (co_ < a.cf) a.yf[pp * a.cf + co_] = v + a.bias[co_];
    }
  }
}

// The same convolution with the input staged in LDS (levels whose width is
// a multiple of 32 and ci <= LDS_CI_MAX): a block computes 4 image rows x
// 32 pixels x 32 output channels (blockIdx.y = output tile), one wave per
// row.  The block first loads the 6 x 34 input pixels of its rows and their
// one-pixel halo (zero outside the image; the upsampled channels first, read
// at half resolution) with 16-byte loads, then every tap reads its A operand
// from LDS, so each input element comes from HBM/L2 once per block instead
// of once per tap.  Pixel stride in LDS: ci + 8 elements.
constexpr int LDS_CI_MAX = 128;
template <int MODE>
__global__ void __launch_bounds__(256) k_conv_lds(const ConvArgs a) {
  extern __shared__ __align__(16) u16 xs[];  // [6][34][ci + 8]
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int ci = a.ca + a.cs, cst = ci + 8;
  const int xt = a.wd >> 5, ry = (a.h + 3) >> 2;
  const int blk = blockIdx.x;
  const int x0 = (blk % xt) * 32;
  const int y0 = ((blk / xt) % ry) * 4;
  const int b = blk / (xt * ry);
  const int co0 = blockIdx.y * 32;
  const int h2 = a.h >> 1, w2 = a.wd >> 1;
  // stage rows y0-1 .. y0+4, columns x0-1 .. x0+32, 8 channels per load
  const int cg = ci >> 3;
  for (int e = threadIdx.x; e < 6 * 34 * cg; e += 256) {
    const int c = (e % cg) << 3;
    const int pix = e / cg;
    const int yy = y0 - 1 + pix / 34, xx = x0 - 1 + pix % 34;
    u16x8 v = {0, 0, 0, 0, 0, 0, 0, 0};
    if (yy >= 0 && yy < a.h && xx >= 0 && xx < a.wd) {
      if (c < a.ca)
        v = *reinterpret_cast<const u16x8 *>(a.a + (((long long)b * h2 + (yy >> 1)) * w2 + (xx >> 1)) * a.ca + c);
      else
        v = *reinterpret_cast<const u16x8 *>(a.s + (((long long)b * a.h + yy) * a.wd + xx) * a.cs + (c - a.ca));
    }
    *reinterpret_cast<u16x8 *>(xs + pix * cst + c) = v;
  }
  __syncthreads();
  const int y = y0 + wave;
  if (y >= a.h) return;  // after the only barrier
  const int m = lane & 31, g = lane >> 5;
  const u16 *wrow = a.w + (long long)(co0 + m) * 16 + 8 * g;
  const long long wstep = (long long)a.co * 16;  // elements per K step
  f32x16 acc;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.f;
  const int cpt = ci >> 4, nks = 9 * cpt;
  u16x8 bv = *reinterpret_cast<const u16x8 *>(wrow);
  for (int k = 0; k < nks; ++k) {
    const int tap = k / cpt, c = (k - tap * cpt) << 4;
    u16x8 bv1 = bv;
    if (k + 1 < nks) bv1 = *reinterpret_cast<const u16x8 *>(wrow + (k + 1) * wstep);
    // pixel (y + dy, x0 + m + dx) sits at LDS row wave + 1 + dy, column m + 1 + dx
    const int pix = (wave + tap / 3) * 34 + m + tap % 3;
    const u16x8 av = *reinterpret_cast<const u16x8 *>(xs + pix * cst + c + 8 * g);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, av), __builtin_bit_cast(bf16x8, bv), acc,
                                                  0, 0, 0);
    bv = bv1;
  }
  // lane holds C[mm][n], n = lane % 32, mm = 8 (r / 4) + 4 (lane / 32) + r % 4
  const int co_ = co0 + m;
  const long long prow = ((long long)b * a.h + y) * a.wd + x0;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const long long pp = prow + ((r >> 2) << 3) + 4 * g + (r & 3);
    float v = acc[r];
    if (MODE == 0) {
      v += a.bias[co_];
      if (a.relu) v = fmaxf(v, 0.f);
      a.y[pp * a.co + co_] = f2bf(v);
    } else if (MODE == 1) {
      if (co_ < a.c0) {
        a.g0[pp * a.c0 + co_] = v;
      } else {
        float *d = a.g1 + pp * (a.co - a.c0) + (co_ - a.c0);
        *d = a.add1 ? *d + v : v;
      }
    } else {
      if (co_ < a.cf) a.yf[pp * a.cf + co_] = v + a.bias[co_];
    }
  }
}

struct WgradArgs {
  const u16 *gy;  // [p][co] bf16
  const u16 *a;   // upsampled input at half resolution (ca channels), may be null
  const u16 *s;   // same-resolution input (cs channels)
  int ca, cs, co;
  int nb, h, wd;
  int nchunks;
  float *part;  // [chunk][co][9][ci]
};

// One wave: (co tile, ci tile, tap, pixel chunk).  Operands are gathered
// along the pixel axis (strided 2-byte loads), the pixel coordinates carried
// from pixel to pixel; the sum over the chunk runs in pixel order, so the
// gradient is the same bit for bit however the loads are scheduled.
__global__ void __launch_bounds__(256) k_wgrad(const WgradArgs a) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int ci = a.ca + a.cs;
  const int ntco = a.co >> 5, ntci = (ci + 31) >> 5;
  long long t = (long long)blockIdx.x * 4 + wave;
  const long long total = (long long)ntco * ntci * 9 * a.nchunks;
  if (t >= total) return;
  const int chunk = (int)(t % a.nchunks);
  t /= a.nchunks;
  const int tap = (int)(t % 9);
  t /= 9;
  const int cit = (int)(t % ntci);
  const int cot = (int)(t / ntci);
  const int co0 = cot * 32, ci0 = cit * 32;
  const int m = lane & 31, g = lane >> 5;
  const long long npix = (long long)a.nb * a.h * a.wd;
  const long long pbeg = (long long)chunk * WGCHUNK;
  const long long pend = pbeg + WGCHUNK < npix ? pbeg + WGCHUNK : npix;
  const int dy = tap / 3 - 1, dx = tap % 3 - 1;
  const int cin = ci0 + m;  // this lane's B column (input channel)
  const bool cok = cin < ci;
  const bool upart = cin < a.ca;
  const int h2 = a.h >> 1, w2 = a.wd >> 1;
  f32x16 acc;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.f;
  // image, row and column of this lane's first pixel pk + 8 g, advanced by
  // carries instead of divisions
  int bx, by, bb;
  {
    const long long p0 = pbeg + 8 * g;
    bx = (int)(p0 % a.wd);
    by = (int)((p0 / a.wd) % a.h);
    bb = (int)(p0 / ((long long)a.wd * a.h));
  }
  // operands of the 16 pixels from pk: this lane's 8 pixels pk + 8 g + i
  auto gather = [&](long long pk, u16x8 &av, u16x8 &bv) {
    int x = bx, y = by, b = bb;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const long long p = pk + 8 * g + i;
      u16 va = 0, vb = 0;
      if (p < pend) {
        va = a.gy[p * a.co + co0 + m];
        if (cok) {
          const int ys = y + dy, xs = x + dx;
          if (ys >= 0 && ys < a.h && xs >= 0 && xs < a.wd) {
            if (upart) vb = a.a[(((long long)b * h2 + (ys >> 1)) * w2 + (xs >> 1)) * a.ca + cin];
            else vb = a.s[(((long long)b * a.h + ys) * a.wd + xs) * a.cs + (cin - a.ca)];
          }
        }
      }
      av[i] = va;
      bv[i] = vb;
      if (++x == a.wd) {
        x = 0;
        if (++y == a.h) {
          y = 0;
          ++b;
        }
      }
    }
    bx += 16;  // the next step starts 16 pixels on
    while (bx >= a.wd) {
      bx -= a.wd;
      if (++by == a.h) {
        by = 0;
        ++bb;
      }
    }
  };
  // two steps' loads in flight before their two products, which are
  // accumulated in pixel order as one step at a time would (four steps in
  // flight measured no faster)
  for (long long pk = pbeg; pk < pend; pk += 32) {
    u16x8 av0, bv0, av1, bv1;
    gather(pk, av0, bv0);
    const bool two = pk + 16 < pend;  // wave-uniform
    if (two) gather(pk + 16, av1, bv1);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, av0), __builtin_bit_cast(bf16x8, bv0),
                                                  acc, 0, 0, 0);
    if (two)
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, av1), __builtin_bit_cast(bf16x8, bv1),
                                                    acc, 0, 0, 0);
  }
  // C[mm][n]: mm = output channel row, n = input channel column (lane % 32)
  if (!cok) return;
  float *dst = a.part + (long long)chunk * a.co * 9 * ci;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int mm = ((r >> 2) << 3) + 4 * g + (r & 3);
    dst[((long long)(co0 + mm) * 9 + tap) * ci + cin] = acc[r];
  }
}

// ---------------------------------------------------------------------------
// Weight gradient on channel-major operands.  The pixel sum is the GEMM's K
// axis, and NHWC keeps pixels strided; two staging kernels write the layer's
// output gradient as gT[p / 8][co][8] and its input, zero padded, as
// xT[b][h + 2][(w + 8) / 8][ci][8] (one pixel of halo on each side, the rows
// padded to a multiple of 8 elements), both through an LDS tile transpose:
// groups of 8 consecutive pixels of one channel, the channels of a group
// next to each other, so the MFMA operands of k_wgrad_t are 16-byte loads of
// 8 consecutive pixels and the 32 lanes of a channel tile read 512
// contiguous bytes.  Used where the width at the layer's level is a
// multiple of 8.
constexpr int TP = 64;  // pixels per transpose tile

// From the FP32 gradient g [p][co] of a layer's output (relu mask from its
// bf16 activation y), or from a bf16 gradient gin: gb = bf16 values [p][co]
// (not written when gin is given), gT [p / 8][co][8], and per-tile bias sums
// bpart[tile][co] of the bf16 values (8-pixel groups, then the groups in
// order).  Block: 64 pixels x 32 channels; grid (ceil(P / 64), co / 32).
__global__ void __launch_bounds__(256) k_gy_prep(const float *g, const u16 *y, int relu, const u16 *gin, u16 *gb,
                                                 u16 *gT, float *bpart, long long P, int co) {
  __shared__ u16 tile[32][TP + 8];
  __shared__ float red[8][32];
  const int t = threadIdx.x;
  const int px = t >> 2, k = t & 3;
  const long long p0 = (long long)blockIdx.x * TP;
  const int c0 = blockIdx.y * 32;
  const long long p = p0 + px;
  u16x8 v = {0, 0, 0, 0, 0, 0, 0, 0};
  if (p < P) {
    const long long off = p * co + c0 + 8 * k;
    if (gin) {
      v = *reinterpret_cast<const u16x8 *>(gin + off);
    } else {
      const float4 lo = *reinterpret_cast<const float4 *>(g + off);
      const float4 hi = *reinterpret_cast<const float4 *>(g + off + 4);
      float f[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
      u16x8 yy = {0, 0, 0, 0, 0, 0, 0, 0};
      if (relu) yy = *reinterpret_cast<const u16x8 *>(y + off);
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        if (relu && !(bf2f(yy[i]) > 0.f)) f[i] = 0.f;
        v[i] = f2bf(f[i]);
      }
      *reinterpret_cast<u16x8 *>(gb + off) = v;
    }
  }
#pragma unroll
  for (int i = 0; i < 8; ++i) tile[8 * k + i][px] = v[i];
  __syncthreads();
  const int ch = t >> 3, j = t & 7;
  const long long pp = p0 + 8 * j;
  u16x8 w;
#pragma unroll
  for (int i = 0; i < 8; ++i) w[i] = tile[ch][8 * j + i];
  if (pp < P) *reinterpret_cast<u16x8 *>(gT + ((pp >> 3) * co + c0 + ch) * 8) = w;
  float sum = 0.f;
#pragma unroll
  for (int i = 0; i < 8; ++i) sum += bf2f(w[i]);  // zero beyond P
  red[j][ch] = sum;
  __syncthreads();
  if (t < 32) {
    float acc = 0.f;
#pragma unroll
    for (int q = 0; q < 8; ++q) acc += red[q][t];
    bpart[(long long)blockIdx.x * co + c0 + t] = acc;
  }
}

// xT[b][yp][xp / 8][c][xp % 8] = input channel c of pixel (b, yp - 1,
// xp - 1), zero outside the image; channels < ca come from the
// half-resolution tensor A (nearest upsampling), the rest from S.  Block: 64
// padded columns x 32 channels of one padded row; grid (ceil((w + 8) / 64),
// nb (h + 2), ceil(ci / 32)).  Requires ca, cs multiples of 8.
__global__ void __launch_bounds__(256) k_x_prep(const u16 *A, int ca, const u16 *S, int cs, int nb, int h, int w,
                                                u16 *xT) {
  __shared__ u16 tile[32][TP + 8];
  const int ci = ca + cs, wp = w + 8, hp = h + 2;
  const int t = threadIdx.x, xl = t >> 2, k = t & 3;
  const int row = blockIdx.y;  // b * hp + yp
  const int b = row / hp, yp = row % hp;
  const int c0 = blockIdx.z * 32, c = c0 + 8 * k;
  const int x = (int)blockIdx.x * TP + xl - 1, yy = yp - 1;
  u16x8 v = {0, 0, 0, 0, 0, 0, 0, 0};
  if (c < ci && yy >= 0 && yy < h && x >= 0 && x < w) {
    if (c < ca)
      v = *reinterpret_cast<const u16x8 *>(A + (((long long)b * (h >> 1) + (yy >> 1)) * (w >> 1) + (x >> 1)) * ca + c);
    else
      v = *reinterpret_cast<const u16x8 *>(S + (((long long)b * h + yy) * w + x) * cs + (c - ca));
  }
#pragma unroll
  for (int i = 0; i < 8; ++i) tile[8 * k + i][xl] = v[i];
  __syncthreads();
  const int ch = t >> 3, j = t & 7;
  const int xq = (int)blockIdx.x * TP + 8 * j;
  if (c0 + ch < ci && xq < wp) {
    u16x8 o;
#pragma unroll
    for (int i = 0; i < 8; ++i) o[i] = tile[ch][8 * j + i];
    *reinterpret_cast<u16x8 *>(xT + (((long long)row * (wp >> 3) + (xq >> 3)) * ci + c0 + ch) * 8) = o;
  }
}

constexpr int WG_DEPTH = 4;  // K steps of k_wgrad_t in flight per wave

struct WgradTArgs {
  const u16 *gT;  // [P / 8][co][8]
  const u16 *xT;  // [nb][h + 2][(w + 8) / 8][ci][8]
  long long P;
  int co, ci, nb, h, w;
  int cseg;     // 8-pixel segments per chunk (even)
  int nchunks;  // a multiple of 4: the 4 waves of a block take 4 chunks of one tile
  float *part;  // [chunk / 4][co][9][ci]
};

// 8 bf16 values starting `sh` elements (0, 1 or 2) into lo:hi
__device__ __forceinline__ bf16x8 shifted8(const u16x8 lo, const u16x8 hi, int sh) {
  typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
  const u32x4 a = __builtin_bit_cast(u32x4, lo), b = __builtin_bit_cast(u32x4, hi);
  const unsigned d[5] = {a[0], a[1], a[2], a[3], b[0]};
  u32x4 r;
  if (sh == 0) {
    r = a;
  } else if (sh == 2) {
    r = u32x4{d[1], d[2], d[3], d[4]};
  } else {
#pragma unroll
    for (int q = 0; q < 4; ++q) r[q] = (unsigned)((((unsigned long long)d[q + 1] << 32) | d[q]) >> 16);
  }
  return __builtin_bit_cast(bf16x8, r);
}

// One wave: (co tile, ci tile, tap row dy, chunk of cseg 8-pixel segments),
// the three taps of the row (three 32x32 accumulators).  A K step is two
// segments (lane group g takes segment 2k + g; a segment never crosses an
// image row since w % 8 == 0): the gradient operand is one 16-byte load,
// the input row two aligned loads serving the three dx shifts; the loads
// run WG_DEPTH steps ahead of the products.  The chunk sums in
// pixel order, so the gradient is the same bit for bit however the loads
// are scheduled.  The 4 waves of a block take 4 consecutive chunks of one
// tile and add their sums through LDS in wave order before one partial per
// block is written.
__global__ void __launch_bounds__(256) k_wgrad_t(const WgradTArgs a) {
  __shared__ float red[3][3][16][64];  // waves 1..3: [wave - 1][dx][r][lane]
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int ntco = a.co >> 5, ntci = (a.ci + 31) >> 5;
  long long t = (long long)blockIdx.x * 4 + wave;  // the grid is exactly the tiles x chunks
  const int chunk = (int)(t % a.nchunks);
  t /= a.nchunks;
  const int dy = (int)(t % 3);
  t /= 3;
  const int cit = (int)(t % ntci), cot = (int)(t / ntci);
  const int co0 = cot * 32, ci0 = cit * 32;
  const int m = lane & 31, g = lane >> 5;
  const int wp = a.w + 8, hp = a.h + 2;
  const bool cok = ci0 + m < a.ci;
  const u16 *arow = a.gT + (long long)(co0 + m) * 8;                       // + segment * co * 8
  const u16 *xch = a.xT + (long long)(cok ? ci0 + m : 0) * 8 + (long long)dy * (wp >> 3) * a.ci * 8;  // + group * ci * 8
  const int wg = wp >> 3;
  const long long nseg = a.P >> 3;
  const long long sbeg = (long long)chunk * a.cseg;
  const long long send = sbeg + a.cseg < nseg ? sbeg + a.cseg : nseg;
  f32x16 acc[3];
#pragma unroll
  for (int q = 0; q < 3; ++q)
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[q][r] = 0.f;
  // pixel of this lane's segment, advanced by carries
  long long s = sbeg + g;
  int x, y, b;
  {
    const long long p = s << 3;
    x = (int)(p % a.w);
    const long long r = p / a.w;
    y = (int)(r % a.h);
    b = (int)(r / a.h);
  }
  const u16x8 zero = {0, 0, 0, 0, 0, 0, 0, 0};
  auto load = [&](u16x8 &av, u16x8 &lo, u16x8 &hi) {
    av = zero;
    lo = zero;
    hi = zero;
    if (s < send) {
      av = *reinterpret_cast<const u16x8 *>(arow + s * a.co * 8);
      if (cok) {
        // padded columns x .. x + 15: the groups x / 8 and x / 8 + 1 of the row
        const u16 *xr = xch + (((long long)b * hp + y) * wg + (x >> 3)) * a.ci * 8;
        lo = *reinterpret_cast<const u16x8 *>(xr);
        hi = *reinterpret_cast<const u16x8 *>(xr + (long long)a.ci * 8);
      }
    }
    s += 2;
    x += 16;
    while (x >= a.w) {
      x -= a.w;
      if (++y == a.h) {
        y = 0;
        ++b;
      }
    }
  };
  // a ring of WG_DEPTH steps in flight: step k's operands are loaded while
  // the products of steps k - WG_DEPTH + 1 .. k - 1 run (a wave's chain of
  // loads and products is the whole cost of this kernel at two waves per
  // SIMD); the steps still run in order
  u16x8 av[WG_DEPTH], lo[WG_DEPTH], hi[WG_DEPTH];
#pragma unroll
  for (int i = 0; i < WG_DEPTH; ++i) load(av[i], lo[i], hi[i]);
  for (long long sk = sbeg; sk < send; sk += 2 * WG_DEPTH) {
#pragma unroll
    for (int i = 0; i < WG_DEPTH; ++i) {
      if (sk + 2 * i < send) {  // wave-uniform
        const bf16x8 A = __builtin_bit_cast(bf16x8, av[i]);
#pragma unroll
        for (int dx = 0; dx < 3; ++dx)
          acc[dx] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A, shifted8(lo[i], hi[i], dx), acc[dx], 0, 0, 0);
        load(av[i], lo[i], hi[i]);  // the step WG_DEPTH ahead (zeros past the chunk)
      }
    }
  }
  if (wave > 0)
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
      for (int r = 0; r < 16; ++r) red[wave - 1][q][r][lane] = acc[q][r];
  __syncthreads();
  if (wave > 0 || !cok) return;
#pragma unroll
  for (int w = 0; w < 3; ++w)
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[q][r] += red[w][q][r][lane];
  float *dst = a.part + (long long)(chunk >> 2) * a.co * 9 * a.ci;
#pragma unroll
  for (int q = 0; q < 3; ++q)
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int mm = ((r >> 2) << 3) + 4 * g + (r & 3);
      dst[((long long)(co0 + mm) * 9 + 3 * dy + q) * a.ci + ci0 + m] = acc[q][r];
    }
}

// dst[c] = sum over rows r of part[r][c] (n columns): block c, thread t sums
// rows t, t + 256, ... in order, then a fixed tree
__device__ void reduce_rows(const float *part, long long nrows, int n, float *dst, int c) {
  __shared__ float red[256];
  float s = 0.f;
  for (long long r = threadIdx.x; r < nrows; r += 256) s += part[r * n + c];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) dst[c] = red[0];
}

// dst[i] = sum over chunks of part[chunk][i]: a block takes 32 elements,
// its eight 32-lane groups each sum one eighth of the chunks in order, and
// the eight partial sums are added in group order
__device__ void reduce_chunks(const float *part, long long n, int nchunks, float *dst, long long blk) {
  __shared__ float red[8][32];
  const int e = threadIdx.x & 31, q = threadIdx.x >> 5;
  const long long i = blk * 32LL + e;
  float s = 0.f;
  if (i < n) {
    const int c1 = (int)((long long)(q + 1) * nchunks / 8);
    int c = (int)((long long)q * nchunks / 8);
    for (; c + 4 <= c1; c += 4) {
      float v[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) v[k] = part[(long long)(c + k) * n + i];
#pragma unroll
      for (int k = 0; k < 4; ++k) s += v[k];
    }
    for (; c < c1; ++c) s += part[(long long)c * n + i];
  }
  red[q][e] = s;
  __syncthreads();
  if (q == 0 && i < n) {
    float t = 0.f;
#pragma unroll
    for (int k = 0; k < 8; ++k) t += red[k][e];
    dst[i] = t;
  }
}

__global__ void __launch_bounds__(256) k_reduce_chunks(const float *part, long long n, int nchunks, float *dst) {
  reduce_chunks(part, n, nchunks, dst, blockIdx.x);
}

// one launch for a layer's two parameter gradients: blocks below nbw sum the
// weight-gradient chunks (32 weights a block), the co blocks after them the
// bias partial rows
__global__ void __launch_bounds__(256) k_reduce_wb(const float *part, long long nw, int nchunks, float *dw,
                                                   const float *bpart, long long nrows, int co, float *db) {
  const long long nbw = (nw + 31) / 32;
  if ((long long)blockIdx.x < nbw) reduce_chunks(part, nw, nchunks, dw, blockIdx.x);
  else reduce_rows(bpart, nrows, co, db, (int)(blockIdx.x - nbw));
}

// bias gradient partials: part[chunk][c] = sum over the chunk's pixels of
// gy[p][c]; thread (r, c) of the block sums pixels r, r + 256/co, ... of
// the chunk, then the rows are added in a fixed tree (co = 32..256, a power
// of two)
__global__ void __launch_bounds__(256) k_bias_part(const u16 *gy, long long npix, int co, float *part) {
  __shared__ float red[256];
  const int rows = 256 / co;
  const int c = threadIdx.x % co, r = threadIdx.x / co;
  const long long beg = (long long)blockIdx.x * WCHUNK;
  const long long end = beg + WCHUNK < npix ? beg + WCHUNK : npix;
  float s = 0.f;
  long long p = beg + r;
  // eight loads in flight, added in the same order as one at a time
  for (; p + 7LL * rows < end; p += 8LL * rows) {
    float v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = bf2f(gy[(p + (long long)k * rows) * co + c]);
#pragma unroll
    for (int k = 0; k < 8; ++k) s += v[k];
  }
  for (; p < end; p += rows) s += bf2f(gy[p * co + c]);
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = rows / 2; k > 0; k >>= 1) {
    if (r < k) red[threadIdx.x] += red[threadIdx.x + k * co];
    __syncthreads();
  }
  if (r == 0) part[(long long)blockIdx.x * co + c] = red[c];
}

// gb = bf16(g * [y > 0]) (relu) or bf16(g)
__global__ void __launch_bounds__(256) k_mask_cast(const float *g, const u16 *y, u16 *gb, long long n, int relu) {
  const long long i = blockIdx.x * 256LL + threadIdx.x;
  if (i >= n) return;
  float v = g[i];
  if (relu && !(bf2f(y[i]) > 0.f)) v = 0.f;
  gb[i] = f2bf(v);
}

// 2x2 average pooling, NHWC, (h, w) = output size
__global__ void __launch_bounds__(256) k_pool(const u16 *x, u16 *y, int nb, int h, int w, int c) {
  const long long i = blockIdx.x * 256LL + threadIdx.x;
  const long long n = (long long)nb * h * w * c;
  if (i >= n) return;
  const int ch = (int)(i % c);
  long long p = i / c;
  const int xx = (int)(p % w), yy = (int)((p / w) % h), b = (int)(p / ((long long)w * h));
  const long long base = (((long long)b * 2 * h + 2 * yy) * 2 * w + 2 * xx) * c + ch;
  const float s = bf2f(x[base]) + bf2f(x[base + c]) + bf2f(x[base + 2LL * w * c]) + bf2f(x[base + 2LL * w * c + c]);
  y[i] = f2bf(0.25f * s);
}

// gx (full resolution, size 2h x 2w) = gy / 4 at the parent cell, added to
// what gx holds when add is set
__global__ void __launch_bounds__(256) k_pool_bwd(const float *gy, float *gx, int nb, int h, int w, int c, int add) {
  const long long i = blockIdx.x * 256LL + threadIdx.x;
  const long long n = (long long)nb * 2 * h * 2 * w * c;
  if (i >= n) return;
  const int ch = (int)(i % c);
  long long p = i / c;
  const int xx = (int)(p % (2 * w)), yy = (int)((p / (2 * w)) % (2 * h)), b = (int)(p / (4LL * w * h));
  const float v = 0.25f * gy[(((long long)b * h + (yy >> 1)) * w + (xx >> 1)) * c + ch];
  gx[i] = add ? gx[i] + v : v;
}

// ga (half resolution h x w) = sum of the 2x2 children of gup (nearest
// upsampling backward), added to what ga holds when add is set
__global__ void __launch_bounds__(256) k_up_bwd(const float *gup, float *ga, int nb, int h, int w, int c, int add) {
  const long long i = blockIdx.x * 256LL + threadIdx.x;
  const long long n = (long long)nb * h * w * c;
  if (i >= n) return;
  const int ch = (int)(i % c);
  long long p = i / c;
  const int xx = (int)(p % w), yy = (int)((p / w) % h), b = (int)(p / ((long long)w * h));
  const long long base = (((long long)b * 2 * h + 2 * yy) * 2 * w + 2 * xx) * c + ch;
  const float v = gup[base] + gup[base + c] + gup[base + 2LL * w * c] + gup[base + 2LL * w * c + c];
  ga[i] = add ? ga[i] + v : v;
}

// features [p][cin] FP32 -> bf16 [p][16] (zero padded)
__global__ void __launch_bounds__(256) k_cast_in(const float *xin, int cin, u16 *x, long long npix) {
  const long long i = blockIdx.x * 256LL + threadIdx.x;
  if (i >= npix * 16) return;
  const int c = (int)(i & 15);
  const long long p = i >> 4;
  x[i] = c < cin ? f2bf(xin[p * cin + c]) : (u16)0;
}

// squared-error loss on fluid cells: per-pixel contribution to part[block],
// gradient 2 (o - t) * scale into gout [p][32] bf16 (channels >= 3 zero)
__global__ void __launch_bounds__(256) k_loss(const float *o, const float *t, const unsigned char *mask, long long npix,
                                              float scale, u16 *gout, double *part) {
  __shared__ double red[256];
  const long long p = blockIdx.x * 256LL + threadIdx.x;
  double l = 0;
  if (p < npix) {
    const bool fl = mask[p] != 0;
    for (int c = 0; c < 32; ++c) {
      float gv = 0.f;
      if (c < 3 && fl) {
        const float d = o[p * 3 + c] - t[p * 3 + c];
        l += (double)d * d;
        gv = 2.f * d * scale;
      }
      gout[p * 32 + c] = f2bf(gv);
    }
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256) k_sum_parts(const double *part, int n, double *out) {
  __shared__ double red[256];
  double s = 0;
  for (int i = threadIdx.x; i < n; i += 256) s += part[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

__global__ void __launch_bounds__(256) k_adam(float *p, const float *g, float *m, float *v, long long n, float lr,
                                              float b1, float b2, float eps, float c1, float c2) {
  const long long i = blockIdx.x * 256LL + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// bf16 working copies of every layer in one launch (blockIdx.y = layer),
// in the convolutions' operand order [tap][in / 16][out][16] (K step
// tap * (in / 16) + c / 16 of output channel o: 16 consecutive input
// channels, the output channels of a step next to each other): wb, in =
// ci, out = co; the dgrad form wd, in = co, out = ci, taps flipped
constexpr int MAXLAYERS = 32;
struct PrepLayer {
  const float *wf;
  int co, ci;
  u16 *wb, *wd;
};
struct PrepAll {
  PrepLayer l[MAXLAYERS];
  int n;
};
__global__ void __launch_bounds__(256) k_prep_all(const PrepAll a) {
  const PrepLayer &L = a.l[blockIdx.y];
  const long long i = blockIdx.x * 256LL + threadIdx.x;
  const long long n = (long long)L.co * 9 * L.ci;
  if (i >= n) return;
  const int c = (int)(i % L.ci);
  const int tap = (int)((i / L.ci) % 9);
  const int o = (int)(i / (9LL * L.ci));
  const u16 v = f2bf(L.wf[i]);
  L.wb[(((long long)tap * (L.ci >> 4) + (c >> 4)) * L.co + o) * 16 + (c & 15)] = v;
  L.wd[(((long long)(8 - tap) * (L.co >> 4) + (o >> 4)) * L.ci + c) * 16 + (o & 15)] = v;
}

// one batch from the device dataset: features, targets and masks of
// samples idx[0..nb) (blockIdx.y = k) into the batch buffers
constexpr int MAXGATHER = 64;
struct GatherArgs {
  long long idx[MAXGATHER];
  long long ps;  // pixels per sample
  int cin;
  const float *dx, *dt;
  const unsigned char *dm;
  float *xin, *tgt;
  unsigned char *mask;
};
__global__ void __launch_bounds__(256) k_gather(const GatherArgs a) {
  const int k = blockIdx.y;
  const long long i = a.idx[k];
  const long long nx = a.ps * a.cin, nt = a.ps * 3, n = nx + nt + a.ps;
  for (long long e = blockIdx.x * 256LL + threadIdx.x; e < n; e += gridDim.x * 256LL) {
    if (e < nx) {
      a.xin[k * nx + e] = a.dx[i * nx + e];
    } else if (e < nx + nt) {
      const long long f = e - nx;
      a.tgt[k * nt + f] = a.dt[i * nt + f];
    } else {
      const long long f = e - nx - nt;
      a.mask[k * a.ps + f] = a.dm[i * a.ps + f];
    }
  }
}

__global__ void __launch_bounds__(256) k_fill(float *p, long long n, float v) {
  const long long i = blockIdx.x * 256LL + threadIdx.x;
  if (i < n) p[i] = v;
}

inline unsigned nblk(long long n) { return (unsigned)((n + 255) / 256); }
inline unsigned nblk32(long long n) { return (unsigned)((n + 31) / 32); }

}  // namespace sur

// ============================================================================
using namespace sur;

namespace {

struct Tensor {
  int level = 0, c = 0;
  bool relu = false;
  u16 *act = nullptr;  // bf16 [B][H_l][W_l][c]
  float *grad = nullptr;
};

struct Layer {
  int a = -1, s = -1;  // input tensors (a upsampled from the next coarser level), output tensor
  int out = -1;
  int ca = 0, cs = 0, co = 0;
  bool relu = true;
  long long woff = 0, boff = 0;  // offsets into the flat parameter array
  u16 *wb = nullptr, *wd = nullptr;
  bool last = false;
  // the input gradient of this layer is the first write into the s (a)
  // tensor's gradient in backward order: stored, not added
  bool s_first = false, a_first = false;
};

}  // namespace

struct lbm_sur {
  int B = 0, H = 0, W = 0, cin = 0;
  std::vector<Tensor> t;
  std::vector<Layer> L;
  std::vector<int> pools;  // tensor pairs (src, dst) of the pooling steps, in forward order
  std::vector<std::pair<int, int>> pool_pairs;
  std::vector<char> pool_first;  // pooling backward is the first write into the source's gradient
  long long nparam = 0;
  float *p = nullptr, *gp = nullptr, *m = nullptr, *v = nullptr;  // params, grads, Adam moments
  long long adam_t = 0;
  float *out = nullptr;   // [B][H][W][3]
  float *tgt = nullptr;   // [B][H][W][3]
  unsigned char *mask = nullptr;
  float *xin = nullptr;   // [B][H][W][cin]
  u16 *gout = nullptr;    // [B][H][W][32]
  float *part = nullptr;  // wgrad chunk partials
  long long part_cap = 0;
  float *gup = nullptr;   // full-resolution gradient of an upsampled input
  long long gup_cap = 0;
  float *grad_arena = nullptr;  // every Tensor::grad, B samples each
  long long grad_n = 0;
  u16 *gT = nullptr, *xT = nullptr;  // channel-major operands of k_wgrad_t (gT: two buffers)
  float *bpart = nullptr;            // bias partial sums of k_gy_prep (two buffers)
  long long gT_cap = 0, xT_cap = 0, bpart_cap = 0;
  double *lpart = nullptr, *lsum = nullptr;
  // dataset
  float *dx = nullptr, *dt = nullptr;
  unsigned char *dm = nullptr;
  long long dn = 0;
  std::vector<long long> dfluid;
  hipStream_t st = nullptr;
  // the weight gradients run on a second stream beside the input-gradient
  // chain: wg waits for ev_gy[k] (gT/bpart buffer k written), st for ev_wg[k]
  // (buffer k read) before writing it again, and for ev_wg of the last one
  // before the backward pass ends
  hipStream_t wg = nullptr;
  hipEvent_t ev_fwd = nullptr, ev_gy[2] = {nullptr, nullptr}, ev_wg[2] = {nullptr, nullptr};
  std::string err;
  bool ok_order = true;  // every gradient is written before it is read
};

static thread_local std::string g_sur_err;

static int sfail(lbm_sur *s, int code, const std::string &msg) {
  if (s) s->err = msg;
  g_sur_err = msg;
  return code;
}

#define SCHK(s, call)                                                                       \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess) return sfail((s), LBM_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

static long long npix_at(const lbm_sur *s, int level, int nb) {
  return (long long)nb * (s->H >> level) * (s->W >> level);
}

static int add_tensor(lbm_sur *s, int level, int c, bool relu) {
  Tensor T;
  T.level = level;
  T.c = c;
  T.relu = relu;
  s->t.push_back(T);
  return (int)s->t.size() - 1;
}

// network: six resolution levels (widths 32, 32, 64, 96, 128, 160; the
// coarsest level is 1/32 of the input, so the receptive field spans the
// whole channel), two 3x3 convolutions per level, 2x2 average pooling down,
// nearest upsampling and skip concatenation up, 3 linear outputs
constexpr int NLEV = 6;
constexpr int WIDTH[NLEV] = {32, 32, 64, 96, 128, 160};

static void build_net(lbm_sur *s) {
  const int x0 = add_tensor(s, 0, 16, false);
  auto conv = [&](int a, int sidx, int co, bool relu, bool last = false) {
    Layer l;
    l.a = a;
    l.s = sidx;
    l.ca = a >= 0 ? s->t[a].c : 0;
    l.cs = s->t[sidx].c;
    l.co = co;
    l.relu = relu;
    l.last = last;
    l.out = last ? -1 : add_tensor(s, s->t[sidx].level, co, relu);
    s->L.push_back(l);
    return l.out;
  };
  auto pool = [&](int src) {
    const int dst = add_tensor(s, s->t[src].level + 1, s->t[src].c, false);
    s->pool_pairs.push_back({src, dst});
    return dst;
  };
  int enc[NLEV];
  int cur = x0;
  for (int lv = 0; lv < NLEV; ++lv) {
    if (lv > 0) cur = pool(cur);
    cur = conv(-1, conv(-1, cur, WIDTH[lv], true), WIDTH[lv], true);
    enc[lv] = cur;
  }
  for (int lv = NLEV - 2; lv >= 0; --lv) cur = conv(-1, conv(cur, enc[lv], WIDTH[lv], true), WIDTH[lv], true);
  conv(-1, cur, 32, false, true);  // 3 used outputs, padded to one 32-channel tile
  // which write into each tensor gradient comes first in backward order
  // (the order of loss_backward): that one stores, the later ones add, so no
  // gradient is cleared before the backward pass; every gradient a layer
  // reads (its output's) has been written by then
  std::vector<char> seen(s->t.size(), 0);
  s->pool_first.assign(s->pool_pairs.size(), 0);
  size_t pi = s->pool_pairs.size();
  for (int li = (int)s->L.size() - 1; li >= 0; --li) {
    Layer &l = s->L[li];
    if (!l.last && !seen[l.out]) s->ok_order = false;
    if (li > 0) {
      l.s_first = !seen[l.s];
      seen[l.s] = 1;
      if (l.a >= 0) {
        l.a_first = !seen[l.a];
        seen[l.a] = 1;
      }
    }
    while (pi > 0 && s->pool_pairs[pi - 1].second == l.s) {
      --pi;
      const int src = s->pool_pairs[pi].first;
      s->pool_first[pi] = !seen[src];
      seen[src] = 1;
    }
  }
  long long off = 0;
  for (Layer &l : s->L) {
    l.woff = off;
    off += (long long)l.co * 9 * (l.ca + l.cs);
    l.boff = off;
    off += l.co;
  }
  s->nparam = off;
}

// the weight gradient runs on channel-major operands (k_wgrad_t) where the
// level's width is a multiple of 8 (a segment of 8 pixels stays in one row)
// and the channel counts are multiples of 8; else on k_wgrad
static bool wgrad_transposed(const Layer &l, int w) { return w % 8 == 0 && l.ca % 8 == 0 && l.cs % 8 == 0; }

// 8-pixel segments per k_wgrad_t chunk: up to 128 (1024 pixels), fewer on
// the small levels so that a launch has about 2048 waves, at least 8 (a
// wave's serial chain of loads and products stays short)
constexpr long long WG_WAVES = 2048;  // 1024 measured the same, 4096 7 % slower per training step
static int wgrad_cseg(long long np, int co, int ci) {
  const long long nseg = np / 8;
  const long long tiles = (long long)(co / 32) * ((ci + 31) / 32) * 3;
  long long c = (nseg * tiles + WG_WAVES - 1) / WG_WAVES;
  c = std::max(8LL, std::min(128LL, c));
  return (int)((c + 1) & ~1LL);
}

static int alloc_all(lbm_sur *s) {
  for (Tensor &T : s->t) {
    const long long n = npix_at(s, T.level, s->B) * T.c;
    SCHK(s, hipMalloc(&T.act, n * sizeof(u16)));
    s->grad_n += (n + 63) / 64 * 64;
  }
  for (Layer &l : s->L) {
    const long long n = (long long)l.co * 9 * (l.ca + l.cs);
    SCHK(s, hipMalloc(&l.wb, n * sizeof(u16)));
    SCHK(s, hipMalloc(&l.wd, n * sizeof(u16)));
    const long long np = npix_at(s, s->t[l.s].level, s->B);
    const long long nch = (np + WGCHUNK - 1) / WGCHUNK;  // WCHUNK >= WGCHUNK: covers the bias partials
    s->part_cap = std::max(s->part_cap, nch * n);
    if (l.ca > 0) s->gup_cap = std::max(s->gup_cap, np * l.ca);
    const int h = s->H >> s->t[l.s].level, w = s->W >> s->t[l.s].level;
    if (wgrad_transposed(l, w)) {
      for (int nb = 1; nb <= s->B; ++nb) {  // the chunking depends on the batch in use
        const long long npb = npix_at(s, s->t[l.s].level, nb);
        const long long cseg = wgrad_cseg(npb, l.co, l.ca + l.cs);
        const long long nch = ((npb / 8 + cseg - 1) / cseg + 3) / 4 * 4;  // chunks, a multiple of 4
        s->part_cap = std::max(s->part_cap, nch / 4 * n);
      }
      s->gT_cap = std::max(s->gT_cap, np * l.co);
      s->xT_cap = std::max(s->xT_cap, (long long)(l.ca + l.cs) * s->B * (h + 2) * (w + 8));
      s->bpart_cap = std::max(s->bpart_cap, (np + TP - 1) / TP * l.co);
    }
  }
  // one arena for every tensor's gradient (each written by its first writer
  // in backward order before anything reads it)
  SCHK(s, hipMalloc(&s->grad_arena, s->grad_n * sizeof(float)));
  {
    long long off = 0;
    for (Tensor &T : s->t) {
      T.grad = s->grad_arena + off;
      off += (npix_at(s, T.level, s->B) * T.c + 63) / 64 * 64;
    }
  }
  const long long np0 = npix_at(s, 0, s->B);
  SCHK(s, hipMalloc(&s->p, s->nparam * sizeof(float)));
  SCHK(s, hipMalloc(&s->gp, s->nparam * sizeof(float)));
  SCHK(s, hipMalloc(&s->m, s->nparam * sizeof(float)));
  SCHK(s, hipMalloc(&s->v, s->nparam * sizeof(float)));
  SCHK(s, hipMalloc(&s->out, np0 * 3 * sizeof(float)));
  SCHK(s, hipMalloc(&s->tgt, np0 * 3 * sizeof(float)));
  SCHK(s, hipMalloc(&s->mask, np0));
  SCHK(s, hipMalloc(&s->xin, np0 * s->cin * sizeof(float)));
  SCHK(s, hipMalloc(&s->gout, np0 * 32 * sizeof(u16)));
  SCHK(s, hipMalloc(&s->part, s->part_cap * sizeof(float)));
  SCHK(s, hipMalloc(&s->gup, std::max(1LL, s->gup_cap) * sizeof(float)));
  SCHK(s, hipMalloc(&s->gT, 2 * std::max(1LL, s->gT_cap) * sizeof(u16)));
  SCHK(s, hipMalloc(&s->xT, std::max(1LL, s->xT_cap) * sizeof(u16)));
  SCHK(s, hipMalloc(&s->bpart, 2 * std::max(1LL, s->bpart_cap) * sizeof(float)));
  SCHK(s, hipMalloc(&s->lpart, ((np0 + 255) / 256) * sizeof(double)));
  SCHK(s, hipMalloc(&s->lsum, sizeof(double)));
  return LBM_OK;
}

static int prep_weights(lbm_sur *s) {
  if ((int)s->L.size() > MAXLAYERS) return sfail(s, LBM_EINVAL, "lbm_sur: too many layers");
  PrepAll pa;
  long long nmax = 0;
  pa.n = (int)s->L.size();
  for (int i = 0; i < pa.n; ++i) {
    const Layer &l = s->L[i];
    pa.l[i] = PrepLayer{s->p + l.woff, l.co, l.ca + l.cs, l.wb, l.wd};
    nmax = std::max(nmax, (long long)l.co * 9 * (l.ca + l.cs));
  }
  hipLaunchKernelGGL(k_prep_all, dim3(nblk(nmax), pa.n), dim3(256), 0, s->st, pa);
  SCHK(s, hipGetLastError());
  return LBM_OK;
}

// waves per output tile of k_conv: the small levels split K over 2 or 4
// waves so that a launch keeps enough waves in flight
static int conv_ksplit(long long npix, int co) {
  const long long tiles = (npix + 31) / 32 * (co / 32);
  return tiles >= 8192 ? 1 : tiles >= 2048 ? 2 : 4;
}
static unsigned conv_blocks(long long npix, int co) {
  const long long tiles = (npix + 31) / 32 * (co / 32);
  const int per = 4 / conv_ksplit(npix, co);
  return (unsigned)((tiles + per - 1) / per);
}

// the large levels (at least 65536 pixels, width a multiple of 32, at most
// LDS_CI_MAX input channels) run k_conv_lds
// (16384 pixels measured the same per training step, 4096 3 % slower)
static bool conv_lds(long long npix, int w, int ci) { return npix >= 65536 && w % 32 == 0 && ci <= LDS_CI_MAX; }

static void launch_conv(int mode, const ConvArgs &a, hipStream_t st) {
  const long long np = (long long)a.nb * a.h * a.wd;
  const int ci = a.ca + a.cs;
  if (conv_lds(np, a.wd, ci)) {
    const dim3 grid((unsigned)(a.nb * ((a.h + 3) / 4) * (a.wd / 32)), (unsigned)(a.co / 32));
    const size_t lds = (size_t)6 * 34 * (ci + 8) * sizeof(u16);
    if (mode == 0) hipLaunchKernelGGL(k_conv_lds<0>, grid, dim3(256), lds, st, a);
    else if (mode == 1) hipLaunchKernelGGL(k_conv_lds<1>, grid, dim3(256), lds, st, a);
    else hipLaunchKernelGGL(k_conv_lds<2>, grid, dim3(256), lds, st, a);
    return;
  }
  const int ks = conv_ksplit(np, a.co);
  const dim3 grid(conv_blocks(np, a.co));
  if (mode == 0) hipLaunchKernelGGL(k_conv<0>, grid, dim3(256), 0, st, a, ks);
  else if (mode == 1) hipLaunchKernelGGL(k_conv<1>, grid, dim3(256), 0, st, a, ks);
  else hipLaunchKernelGGL(k_conv<2>, grid, dim3(256), 0, st, a, ks);
}

// forward of nb <= B samples whose features are in s->xin
static int forward(lbm_sur *s, int nb) {
  const long long np0 = npix_at(s, 0, nb);
  hipLaunchKernelGGL(k_cast_in, dim3(nblk(np0 * 16)), dim3(256), 0, s->st, (const float *)s->xin, s->cin,
                     s->t[0].act, np0);
  size_t pi = 0;
  for (const Layer &l : s->L) {
    // pooling steps whose output this layer reads come first
    while (pi < s->pool_pairs.size() && s->pool_pairs[pi].second == l.s) {
      const Tensor &src = s->t[s->pool_pairs[pi].first];
      const Tensor &dst = s->t[s->pool_pairs[pi].second];
      const int h = s->H >> dst.level, w = s->W >> dst.level;
      hipLaunchKernelGGL(k_pool, dim3(nblk((long long)nb * h * w * dst.c)), dim3(256), 0, s->st,
                         (const u16 *)src.act, dst.act, nb, h, w, dst.c);
      ++pi;
    }
    const Tensor &S = s->t[l.s];
    ConvArgs a;
    memset(&a, 0, sizeof(a));
    a.a = l.a >= 0 ? s->t[l.a].act : nullptr;
    a.s = S.act;
    a.ca = l.ca;
    a.cs = l.cs;
    a.w = l.wb;
    a.bias = s->p + l.boff;
    a.nb = nb;
    a.h = s->H >> S.level;
    a.wd = s->W >> S.level;
    a.co = l.co;
    const long long np = (long long)nb * a.h * a.wd;
    if (l.last) {
      a.yf = s->out;
      a.cf = 3;
      launch_conv(2, a, s->st);
    } else {
      a.y = s->t[l.out].act;
      a.relu = l.relu;
      launch_conv(0, a, s->st);
    }
  }
  SCHK(s, hipGetLastError());
  return LBM_OK;
}

// loss of nb samples (targets and masks already on the device) and, when
// grads is set, backward into s->gp
static int loss_backward(lbm_sur *s, int nb, long long nfluid, bool grads) {
  const long long np0 = npix_at(s, 0, nb);
  const float scale = 1.f / (3.f * (float)std::max(1LL, nfluid));
  const int nbl = (int)nblk(np0);
  hipLaunchKernelGGL(k_loss, dim3(nbl), dim3(256), 0, s->st, (const float *)s->out, (const float *)s->tgt,
                     (const unsigned char *)s->mask, np0, scale, s->gout, s->lpart);
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(256), 0, s->st, (const double *)s->lpart, nbl, s->lsum);
  SCHK(s, hipGetLastError());
  if (!grads) return LBM_OK;
  size_t pi = s->pool_pairs.size();
  // the activations of the forward pass are read on the weight-gradient stream
  SCHK(s, hipEventRecord(s->ev_fwd, s->st));
  SCHK(s, hipStreamWaitEvent(s->wg, s->ev_fwd, 0));
  bool pending[2] = {false, false};
  int ntr = 0, last = -1;
  for (int li = (int)s->L.size() - 1; li >= 0; --li) {
    const Layer &l = s->L[li];
    const Tensor &S = s->t[l.s];
    const int h = s->H >> S.level, w = s->W >> S.level;
    const long long np = (long long)nb * h * w;
    const int ci = l.ca + l.cs;
    // gradient of this layer's output, bf16 with the relu mask applied
    // (gout holds it), and the weight and bias gradients
    const u16 *gy = s->gout;
    const long long nw = (long long)l.co * 9 * ci;
    const int nchunks = (int)((np + WGCHUNK - 1) / WGCHUNK);
    if (wgrad_transposed(l, w)) {
      const Tensor &O = s->t[l.last ? l.s : l.out];
      const unsigned ptiles = (unsigned)((np + TP - 1) / TP);
      const int k = ntr++ & 1;
      u16 *gT = s->gT + k * std::max(1LL, s->gT_cap);
      float *bpart = s->bpart + k * std::max(1LL, s->bpart_cap);
      if (pending[k]) SCHK(s, hipStreamWaitEvent(s->st, s->ev_wg[k], 0));
      hipLaunchKernelGGL(k_gy_prep, dim3(ptiles, l.co / 32), dim3(256), 0, s->st, l.last ? nullptr : (const float *)O.grad,
                         l.last ? nullptr : (const u16 *)O.act, (!l.last && O.relu) ? 1 : 0,
                         l.last ? (const u16 *)s->gout : nullptr, s->gout, gT, bpart, np, l.co);
      SCHK(s, hipEventRecord(s->ev_gy[k], s->st));
      hipLaunchKernelGGL(k_x_prep, dim3((unsigned)((w + 8 + TP - 1) / TP), (unsigned)(nb * (h + 2)), (unsigned)((ci + 31) / 32)),
                         dim3(256), 0, s->wg, l.a >= 0 ? (const u16 *)s->t[l.a].act : nullptr, l.ca,
                         (const u16 *)S.act, l.cs, nb, h, w, s->xT);
      SCHK(s, hipStreamWaitEvent(s->wg, s->ev_gy[k], 0));
      WgradTArgs wa;
      wa.gT = gT;
      wa.xT = s->xT;
      wa.P = np;
      wa.co = l.co;
      wa.ci = ci;
      wa.nb = nb;
      wa.h = h;
      wa.w = w;
      wa.cseg = wgrad_cseg(np, l.co, ci);
      wa.nchunks = (int)(((np / 8 + wa.cseg - 1) / wa.cseg + 3) / 4 * 4);
      wa.part = s->part;
      const long long waves = (long long)(l.co / 32) * ((ci + 31) / 32) * 3 * wa.nchunks;
      hipLaunchKernelGGL(k_wgrad_t, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, s->wg, wa);
      hipLaunchKernelGGL(k_reduce_wb, dim3(nblk32(nw) + l.co), dim3(256), 0, s->wg, (const float *)s->part, nw,
                         wa.nchunks / 4, s->gp + l.woff, (const float *)bpart, (long long)ptiles, l.co, s->gp + l.boff);
      SCHK(s, hipEventRecord(s->ev_wg[k], s->wg));
      pending[k] = true;
      last = k;
    } else {
      // s->part is the weight-gradient stream's too
      if (last >= 0) SCHK(s, hipStreamWaitEvent(s->st, s->ev_wg[last], 0));
      if (!l.last) {
        Tensor &O = s->t[l.out];
        hipLaunchKernelGGL(k_mask_cast, dim3(nblk(np * l.co)), dim3(256), 0, s->st, (const float *)O.grad,
                           (const u16 *)O.act, s->gout, np * l.co, O.relu ? 1 : 0);
      }
      WgradArgs wa;
      wa.gy = gy;
      wa.a = l.a >= 0 ? s->t[l.a].act : nullptr;
      wa.s = S.act;
      wa.ca = l.ca;
      wa.cs = l.cs;
      wa.co = l.co;
      wa.nb = nb;
      wa.h = h;
      wa.wd = w;
      wa.nchunks = nchunks;
      wa.part = s->part;
      const long long waves = (long long)(l.co / 32) * ((ci + 31) / 32) * 9 * nchunks;
      hipLaunchKernelGGL(k_wgrad, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, s->st, wa);
      hipLaunchKernelGGL(k_reduce_chunks, dim3(nblk32(nw)), dim3(256), 0, s->st, (const float *)s->part, nw, nchunks,
                         s->gp + l.woff);
      const int bchunks = (int)((np + WCHUNK - 1) / WCHUNK);
      hipLaunchKernelGGL(k_bias_part, dim3(bchunks), dim3(256), 0, s->st, gy, np, l.co, s->part);
      hipLaunchKernelGGL(k_reduce_chunks, dim3(nblk32(l.co)), dim3(256), 0, s->st, (const float *)s->part,
                         (long long)l.co, bchunks, s->gp + l.boff);
    }
    // input gradient (not for the network input)
    if (li > 0) {
      ConvArgs a;
      memset(&a, 0, sizeof(a));
      a.s = gy;
      a.cs = l.co;
      a.w = l.wd;
      a.nb = nb;
      a.h = h;
      a.wd = w;
      a.co = ci;
      a.c0 = l.ca;
      a.g0 = l.ca > 0 ? s->gup : nullptr;
      a.g1 = S.grad;
      a.add1 = l.s_first ? 0 : 1;
      launch_conv(1, a, s->st);
      if (l.ca > 0) {
        Tensor &A = s->t[l.a];
        hipLaunchKernelGGL(k_up_bwd, dim3(nblk(np / 4 * l.ca)), dim3(256), 0, s->st, (const float *)s->gup, A.grad,
                           nb, h / 2, w / 2, l.ca, l.a_first ? 0 : 1);
      }
    }
    // pooling steps that produced this layer's input run backward next
    while (pi > 0 && s->pool_pairs[pi - 1].second == l.s) {
      --pi;
      Tensor &src = s->t[s->pool_pairs[pi].first];
      const Tensor &dst = s->t[s->pool_pairs[pi].second];
      const int hh = s->H >> dst.level, ww = s->W >> dst.level;
      hipLaunchKernelGGL(k_pool_bwd, dim3(nblk((long long)nb * 4 * hh * ww * dst.c)), dim3(256), 0, s->st,
                         (const float *)dst.grad, src.grad, nb, hh, ww, dst.c, s->pool_first[pi] ? 0 : 1);
    }
  }
  if (last >= 0) SCHK(s, hipStreamWaitEvent(s->st, s->ev_wg[last], 0));
  SCHK(s, hipGetLastError());
  return LBM_OK;
}

static int check_n(lbm_sur *s, int n) {
  if (n < 1 || n > s->B) return sfail(s, LBM_EINVAL, "lbm_sur: n must be in 1..batch");
  return LBM_OK;
}

extern "C" {

int lbm_sur_create(int32_t batch, int32_t height, int32_t width, int32_t in_channels, uint64_t seed, lbm_sur **out) {
  if (!out || batch < 1 || height < 32 || width < 32 || height % 32 || width % 32 || in_channels < 1 ||
      in_channels > 16)
    return sfail(nullptr, LBM_EINVAL, "lbm_sur_create: batch >= 1, height and width multiples of 32, in_channels 1..16");
  lbm_sur *s = new lbm_sur();
  s->B = batch;
  s->H = height;
  s->W = width;
  s->cin = in_channels;
  build_net(s);
  auto bail = [&](int rc) {
    lbm_sur_destroy(s);
    return rc;
  };
  if (!s->ok_order) return bail(sfail(nullptr, LBM_EINVAL, "lbm_sur_create: a gradient is read before it is written"));
  if (hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&s->wg, hipStreamNonBlocking) != hipSuccess)
    return bail(sfail(nullptr, LBM_EHIP, "lbm_sur_create: stream"));
  for (hipEvent_t *e : {&s->ev_fwd, &s->ev_gy[0], &s->ev_gy[1], &s->ev_wg[0], &s->ev_wg[1]})
    if (hipEventCreateWithFlags(e, hipEventDisableTiming) != hipSuccess)
      return bail(sfail(nullptr, LBM_EHIP, "lbm_sur_create: event"));
  int rc = alloc_all(s);
  if (rc) return bail(sfail(nullptr, rc, s->err));
  // He-normal weights, zero biases; the output layer starts 10x smaller
  std::vector<float> p(s->nparam, 0.f);
  std::mt19937_64 gen(seed);
  for (const Layer &l : s->L) {
    const int ci = l.ca + l.cs;
    std::normal_distribution<float> nd(0.f, std::sqrt(2.f / (9.f * ci)) * (l.last ? 0.1f : 1.f));
    for (long long i = 0; i < (long long)l.co * 9 * ci; ++i) p[l.woff + i] = nd(gen);
    if (l.last)  // only 3 of the 32 output channels are used
      for (int o = 3; o < l.co; ++o)
        for (int k = 0; k < 9 * ci; ++k) p[l.woff + (long long)o * 9 * ci + k] = 0.f;
  }
  rc = lbm_sur_set_params(s, p.data());
  if (rc) return bail(rc);
  *out = s;
  return LBM_OK;
}

int lbm_sur_param_count(const lbm_sur *s, int64_t *n) {
  if (!s || !n) return sfail(nullptr, LBM_EINVAL, "lbm_sur_param_count: null argument");
  *n = s->nparam;
  return LBM_OK;
}

int lbm_sur_get_params(lbm_sur *s, float *params) {
  if (!s || !params) return sfail(s, LBM_EINVAL, "lbm_sur_get_params: null argument");
  SCHK(s, hipMemcpyAsync(params, s->p, s->nparam * sizeof(float), hipMemcpyDeviceToHost, s->st));
  SCHK(s, hipStreamSynchronize(s->st));
  return LBM_OK;
}

int lbm_sur_set_params(lbm_sur *s, const float *params) {
  if (!s || !params) return sfail(s, LBM_EINVAL, "lbm_sur_set_params: null argument");
  SCHK(s, hipMemcpyAsync(s->p, params, s->nparam * sizeof(float), hipMemcpyHostToDevice, s->st));
  hipLaunchKernelGGL(k_fill, dim3(nblk(s->nparam)), dim3(256), 0, s->st, s->m, s->nparam, 0.f);
  hipLaunchKernelGGL(k_fill, dim3(nblk(s->nparam)), dim3(256), 0, s->st, s->v, s->nparam, 0.f);
  s->adam_t = 0;
  int rc = prep_weights(s);
  if (rc) return rc;
  SCHK(s, hipStreamSynchronize(s->st));
  return LBM_OK;
}

int lbm_sur_forward(lbm_sur *s, const float *x, int32_t n, float *y) {
  if (!s || !x || !y) return sfail(s, LBM_EINVAL, "lbm_sur_forward: null argument");
  int rc = check_n(s, n);
  if (rc) return rc;
  const long long np0 = npix_at(s, 0, n);
  SCHK(s, hipMemcpyAsync(s->xin, x, np0 * s->cin * sizeof(float), hipMemcpyHostToDevice, s->st));
  rc = forward(s, n);
  if (rc) return rc;
  SCHK(s, hipMemcpyAsync(y, s->out, np0 * 3 * sizeof(float), hipMemcpyDeviceToHost, s->st));
  SCHK(s, hipStreamSynchronize(s->st));
  return LBM_OK;
}

int lbm_sur_loss_grad(lbm_sur *s, const float *x, const float *t, const uint8_t *mask, int32_t n, double *loss,
                      float *grad) {
  if (!s || !x || !t || !mask || !loss) return sfail(s, LBM_EINVAL, "lbm_sur_loss_grad: null argument");
  int rc = check_n(s, n);
  if (rc) return rc;
  const long long np0 = npix_at(s, 0, n);
  long long nfl = 0;
  for (long long i = 0; i < np0; ++i) nfl += mask[i] != 0;
  SCHK(s, hipMemcpyAsync(s->xin, x, np0 * s->cin * sizeof(float), hipMemcpyHostToDevice, s->st));
  SCHK(s, hipMemcpyAsync(s->tgt, t, np0 * 3 * sizeof(float), hipMemcpyHostToDevice, s->st));
  SCHK(s, hipMemcpyAsync(s->mask, mask, np0, hipMemcpyHostToDevice, s->st));
  rc = forward(s, n);
  if (rc) return rc;
  rc = loss_backward(s, n, nfl, grad != nullptr);
  if (rc) return rc;
  double l = 0;
  SCHK(s, hipMemcpyAsync(&l, s->lsum, sizeof(double), hipMemcpyDeviceToHost, s->st));
  if (grad) SCHK(s, hipMemcpyAsync(grad, s->gp, s->nparam * sizeof(float), hipMemcpyDeviceToHost, s->st));
  SCHK(s, hipStreamSynchronize(s->st));
  *loss = l / (3.0 * (double)std::max(1LL, nfl));
  return LBM_OK;
}

int lbm_sur_set_dataset(lbm_sur *s, const float *x, const float *t, const uint8_t *mask, int64_t n) {
  if (!s || !x || !t || !mask || n < 1) return sfail(s, LBM_EINVAL, "lbm_sur_set_dataset: bad argument");
  if (s->dx) (void)hipFree(s->dx);
  if (s->dt) (void)hipFree(s->dt);
  if (s->dm) (void)hipFree(s->dm);
  s->dx = nullptr;
  s->dt = nullptr;
  s->dm = nullptr;
  const long long ps = (long long)s->H * s->W;
  SCHK(s, hipMalloc(&s->dx, n * ps * s->cin * sizeof(float)));
  SCHK(s, hipMalloc(&s->dt, n * ps * 3 * sizeof(float)));
  SCHK(s, hipMalloc(&s->dm, n * ps));
  SCHK(s, hipMemcpy(s->dx, x, n * ps * s->cin * sizeof(float), hipMemcpyHostToDevice));
  SCHK(s, hipMemcpy(s->dt, t, n * ps * 3 * sizeof(float), hipMemcpyHostToDevice));
  SCHK(s, hipMemcpy(s->dm, mask, n * ps, hipMemcpyHostToDevice));
  s->dn = n;
  s->dfluid.assign(n, 0);
  for (long long i = 0; i < n; ++i) {
    long long c = 0;
    for (long long k = 0; k < ps; ++k) c += mask[i * ps + k] != 0;
    s->dfluid[i] = c;
  }
  return LBM_OK;
}

static int gather_batch(lbm_sur *s, const int64_t *idx, int nb, long long *nfl) {
  const long long ps = (long long)s->H * s->W;
  *nfl = 0;
  if (nb <= MAXGATHER) {  // one launch
    GatherArgs g;
    for (int k = 0; k < nb; ++k) {
      if (idx[k] < 0 || idx[k] >= s->dn) return sfail(s, LBM_EINVAL, "lbm_sur: dataset index out of range");
      g.idx[k] = idx[k];
      *nfl += s->dfluid[idx[k]];
    }
    g.ps = ps;
    g.cin = s->cin;
    g.dx = s->dx;
    g.dt = s->dt;
    g.dm = s->dm;
    g.xin = s->xin;
    g.tgt = s->tgt;
    g.mask = s->mask;
    hipLaunchKernelGGL(k_gather, dim3(256, nb), dim3(256), 0, s->st, g);
    SCHK(s, hipGetLastError());
    return LBM_OK;
  }
  for (int k = 0; k < nb; ++k) {
    const long long i = idx[k];
    if (i < 0 || i >= s->dn) return sfail(s, LBM_EINVAL, "lbm_sur: dataset index out of range");
    SCHK(s, hipMemcpyAsync(s->xin + k * ps * s->cin, s->dx + i * ps * s->cin, ps * s->cin * sizeof(float),
                           hipMemcpyDeviceToDevice, s->st));
    SCHK(s, hipMemcpyAsync(s->tgt + k * ps * 3, s->dt + i * ps * 3, ps * 3 * sizeof(float), hipMemcpyDeviceToDevice,
                           s->st));
    SCHK(s, hipMemcpyAsync(s->mask + k * ps, s->dm + i * ps, ps, hipMemcpyDeviceToDevice, s->st));
    *nfl += s->dfluid[i];
  }
  return LBM_OK;
}

int lbm_sur_train(lbm_sur *s, const int64_t *idx, int32_t nsteps, float lr, double *losses) {
  if (!s || !idx || nsteps < 1 || !(lr > 0)) return sfail(s, LBM_EINVAL, "lbm_sur_train: bad argument");
  if (!s->dn) return sfail(s, LBM_EINVAL, "lbm_sur_train: no dataset");
  for (int k = 0; k < nsteps; ++k) {
    long long nfl = 0;
    int rc = gather_batch(s, idx + (long long)k * s->B, s->B, &nfl);
    if (rc) return rc;
    rc = forward(s, s->B);
    if (rc) return rc;
    rc = loss_backward(s, s->B, nfl, true);
    if (rc) return rc;
    s->adam_t++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)s->adam_t), c2 = 1.f - std::pow(b2, (float)s->adam_t);
    hipLaunchKernelGGL(k_adam, dim3(nblk(s->nparam)), dim3(256), 0, s->st, s->p, (const float *)s->gp, s->m, s->v,
                       s->nparam, lr, b1, b2, 1e-8f, c1, c2);
    rc = prep_weights(s);
    if (rc) return rc;
    if (losses) {
      double l = 0;
      SCHK(s, hipMemcpyAsync(&l, s->lsum, sizeof(double), hipMemcpyDeviceToHost, s->st));
      SCHK(s, hipStreamSynchronize(s->st));
      losses[k] = l / (3.0 * (double)std::max(1LL, nfl));
    }
  }
  SCHK(s, hipStreamSynchronize(s->st));
  return LBM_OK;
}

int lbm_sur_predict_dataset(lbm_sur *s, int64_t i0, int32_t n, float *y) {
  if (!s || !y) return sfail(s, LBM_EINVAL, "lbm_sur_predict_dataset: null argument");
  int rc = check_n(s, n);
  if (rc) return rc;
  if (i0 < 0 || i0 + n > s->dn) return sfail(s, LBM_EINVAL, "lbm_sur_predict_dataset: range outside the dataset");
  std::vector<int64_t> idx(n);
  for (int k = 0; k < n; ++k) idx[k] = i0 + k;
  long long nfl = 0;
  rc = gather_batch(s, idx.data(), n, &nfl);
  if (rc) return rc;
  rc = forward(s, n);
  if (rc) return rc;
  SCHK(s, hipMemcpyAsync(y, s->out, npix_at(s, 0, n) * 3 * sizeof(float), hipMemcpyDeviceToHost, s->st));
  SCHK(s, hipStreamSynchronize(s->st));
  return LBM_OK;
}

const char *lbm_sur_last_error(const lbm_sur *s) { return s ? s->err.c_str() : g_sur_err.c_str(); }

int lbm_sur_destroy(lbm_sur *s) {
  if (!s) return LBM_OK;
  if (s->st) (void)hipStreamSynchronize(s->st);
  if (s->wg) (void)hipStreamSynchronize(s->wg);
  for (Tensor &T : s->t) {
    if (T.act) (void)hipFree(T.act);
    T.grad = nullptr;  // in the arena
  }
  for (Layer &l : s->L) {
    if (l.wb) (void)hipFree(l.wb);
    if (l.wd) (void)hipFree(l.wd);
  }
  void *ptrs[] = {s->p, s->gp, s->m, s->v, s->out, s->tgt, s->mask, s->xin, s->gout, s->part, s->gup,
                  s->gT, s->xT, s->bpart, s->grad_arena, s->lpart, s->lsum, s->dx, s->dt, s->dm};
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  for (hipEvent_t e : {s->ev_fwd, s->ev_gy[0], s->ev_gy[1], s->ev_wg[0], s->ev_wg[1]})
    if (e) (void)hipEventDestroy(e);
  if (s->st) (void)hipStreamDestroy(s->st);
  if (s->wg) (void)hipStreamDestroy(s->wg);
  delete s;
  return LBM_OK;
}

}  // extern "C"
