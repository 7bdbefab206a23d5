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
  const u16 *w;  // [co][9][ca + cs]
  const float *bias;
  int nb, h, wd, co;
  // MODE 0: y = bf16(act(acc + bias)), [p][co]
  u16 *y;
  int relu;
  // MODE 1: channels < c0 accumulate into g0 [p][c0], the rest into g1 [p][co - c0]
  float *g0, *g1;
  int c0;
  // MODE 2: yf = acc + bias for the first cf channels, [p][cf]
  float *yf;
  int cf;
};

template <int MODE>
__global__ void __launch_bounds__(256) k_conv(const ConvArgs a) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long long npix = (long long)a.nb * a.h * a.wd;
  const int ntc = a.co >> 5;
  const long long tile = (long long)blockIdx.x * 4 + wave;
  const long long p0 = (tile / ntc) * 32;
  if (p0 >= npix) return;  // wave-uniform; no block barrier in this kernel
  const int co0 = (int)(tile % ntc) * 32;
  const int ci = a.ca + a.cs;
  const int m = lane & 31, g = lane >> 5;
  const long long p = p0 + m;
  const bool pin = p < npix;
  int b = 0, y = 0, x = 0;
  if (pin) {
    x = (int)(p % a.wd);
    y = (int)((p / a.wd) % a.h);
    b = (int)(p / ((long long)a.wd * a.h));
  }
  const u16 *wrow = a.w + (long long)(co0 + m) * 9 * ci + 8 * g;
  f32x16 acc;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.f;
  const int h2 = a.h >> 1, w2 = a.wd >> 1;
  for (int tap = 0; tap < 9; ++tap) {
    const int ys = y + tap / 3 - 1, xs = x + tap % 3 - 1;
    const bool ok = pin && ys >= 0 && ys < a.h && xs >= 0 && xs < a.wd;
    const u16 *wt = wrow + tap * ci;
    if (a.ca > 0) {
      const u16 *src = a.a + (((long long)b * h2 + (ys >> 1)) * w2 + (xs >> 1)) * a.ca + 8 * g;
      for (int c = 0; c < a.ca; c += 16) {
        bf16x8 av;
        if (ok) av = ld8(src + c);
        else av = __builtin_bit_cast(bf16x8, u16x8{0, 0, 0, 0, 0, 0, 0, 0});
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, ld8(wt + c), acc, 0, 0, 0);
      }
    }
    const u16 *src = a.s + (((long long)b * a.h + ys) * a.wd + xs) * a.cs + 8 * g;
    for (int c = 0; c < a.cs; c += 16) {
      bf16x8 av;
      if (ok) av = ld8(src + c);
      else av = __builtin_bit_cast(bf16x8, u16x8{0, 0, 0, 0, 0, 0, 0, 0});
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, ld8(wt + a.ca + c), acc, 0, 0, 0);
    }
  }
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
      if (co_ < a.c0) a.g0[pp * a.c0 + co_] += v;
      else a.g1[pp * (a.co - a.c0) + (co_ - a.c0)] += v;
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

// dst[i] = sum over chunks of part[chunk][i], chunks in order
__global__ void __launch_bounds__(256) k_reduce_chunks(const float *part, long long n, int nchunks, float *dst) {
  const long long i = blockIdx.x * 256LL + threadIdx.x;
  if (i >= n) return;
  float s = 0.f;
  int c = 0;
  // eight loads in flight, added in chunk order
  for (; c + 8 <= nchunks; c += 8) {
    float v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = part[(long long)(c + k) * n + i];
#pragma unroll
    for (int k = 0; k < 8; ++k) s += v[k];
  }
  for (; c < nchunks; ++c) s += part[(long long)c * n + i];
  dst[i] = s;
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

// gx (full resolution, size 2h x 2w) += gy / 4 at the parent cell
__global__ void __launch_bounds__(256) k_pool_bwd(const float *gy, float *gx, int nb, int h, int w, int c) {
  const long long i = blockIdx.x * 256LL + threadIdx.x;
  const long long n = (long long)nb * 2 * h * 2 * w * c;
  if (i >= n) return;
  const int ch = (int)(i % c);
  long long p = i / c;
  const int xx = (int)(p % (2 * w)), yy = (int)((p / (2 * w)) % (2 * h)), b = (int)(p / (4LL * w * h));
  gx[i] += 0.25f * gy[(((long long)b * h + (yy >> 1)) * w + (xx >> 1)) * c + ch];
}

// ga (half resolution h x w) += sum of the 2x2 children of gup (nearest
// upsampling backward)
__global__ void __launch_bounds__(256) k_up_bwd(const float *gup, float *ga, int nb, int h, int w, int c) {
  const long long i = blockIdx.x * 256LL + threadIdx.x;
  const long long n = (long long)nb * h * w * c;
  if (i >= n) return;
  const int ch = (int)(i % c);
  long long p = i / c;
  const int xx = (int)(p % w), yy = (int)((p / w) % h), b = (int)(p / ((long long)w * h));
  const long long base = (((long long)b * 2 * h + 2 * yy) * 2 * w + 2 * xx) * c + ch;
  ga[i] += gup[base] + gup[base + c] + gup[base + 2LL * w * c] + gup[base + 2LL * w * c + c];
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

// bf16 working copies of one layer: wb[co][tap][ci] and the dgrad form
// wd[ci][8 - tap][co]
__global__ void __launch_bounds__(256) k_prep_w(const float *wf, int co, int ci, u16 *wb, u16 *wd) {
  const long long i = blockIdx.x * 256LL + threadIdx.x;
  const long long n = (long long)co * 9 * ci;
  if (i >= n) return;
  const int c = (int)(i % ci);
  const int tap = (int)((i / ci) % 9);
  const int o = (int)(i / (9LL * ci));
  const u16 v = f2bf(wf[i]);
  wb[i] = v;
  if (wd) wd[((long long)c * 9 + (8 - tap)) * co + o] = v;
}

__global__ void __launch_bounds__(256) k_fill(float *p, long long n, float v) {
  const long long i = blockIdx.x * 256LL + threadIdx.x;
  if (i < n) p[i] = v;
}

inline unsigned nblk(long long n) { return (unsigned)((n + 255) / 256); }

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
};

}  // namespace

struct lbm_sur {
  int B = 0, H = 0, W = 0, cin = 0;
  std::vector<Tensor> t;
  std::vector<Layer> L;
  std::vector<int> pools;  // tensor pairs (src, dst) of the pooling steps, in forward order
  std::vector<std::pair<int, int>> pool_pairs;
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
  double *lpart = nullptr, *lsum = nullptr;
  // dataset
  float *dx = nullptr, *dt = nullptr;
  unsigned char *dm = nullptr;
  long long dn = 0;
  std::vector<long long> dfluid;
  hipStream_t st = nullptr;
  std::string err;
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
  long long off = 0;
  for (Layer &l : s->L) {
    l.woff = off;
    off += (long long)l.co * 9 * (l.ca + l.cs);
    l.boff = off;
    off += l.co;
  }
  s->nparam = off;
}

static int alloc_all(lbm_sur *s) {
  for (Tensor &T : s->t) {
    const long long n = npix_at(s, T.level, s->B) * T.c;
    SCHK(s, hipMalloc(&T.act, n * sizeof(u16)));
    SCHK(s, hipMalloc(&T.grad, n * sizeof(float)));
  }
  for (Layer &l : s->L) {
    const long long n = (long long)l.co * 9 * (l.ca + l.cs);
    SCHK(s, hipMalloc(&l.wb, n * sizeof(u16)));
    SCHK(s, hipMalloc(&l.wd, n * sizeof(u16)));
    const long long np = npix_at(s, s->t[l.s].level, s->B);
    const long long nch = (np + WGCHUNK - 1) / WGCHUNK;  // WCHUNK >= WGCHUNK: covers the bias partials
    s->part_cap = std::max(s->part_cap, nch * n);
    if (l.ca > 0) s->gup_cap = std::max(s->gup_cap, np * l.ca);
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
  SCHK(s, hipMalloc(&s->lpart, ((np0 + 255) / 256) * sizeof(double)));
  SCHK(s, hipMalloc(&s->lsum, sizeof(double)));
  return LBM_OK;
}

static int prep_weights(lbm_sur *s) {
  for (Layer &l : s->L) {
    const long long n = (long long)l.co * 9 * (l.ca + l.cs);
    hipLaunchKernelGGL(k_prep_w, dim3(nblk(n)), dim3(256), 0, s->st, (const float *)(s->p + l.woff), l.co,
                       l.ca + l.cs, l.wb, l.wd);
  }
  SCHK(s, hipGetLastError());
  return LBM_OK;
}

static unsigned conv_blocks(long long npix, int co) { return (unsigned)(((npix + 31) / 32 * (co / 32) + 3) / 4); }

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
      hipLaunchKernelGGL(k_conv<2>, dim3(conv_blocks(np, l.co)), dim3(256), 0, s->st, a);
    } else {
      a.y = s->t[l.out].act;
      a.relu = l.relu;
      hipLaunchKernelGGL(k_conv<0>, dim3(conv_blocks(np, l.co)), dim3(256), 0, s->st, a);
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
  for (Tensor &T : s->t) {
    const long long n = npix_at(s, T.level, nb) * T.c;
    hipLaunchKernelGGL(k_fill, dim3(nblk(n)), dim3(256), 0, s->st, T.grad, n, 0.f);
  }
  size_t pi = s->pool_pairs.size();
  for (int li = (int)s->L.size() - 1; li >= 0; --li) {
    const Layer &l = s->L[li];
    const Tensor &S = s->t[l.s];
    const int h = s->H >> S.level, w = s->W >> S.level;
    const long long np = (long long)nb * h * w;
    const int ci = l.ca + l.cs;
    // gradient of this layer's output, bf16 with the relu mask applied
    const u16 *gy;
    if (l.last) {
      gy = s->gout;
    } else {
      Tensor &O = s->t[l.out];
      // reuse the activation-gradient scratch: gout holds the bf16 copy
      hipLaunchKernelGGL(k_mask_cast, dim3(nblk(np * l.co)), dim3(256), 0, s->st, (const float *)O.grad,
                         (const u16 *)O.act, s->gout, np * l.co, O.relu ? 1 : 0);
      gy = s->gout;
    }
    // weight and bias gradients
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
    wa.nchunks = (int)((np + WGCHUNK - 1) / WGCHUNK);
    wa.part = s->part;
    const long long nw = (long long)l.co * 9 * ci;
    const long long waves = (long long)(l.co / 32) * ((ci + 31) / 32) * 9 * wa.nchunks;
    hipLaunchKernelGGL(k_wgrad, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, s->st, wa);
    hipLaunchKernelGGL(k_reduce_chunks, dim3(nblk(nw)), dim3(256), 0, s->st, (const float *)s->part, nw, wa.nchunks,
                       s->gp + l.woff);
    const int bchunks = (int)((np + WCHUNK - 1) / WCHUNK);
    hipLaunchKernelGGL(k_bias_part, dim3(bchunks), dim3(256), 0, s->st, gy, np, l.co, s->part);
    hipLaunchKernelGGL(k_reduce_chunks, dim3(nblk(l.co)), dim3(256), 0, s->st, (const float *)s->part, (long long)l.co,
                       bchunks, s->gp + l.boff);
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
      if (l.ca > 0)
        hipLaunchKernelGGL(k_fill, dim3(nblk(np * l.ca)), dim3(256), 0, s->st, s->gup, np * l.ca, 0.f);
      hipLaunchKernelGGL(k_conv<1>, dim3(conv_blocks(np, ci)), dim3(256), 0, s->st, a);
      if (l.ca > 0) {
        Tensor &A = s->t[l.a];
        hipLaunchKernelGGL(k_up_bwd, dim3(nblk(np / 4 * l.ca)), dim3(256), 0, s->st, (const float *)s->gup, A.grad,
                           nb, h / 2, w / 2, l.ca);
      }
    }
    // pooling steps that produced this layer's input run backward next
    while (pi > 0 && s->pool_pairs[pi - 1].second == l.s) {
      --pi;
      Tensor &src = s->t[s->pool_pairs[pi].first];
      const Tensor &dst = s->t[s->pool_pairs[pi].second];
      const int hh = s->H >> dst.level, ww = s->W >> dst.level;
      hipLaunchKernelGGL(k_pool_bwd, dim3(nblk((long long)nb * 4 * hh * ww * dst.c)), dim3(256), 0, s->st,
                         (const float *)dst.grad, src.grad, nb, hh, ww, dst.c);
    }
  }
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
  if (hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking) != hipSuccess)
    return bail(sfail(nullptr, LBM_EHIP, "lbm_sur_create: stream"));
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
  for (Tensor &T : s->t) {
    if (T.act) (void)hipFree(T.act);
    if (T.grad) (void)hipFree(T.grad);
  }
  for (Layer &l : s->L) {
    if (l.wb) (void)hipFree(l.wb);
    if (l.wd) (void)hipFree(l.wd);
  }
  void *ptrs[] = {s->p, s->gp, s->m, s->v, s->out, s->tgt, s->mask, s->xin, s->gout, s->part, s->gup,
                  s->lpart, s->lsum, s->dx, s->dt, s->dm};
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  if (s->st) (void)hipStreamDestroy(s->st);
  delete s;
  return LBM_OK;
}

}  // extern "C"
