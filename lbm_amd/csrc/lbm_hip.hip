// liblbm_hip: fused lattice-Boltzmann step kernels for MI355X (gfx950) and
// the C ABI declared in include/lbm_hip.h.
//
// Scheme.  The reference runs one time step as nine CPU passes
// (src/lbm/solver.cpp:307-320): macroscopic values from fold, feq, BGK
// collision into f, push propagation fold(nbr(c,i),i) = f(c,i) through an
// int64 neighbour table, then boundary conditions that overwrite incoming
// populations of boundary cells.  Here the same arithmetic is one kernel per
// step in pull form.  The stored state S is the post-collision population
// set f^n (structure of arrays, one plane per direction, x fastest).  A
// step at cell x gathers g_j = S(x - c_j, j), which is exactly the
// reference's post-propagation fold(x, j); where the upstream cell is
// missing (domain face or solid obstacle) the boundary rule of the face
// replaces it in registers (bounce-back reads S(x, opp j), the value the
// reference's bounce-back copies, src/lbm/bnd/bnd_dirichlet.h:85-116).
// Then moments, equilibrium and collision run in registers and S' is
// written once: 2*Q loads+stores per cell, no neighbour table on device.
//
// Populations are stored shifted, h = f - w (density deviation form), which
// keeps FP32 and 16-bit storage well conditioned; moments, equilibrium and
// collision run in FP64 for FP32 and FP64 storage and in FP32 for 16-bit
// storage.  Checked steps reduce mass, momentum and the
// one-step change of (u, rho) in FP64 with a fixed-order two-stage tree.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/lbm_hip.h"
#include "lbm_cell.h"
#include "lbm_launch.h"

namespace lbm {

// Fixed-order second stage: block k reduces sum k; thread t sums blocks
// t, t+1024, ... in order, then a fixed pairwise tree.
__global__ void __launch_bounds__(1024) k_finalize(const double *partials, long long nblocks, double *out) {
  __shared__ double s[1024];
  const int k = blockIdx.x;
  const double *p = partials + (long long)k * nblocks;
  double acc = 0;
  for (long long b = threadIdx.x; b < nblocks; b += 1024) acc += p[b];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int st = 512; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) s[threadIdx.x] += s[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[k] = s[0];
}

// Solid-link masks (see LINK_SELF_SOLID) of a padded box; the periodic
// axes wrap, links across a non-periodic face stay clear (faces are handled
// from coordinates).  One thread per padded cell.
template <int Q>
__global__ void __launch_bounds__(256) k_links(const uint8_t *flags, uint32_t *links, int nx, int ny, int nz, int px,
                                               int p0, int p1, int p2) {
  using L = Lat<Q>;
  const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y, z = blockIdx.z;
  if (x >= px) return;
  const long long cell = x + (long long)px * (y + (long long)ny * z);
  uint32_t m = 0;
  if (x < nx) {
    if (flags[cell] != LBM_CELL_FLUID) m = LINK_SELF_SOLID;
#pragma unroll
    for (int j = 0; j < Q - 1; ++j) {
      int sx = x - L::cx[j], sy = y - L::cy[j], sz = z - L::cz[j];
      if (sx < 0 || sx >= nx) { if (!p0) continue; sx = sx < 0 ? sx + nx : sx - nx; }
      if (sy < 0 || sy >= ny) { if (!p1) continue; sy = sy < 0 ? sy + ny : sy - ny; }
      if (sz < 0 || sz >= nz) { if (!p2) continue; sz = sz < 0 ? sz + nz : sz - nz; }
      if (flags[sx + (long long)px * (sy + (long long)ny * sz)] != LBM_CELL_FLUID) m |= 1u << j;
    }
  }
  links[cell] = m;
}

// Initial condition: h = feq(rho, u) - w, then the first collision, i.e. the
// state after the collision of the reference's first time step.
template <int Q, typename ST>
__global__ void __launch_bounds__(256) k_init(KArgs a, const double *rho, const double *u) {
  using L = Lat<Q>;
  using C = typename Store<ST>::C;
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  const int z = blockIdx.z;
  if (x >= a.nx || y >= a.ny) return;
  const long long cell = (long long)x + (long long)a.px * ((long long)y + (long long)a.ny * z);
  const long long cc = (long long)x + (long long)a.nx * ((long long)y + (long long)a.ny * z);
  ST *dst = static_cast<ST *>(a.dst);
  if (a.flags && a.flags[cell] != LBM_CELL_FLUID) {
    for (int j = 0; j < Q; ++j) Store<ST>::st(dst + j * a.plane + cell, C(0));
    return;
  }
  const C r = C(rho[cc]);
  const C u0 = C(u[cc * L::D + 0]), u1 = C(u[cc * L::D + 1]), u2 = L::D == 3 ? C(u[cc * L::D + 2]) : C(0);
  const C usq = u0 * u0 + u1 * u1 + u2 * u2;
  typename Store<ST>::R h[Q];
#pragma unroll
  for (int j = 0; j < Q; ++j) {
    const C cu = C(L::cx[j]) * u0 + C(L::cy[j]) * u1 + C(L::cz[j]) * u2;
    h[j] = typename Store<ST>::R(C(L::w[j]) * ((r - C(1)) + r * (C(3) * cu + C(4.5) * cu * cu - C(1.5) * usq)));
  }
  const C Fx = C(a.F[0]), Fy = C(a.F[1]), Fz = C(a.F[2]);
  const bool force = a.F[0] != 0.0 || a.F[1] != 0.0 || a.F[2] != 0.0;
  const Moments<Q, C> m = moments<Q, C>(h, Fx, Fy, Fz);
  if (a.snap_out) {
    using V4 = typename Vec4<C>::T;
    V4 v;
    v.x = m.ux; v.y = m.uy; v.z = m.uz; v.w = m.rho;
    static_cast<V4 *>(a.snap_out)[cell] = v;
  }
  collide_store<Q, ST, false>(h, m, C(a.omega), Fx, Fy, Fz, force, dst, a.plane, cell);
}

// Macroscopic fields of the current state: rho and u* = (m_S - F/2)/rho.
// Also reduces mass and momentum when partials is set.
template <int Q, typename ST>
__global__ void __launch_bounds__(256) k_macros(KArgs a, double *rho, double *u) {
  using L = Lat<Q>;
  using C = typename Store<ST>::C;
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  const int z = blockIdx.z;
  const long long cell = (long long)x + (long long)a.px * ((long long)y + (long long)a.ny * z);
  const long long cc = (long long)x + (long long)a.nx * ((long long)y + (long long)a.ny * z);
  bool active = x < a.nx && y < a.ny;
  const bool fluid = active && !(a.flags && a.flags[cell] != LBM_CELL_FLUID);
  double red[NSUM];
  for (int k = 0; k < NSUM; ++k) red[k] = 0;
  if (active) {
    const ST *src = static_cast<const ST *>(a.src);
    typename Store<ST>::R h[Q];
#pragma unroll
    for (int j = 0; j < Q; ++j) h[j] = Store<ST>::ld(src + j * a.plane + cell);
    const Moments<Q, C> m = moments<Q, C>(h, C(-a.F[0]), C(-a.F[1]), C(-a.F[2]));
    double r = m.rho, ux = m.ux, uy = m.uy, uz = m.uz;
    if (!fluid) { r = 0; ux = uy = uz = 0; }
    if (rho) {
      rho[cc] = r;
      u[cc * L::D + 0] = ux;
      u[cc * L::D + 1] = uy;
      if (L::D == 3) u[cc * L::D + 2] = uz;
    }
    if (fluid) {
      red[0] = r; red[1] = r * ux; red[2] = r * uy; red[3] = r * uz;
      red[9] = ux * ux + uy * uy + uz * uz;
      red[10] = (isfinite(r) && isfinite(ux) && isfinite(uy) && isfinite(uz)) ? 0.0 : 1.0;
      red[11] = 1.0;
    }
  }
  if (a.partials) {
    const long long nb = (long long)gridDim.x * gridDim.y * gridDim.z;
    const long long b = blockIdx.x + (long long)gridDim.x * (blockIdx.y + (long long)gridDim.y * blockIdx.z);
    block_reduce_store<256>(red, a.partials, b, nb);
  }
}

template <int Q, typename ST>
__global__ void k_get_pops(KArgs a, double *f) {
  using L = Lat<Q>;
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  const int z = blockIdx.z;
  if (x >= a.nx || y >= a.ny) return;
  const long long cell = (long long)x + (long long)a.px * ((long long)y + (long long)a.ny * z);
  const long long cc = (long long)x + (long long)a.nx * ((long long)y + (long long)a.ny * z);
  const long long n = (long long)a.nx * a.ny * a.nz;
  const ST *src = static_cast<const ST *>(a.src);
  for (int j = 0; j < Q; ++j) f[j * n + cc] = (double)Store<ST>::ld(src + j * a.plane + cell) + L::w[j];
}

template <int Q, typename ST>
__global__ void k_set_pops(KArgs a, const double *f) {
  using L = Lat<Q>;
  using C = typename Store<ST>::C;
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  const int z = blockIdx.z;
  if (x >= a.nx || y >= a.ny) return;
  const long long cell = (long long)x + (long long)a.px * ((long long)y + (long long)a.ny * z);
  const long long cc = (long long)x + (long long)a.nx * ((long long)y + (long long)a.ny * z);
  const long long n = (long long)a.nx * a.ny * a.nz;
  ST *dst = static_cast<ST *>(a.dst);
  for (int j = 0; j < Q; ++j) Store<ST>::st(dst + j * a.plane + cell, C(f[j * n + cc] - L::w[j]));
}

// Streaming copy used as the bandwidth ceiling: one 16-byte element per
// thread, nontemporal like the step kernel.
__global__ void __launch_bounds__(256) k_copy(const float4 *__restrict__ in, float4 *__restrict__ out, long long n) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) {
    const float4 *p = in + i;
    float4 v;
    v.x = __builtin_nontemporal_load(&p->x);
    v.y = __builtin_nontemporal_load(&p->y);
    v.z = __builtin_nontemporal_load(&p->z);
    v.w = __builtin_nontemporal_load(&p->w);
    float4 *q = out + i;
    __builtin_nontemporal_store(v.x, &q->x);
    __builtin_nontemporal_store(v.y, &q->y);
    __builtin_nontemporal_store(v.z, &q->z);
    __builtin_nontemporal_store(v.w, &q->w);
  }
}

}  // namespace lbm

// Library-wide scratch of lbm_step_ex_noalloc: the FP64 per-block partial
// sums of a check step (NSUM arrays of up to LBM_NOALLOC_MAX_BLOCKS, block
// order) followed by the NSUM reduced sums.  Every check launch writes all
// partials of its blocks and k_finalize writes all sums, so nothing needs
// clearing between calls.  Each device holds its own copy until the process
// ends; host code reaches it only through the address that
// hipGetSymbolAddress returns for the current device (noalloc_scratch).
constexpr long long NOALLOC_DOUBLES = (long long)LBM_NOALLOC_MAX_BLOCKS * lbm::NSUM + lbm::NSUM;
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

// ============================================================================
// C ABI
using namespace lbm;

struct lbm_handle {
  int Q = 0, dim = 0, storage = 0, esize = 0;
  int nx = 0, ny = 0, nz = 0, px = 0;
  long long ncell = 0, plane = 0;
  void *buf[2] = {nullptr, nullptr};
  int cur = 0;
  uint8_t *flags = nullptr;
  uint32_t *links = nullptr;  // solid-link masks (padded layout)
  bool has_flags = false;
  FaceDev face[6];
  FaceDev *dface = nullptr;  // device copy of face[] read by the kernels
  int per[3] = {1, 1, 1};
  double omega = 1.0, F[3] = {0, 0, 0};
  void *snap[2] = {nullptr, nullptr};
  int snap_last = 0;  // buffer holding the most recent snapshot
  double *partials = nullptr;
  long long partials_cap = 0;
  double *sums_d = nullptr;
  void *io = nullptr;  // FP64 (rho, u) staging, see ensure_io
  size_t io_cap = 0;
  hipStream_t stream = nullptr;
  int bx = 256, by = 1, nt = 1;  // nontemporal loads/stores measured fastest on gfx950
  bool w64 = false;              // 64-bit population offsets (plane * esize >= 2^32)
  bool lds = false;              // LDS-staged x halo (lbm_set_halo)
  bool forcing = false;          // lbm_set_pressure_forcing
  // steps of a domain with faces or solids choose a path per wave: the
  // waves of interior cells only run without the boundary code;
  // LBM_STEP_WAVE_PATHS=0 keeps the general path for every wave
  unsigned long long *lanes = nullptr;
  bool wave_paths = true;
  double p_in = 1.0, p_out = 1.0;
  double *fmac = nullptr;        // (rho, u_x, u_y) of the forced cells before the correction
  long long steps = 0;
  int device = 0;  // device current at lbm_create
  std::string err;
};

static thread_local std::string g_err;

// LBMSolver::forcing (src/lbm/solver.cpp:626-696) on the D2Q9 post-collision
// state in buf: thread 2y + s rewrites cell (x_f, y) of the outlet face
// (s = 0, x_f = nx-1, from v = (1, y), p_out) or the inlet face (s = 1,
// x_f = 0, from v = (nx-2, y), p_in) as f_i = feq_i(p, u_v) + f_i(v) -
// feq_i(rho_v, u_v), feq_i(p, u_v) taking only u_v's x component into c_i.u.
// rho_v and u_v are the moments of v's post-collision populations, which the
// BGK collision (no body force) leaves equal to those of the step.
// Populations are stored shifted by the weights.  The moments of the forced
// cell's own post-collision populations (those of its step) go to
// mac[t][3] first: lbm_get_macros reports them for these cells.
template <typename ST>
__global__ void __launch_bounds__(256) k_forcing(ST *buf, long long plane, int px, int nx, int ny, double p_in,
                                                 double p_out, double *mac) {
  using L = lbm::Lat<9>;
  using C = double;
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= 2 * ny) return;
  const int y = t >> 1, side = t & 1;
  const int vx = side ? nx - 2 : 1, fx = side ? 0 : nx - 1;
  const C p = side ? p_in : p_out;
  const long long vc = vx + (long long)px * y, fc = fx + (long long)px * y;
  C f[9], rho = 0, jx = 0, jy = 0;
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    f[j] = C(L::w[j]) + C(lbm::Store<ST>::ld(buf + j * plane + vc));
    rho += f[j];
    jx += C(L::cx[j]) * f[j];
    jy += C(L::cy[j]) * f[j];
  }
  const C ux = jx / rho, uy = jy / rho, vsq = ux * ux + uy * uy;
  C fr = 0, fjx = 0, fjy = 0;
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    const C g = C(L::w[j]) + C(lbm::Store<ST>::ld(buf + j * plane + fc));
    fr += g;
    fjx += C(L::cx[j]) * g;
    fjy += C(L::cy[j]) * g;
  }
  mac[3 * t] = fr;
  mac[3 * t + 1] = fjx / fr;
  mac[3 * t + 2] = fjy / fr;
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    const C w = C(L::w[j]);
    const C cu = C(L::cx[j]) * ux + C(L::cy[j]) * uy, cux = C(L::cx[j]) * ux;
    const C feqv = w * rho * (C(1) + C(3) * cu + C(4.5) * cu * cu - C(1.5) * vsq);
    const C feqp = w * p * (C(1) + C(3) * cux + C(4.5) * cux * cux - C(1.5) * vsq);
    lbm::Store<ST>::st(buf + j * plane + fc, feqp + f[j] - feqv - w);
  }
}

static int fail(lbm_handle *h, int code, const std::string &msg) {
  if (h) h->err = msg;
  g_err = msg;
  return code;
}

#define HIPCHK(h, call)                                                                      \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      return fail((h), LBM_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));        \
  } while (0)

// A pair of timing events, destroyed on every exit path.
struct EventPair {
  hipEvent_t e[2] = {nullptr, nullptr};
  EventPair() = default;
  EventPair(const EventPair &) = delete;
  EventPair &operator=(const EventPair &) = delete;
  hipError_t create() {
    for (hipEvent_t &x : e) {
      const hipError_t r = hipEventCreate(&x);
      if (r != hipSuccess) return r;
    }
    return hipSuccess;
  }
  ~EventPair() {
    for (hipEvent_t x : e)
      if (x) (void)hipEventDestroy(x);
  }
};

// A transient device buffer, freed on every exit path unless released.
struct DevBuf {
  void *p = nullptr;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  hipError_t alloc(size_t bytes) { return hipMalloc(&p, bytes); }
  void *release() {
    void *q = p;
    p = nullptr;
    return q;
  }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

// Staging buffer for the FP64 (rho, u) fields of lbm_init_fields and
// lbm_get_macros, kept by the handle so that repeated read-backs (output
// and postprocessing hooks) do not allocate.
static int ensure_io(lbm_handle *h, size_t bytes) {
  if (bytes <= h->io_cap) return LBM_OK;
  if (h->io) HIPCHK(h, hipFree(h->io));
  h->io = nullptr;
  h->io_cap = 0;
  HIPCHK(h, hipMalloc(&h->io, bytes));
  h->io_cap = bytes;
  return LBM_OK;
}

// Device address of `scratch` on device dev, resolved once per device with
// that device current (lbm_create resolves it for its device).
constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

static int noalloc_scratch(lbm_handle *h, int dev, double **out) {
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(h, LBM_EINVAL, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    int cur = 0;
    HIPCHK(h, hipGetDevice(&cur));
    if (cur != dev) HIPCHK(h, hipSetDevice(dev));
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (cur != dev) HIPCHK(h, hipSetDevice(cur));
    if (e != hipSuccess) return fail(h, LBM_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double *>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return LBM_OK;
}

static int upload_faces(lbm_handle *h) {
  HIPCHK(h, hipMemcpyAsync(h->dface, h->face, sizeof(h->face), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return LBM_OK;
}

static KArgs make_args(lbm_handle *h) {
  KArgs a;
  memset(&a, 0, sizeof(a));
  a.src = h->buf[h->cur];
  a.dst = h->buf[h->cur ^ 1];
  a.flags = h->has_flags ? h->flags : nullptr;
  a.links = h->has_flags ? h->links : nullptr;
  a.plane = h->plane;
  a.nx = h->nx; a.ny = h->ny; a.nz = h->nz; a.px = h->px;
  for (int i = 0; i < 3; ++i) a.per[i] = h->per[i];
  a.face = h->dface;
  a.omega = h->omega;
  for (int i = 0; i < 3; ++i) a.F[i] = h->F[i];
  a.lanes = h->wave_paths ? h->lanes : nullptr;
  return a;
}

static void launch_dims(const lbm_handle *h, dim3 &grid, dim3 &block) {
  block = dim3(h->bx, h->by, 1);
  grid = dim3((h->nx + h->bx - 1) / h->bx, (h->ny + h->by - 1) / h->by, h->nz);
}

static long long nblocks_of(const lbm_handle *h) {
  dim3 g, b;
  launch_dims(h, g, b);
  return (long long)g.x * g.y * g.z;
}

// dispatch helpers over (Q, storage)
template <template <int, typename> class F, typename... Args>
static void dispatch_qs(int Q, int storage, Args... args) {
  if (Q == 9) {
    switch (storage) {
      case LBM_STORE_F32: F<9, float>::run(args...); break;
      case LBM_STORE_F64: F<9, double>::run(args...); break;
      case LBM_STORE_F16: F<9, __half>::run(args...); break;
      case LBM_STORE_BF16: F<9, __hip_bfloat16>::run(args...); break;
    }
  } else {
    switch (storage) {
      case LBM_STORE_F32: F<19, float>::run(args...); break;
      case LBM_STORE_F64: F<19, double>::run(args...); break;
      case LBM_STORE_F16: F<19, __half>::run(args...); break;
      case LBM_STORE_BF16: F<19, __hip_bfloat16>::run(args...); break;
    }
  }
}

template <int Q, typename ST>
struct InitL {
  static void run(const KArgs &a, dim3 g, dim3 b, hipStream_t s, const double *r, const double *u) {
    hipLaunchKernelGGL((k_init<Q, ST>), g, b, 0, s, a, r, u);
  }
};
template <int Q, typename ST>
struct MacroL {
  static void run(const KArgs &a, dim3 g, dim3 b, hipStream_t s, double *r, double *u) {
    hipLaunchKernelGGL((k_macros<Q, ST>), g, b, 0, s, a, r, u);
  }
};
template <int Q, typename ST>
struct GetL {
  static void run(const KArgs &a, dim3 g, dim3 b, hipStream_t s, double *f) {
    hipLaunchKernelGGL((k_get_pops<Q, ST>), g, b, 0, s, a, f);
  }
};
template <int Q, typename ST>
struct SetL {
  static void run(const KArgs &a, dim3 g, dim3 b, hipStream_t s, const double *f) {
    hipLaunchKernelGGL((k_set_pops<Q, ST>), g, b, 0, s, a, f);
  }
};

static int esize_of(int storage) {
  switch (storage) {
    case LBM_STORE_F32: return 4;
    case LBM_STORE_F64: return 8;
    case LBM_STORE_F16: return 2;
    case LBM_STORE_BF16: return 2;
  }
  return 0;
}

static int ensure_snapshots(lbm_handle *h);

static int ensure_partials(lbm_handle *h) {
  const long long nb = nblocks_of(h);
  if (nb > h->partials_cap) {
    if (h->partials) HIPCHK(h, hipFree(h->partials));
    h->partials = nullptr;
    HIPCHK(h, hipMalloc(&h->partials, (size_t)nb * NSUM * sizeof(double)));
    h->partials_cap = nb;
  }
  if (!h->sums_d) HIPCHK(h, hipMalloc(&h->sums_d, NSUM * sizeof(double)));
  return LBM_OK;
}

template <int Q>
static void launch_links(const uint8_t *flags, uint32_t *links, int nx, int ny, int nz, int px, const int *per,
                         hipStream_t st) {
  hipLaunchKernelGGL((k_links<Q>), dim3((px + 255) / 256, ny, nz), dim3(256), 0, st, flags, links, nx, ny, nz, px,
                     per[0], per[1], per[2]);
}

// lanes[w], one word per wave of the step grid (the step kernels' block
// shape): bit i set when lane i's cell is interior -- inside the domain, one
// cell or more from every face (no wrap) and without a link to a solid or
// solid itself (links == 0).  The step updates a wave whose every bit is
// set without the boundary code.
__global__ void __launch_bounds__(256) k_lanes(const uint32_t *links, int nx, int ny, int nz, int px, int dim,
                                               unsigned long long *lanes) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  const int z = blockIdx.z;
  bool in = x >= 1 && x <= nx - 2 && y >= 1 && y <= ny - 2 && (dim == 2 || (z >= 1 && z <= nz - 2));
  if (in && links) in = links[x + (long long)px * (y + (long long)ny * z)] == 0u;
  const unsigned long long m = __ballot(in);
  const int t = (int)(threadIdx.y * blockDim.x + threadIdx.x);
  const long long wave =
      ((long long)(blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x) * (blockDim.x * blockDim.y / 64) +
      (t >> 6);
  if ((t & 63) == 0) lanes[wave] = m;
}

// (Re)build the solid-link masks and the interior lanes; they depend on the
// periodic axes, so lbm_set_bc calls this again.
static int build_links(lbm_handle *h) {
  if (h->has_flags) {
    if (h->Q == 9) launch_links<9>(h->flags, h->links, h->nx, h->ny, h->nz, h->px, h->per, h->stream);
    else launch_links<19>(h->flags, h->links, h->nx, h->ny, h->nz, h->px, h->per, h->stream);
    HIPCHK(h, hipGetLastError());
  }
  if (h->lanes) {
    dim3 g, b;
    launch_dims(h, g, b);
    hipLaunchKernelGGL(k_lanes, g, b, 0, h->stream, (const uint32_t *)(h->has_flags ? h->links : nullptr), h->nx, h->ny,
                       h->nz, h->px, h->dim, h->lanes);
    HIPCHK(h, hipGetLastError());
  }
  return LBM_OK;
}

extern "C" {

int lbm_create(const lbm_grid_desc *grid, const lbm_params *p, lbm_handle **out) {
  if (!grid || !p || !out) return fail(nullptr, LBM_EINVAL, "lbm_create: null argument");
  *out = nullptr;
  if (p->lattice != LBM_D2Q9 && p->lattice != LBM_D3Q19)
    return fail(nullptr, LBM_EINVAL, "lbm_create: lattice must be 9 (D2Q9) or 19 (D3Q19)");
  const int dim = p->lattice == LBM_D2Q9 ? 2 : 3;
  if (grid->dim != dim) return fail(nullptr, LBM_EINVAL, "lbm_create: grid dim does not match the lattice");
  if (grid->nx < 3 || grid->ny < 3 || grid->nz < 1 || (dim == 2 && grid->nz != 1) || (dim == 3 && grid->nz < 3))
    return fail(nullptr, LBM_EINVAL, "lbm_create: each extent must be >= 3 (nz = 1 in 2D)");
  const int es = esize_of(p->storage);
  if (!es) return fail(nullptr, LBM_EINVAL, "lbm_create: unknown storage");
  if (!(p->omega > 0.0 && p->omega < 2.0)) return fail(nullptr, LBM_EINVAL, "lbm_create: omega must be in (0,2)");
  lbm_handle *h = new lbm_handle();
  h->Q = p->lattice; h->dim = dim; h->storage = p->storage; h->esize = es;
  h->nx = grid->nx; h->ny = grid->ny; h->nz = grid->nz;
  h->px = (grid->nx + 63) / 64 * 64;
  h->ncell = (long long)grid->nx * grid->ny * grid->nz;
  long long pad = 0;
  if (const char *e = getenv("LBM_PLANE_PAD")) pad = atoll(e);
  h->plane = (long long)h->px * h->ny * h->nz + pad;
  h->w64 = (unsigned long long)h->plane * (unsigned long long)es > 0xffffffffull;
  h->omega = p->omega;
  for (int i = 0; i < 3; ++i) h->F[i] = p->force[i];
  for (int f = 0; f < 6; ++f) { memset(&h->face[f], 0, sizeof(FaceDev)); h->face[f].kind = LBM_FACE_PERIODIC; h->face[f].rho = 1.0; }
  h->bx = h->px >= 256 ? 256 : (h->px >= 128 ? 128 : 64);
  h->by = 256 / h->bx;
  int rc = LBM_OK;
  auto bail = [&](int code) { lbm_destroy(h); return code; };
  double *scratch_ptr = nullptr;
  if (hipGetDevice(&h->device) != hipSuccess || noalloc_scratch(nullptr, h->device, &scratch_ptr) != LBM_OK) {
    fail(nullptr, LBM_EHIP, "lbm_create: device set-up failed: " + g_err);
    return bail(LBM_EHIP);
  }
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    fail(nullptr, LBM_EHIP, "lbm_create: hipStreamCreate failed");
    return bail(LBM_EHIP);
  }
  if (hipMalloc(&h->dface, sizeof(h->face)) != hipSuccess || upload_faces(h) != LBM_OK) {
    fail(nullptr, LBM_EHIP, "lbm_create: face table upload failed");
    return bail(LBM_EHIP);
  }
  const size_t bytes = (size_t)h->plane * h->Q * es;
  for (int b = 0; b < 2; ++b) {
    if (hipMalloc(&h->buf[b], bytes) != hipSuccess) {
      fail(nullptr, LBM_ENOMEM, "lbm_create: hipMalloc of populations failed");
      return bail(LBM_ENOMEM);
    }
    if (hipMemsetAsync(h->buf[b], 0, bytes, h->stream) != hipSuccess) {
      fail(nullptr, LBM_EHIP, "lbm_create: hipMemset failed");
      return bail(LBM_EHIP);
    }
  }
  if (grid->solid) {
    std::vector<uint8_t> padded((size_t)h->px * h->ny * h->nz, LBM_CELL_SOLID);
    for (long long z = 0; z < h->nz; ++z)
      for (long long y = 0; y < h->ny; ++y)
        for (long long x = 0; x < h->nx; ++x) {
          const uint8_t v = grid->solid[x + h->nx * (y + h->ny * z)];
          if (v > LBM_CELL_SOLID) {
            fail(nullptr, LBM_EINVAL, "lbm_create: cell flag must be 0 (fluid) or 1 (solid)");
            return bail(LBM_EINVAL);
          }
          padded[x + h->px * (y + h->ny * z)] = v;
        }
    if (hipMalloc(&h->flags, padded.size()) != hipSuccess) {
      fail(nullptr, LBM_ENOMEM, "lbm_create: hipMalloc of flags failed");
      return bail(LBM_ENOMEM);
    }
    if (hipMemcpy(h->flags, padded.data(), padded.size(), hipMemcpyHostToDevice) != hipSuccess) {
      fail(nullptr, LBM_EHIP, "lbm_create: flag upload failed");
      return bail(LBM_EHIP);
    }
    const long long ncp = (long long)h->px * h->ny * h->nz;
    if (hipMalloc(&h->links, ncp * sizeof(uint32_t)) != hipSuccess) {
      fail(nullptr, LBM_ENOMEM, "lbm_create: hipMalloc of link masks failed");
      return bail(LBM_ENOMEM);
    }
    h->has_flags = true;
  }
  if (const char *e = getenv("LBM_STEP_WAVE_PATHS")) h->wave_paths = atoi(e) != 0;
  if (hipMalloc(&h->lanes, nblocks_of(h) * 4 * sizeof(unsigned long long)) != hipSuccess) {
    fail(nullptr, LBM_ENOMEM, "lbm_create: hipMalloc of the interior-lane masks failed");
    return bail(LBM_ENOMEM);
  }
  rc = build_links(h);
  if (rc) return bail(rc);
  rc = ensure_partials(h);
  if (rc) return bail(rc);
  *out = h;
  return LBM_OK;
}

// A NEEM face reads the post-streaming populations of the inward
// neighbour with bounce-back only, so the faces of the other axes must be
// walls or periodic.
static bool neem_neighbours_ok(const FaceDev *face, int dim) {
  for (int f = 0; f < 2 * dim; ++f) {
    if (face[f].kind != LBM_FACE_PRESSURE_NEEM && face[f].kind != LBM_FACE_VELOCITY_NEEM) continue;
    for (int g = 0; g < 2 * dim; ++g)
      if ((g >> 1) != (f >> 1) && face[g].kind != LBM_FACE_PERIODIC && face[g].kind != LBM_FACE_WALL) return false;
  }
  return true;
}

int lbm_set_bc(lbm_handle *h, const lbm_bc_desc *bcs, int32_t n) {
  if (!h || (n > 0 && !bcs) || n < 0) return fail(h, LBM_EINVAL, "lbm_set_bc: bad argument");
  FaceDev face[6];
  int per[3] = {1, 1, 1};
  for (int f = 0; f < 6; ++f) { memset(&face[f], 0, sizeof(FaceDev)); face[f].kind = LBM_FACE_PERIODIC; face[f].rho = 1.0; }
  bool seen[6] = {false, false, false, false, false, false};
  for (int i = 0; i < n; ++i) {
    const lbm_bc_desc &d = bcs[i];
    if (d.face < 0 || d.face >= 2 * h->dim) return fail(h, LBM_EINVAL, "lbm_set_bc: face index out of range");
    if (seen[d.face]) return fail(h, LBM_EINVAL, "lbm_set_bc: face described twice");
    if (d.kind < LBM_FACE_PERIODIC || d.kind > LBM_FACE_WALL_NEBB0) return fail(h, LBM_EINVAL, "lbm_set_bc: unknown face kind");
    if ((d.kind == LBM_FACE_WALL_NEBB || d.kind == LBM_FACE_WALL_NEBB0) && h->Q != 9)
      return fail(h, LBM_EINVAL, "lbm_set_bc: NEBB walls are D2Q9 only");
    if (d.kind == LBM_FACE_WALL_NEBB && d.value[d.face >> 1] != 0.0)
      return fail(h, LBM_EINVAL, "lbm_set_bc: a NEBB wall velocity must be tangential");
    seen[d.face] = true;
    face[d.face].kind = d.kind;
    face[d.face].prio = d.priority;
    for (int k = 0; k < 3; ++k) face[d.face].u[k] = d.value[k];
    face[d.face].rho = d.rho;
  }
  for (int ax = 0; ax < h->dim; ++ax) {
    const bool p0 = face[2 * ax].kind == LBM_FACE_PERIODIC, p1 = face[2 * ax + 1].kind == LBM_FACE_PERIODIC;
    if (p0 != p1) return fail(h, LBM_EINVAL, "lbm_set_bc: a periodic face needs a periodic partner");
    per[ax] = p0 ? 1 : 0;
    if ((face[2 * ax].kind == LBM_FACE_PERIODIC_PRESSURE) != (face[2 * ax + 1].kind == LBM_FACE_PERIODIC_PRESSURE))
      return fail(h, LBM_EINVAL, "lbm_set_bc: a periodic pressure face needs a periodic pressure partner");
    const int na = ax == 0 ? h->nx : (ax == 1 ? h->ny : h->nz);
    for (int s = 0; s < 2; ++s) {
      const int k = face[2 * ax + s].kind;
      if ((k == LBM_FACE_PRESSURE_ABB || k == LBM_FACE_PRESSURE_NEEM || k == LBM_FACE_VELOCITY_NEEM ||
           k == LBM_FACE_WALL_NEEM) && na < 3)
        return fail(h, LBM_EINVAL, "lbm_set_bc: anti-bounce-back and NEEM faces need at least 3 cells along their normal");
      if ((k == LBM_FACE_WALL_EQ || k == LBM_FACE_WALL_NEBB || k == LBM_FACE_WALL_NEBB0) && na < 2)
        return fail(h, LBM_EINVAL, "lbm_set_bc: wet-node walls need at least 2 cells along their normal");
    }
  }
  if (!neem_neighbours_ok(face, h->dim))
    return fail(h, LBM_EINVAL, "lbm_set_bc: the faces next to a NEEM face must be walls or periodic");
  // a NEEM wall reads its inward neighbour's post-streaming populations,
  // which are defined across periodic, periodic-pressure, bounce-back and
  // anti-bounce-back faces (the latter bounced back plainly)
  for (int f = 0; f < 2 * h->dim; ++f) {
    if (face[f].kind != LBM_FACE_WALL_NEEM) continue;
    for (int g = 0; g < 2 * h->dim; ++g) {
      const int k = face[g].kind;
      if ((g >> 1) != (f >> 1) && k != LBM_FACE_PERIODIC && k != LBM_FACE_PERIODIC_PRESSURE && k != LBM_FACE_WALL &&
          k != LBM_FACE_PRESSURE_ABB)
        return fail(h, LBM_EINVAL, "lbm_set_bc: the faces next to a NEEM wall must be periodic, bounce-back or anti-bounce-back");
    }
  }
  memcpy(h->face, face, sizeof(face));
  const int urc = upload_faces(h);
  if (urc) return urc;
  memcpy(h->per, per, sizeof(per));
  return build_links(h);
}

static int apply_forcing(lbm_handle *h);

int lbm_init_fields(lbm_handle *h, const double *rho, const double *u) {
  if (!h || !rho || !u) return fail(h, LBM_EINVAL, "lbm_init_fields: null argument");
  int rc = ensure_io(h, (size_t)h->ncell * (1 + h->dim) * sizeof(double));
  if (rc) return rc;
  double *rd = static_cast<double *>(h->io), *ud = rd + h->ncell;
  HIPCHK(h, hipMemcpyAsync(rd, rho, h->ncell * sizeof(double), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(ud, u, h->ncell * h->dim * sizeof(double), hipMemcpyHostToDevice, h->stream));
  rc = ensure_snapshots(h);
  if (rc) return rc;
  KArgs a = make_args(h);
  a.dst = h->buf[h->cur];
  h->snap_last = 0;
  a.snap_out = h->snap[0];
  dim3 g, b;
  launch_dims(h, g, b);
  dispatch_qs<InitL>(h->Q, h->storage, a, g, b, h->stream, (const double *)rd, (const double *)ud);
  HIPCHK(h, hipGetLastError());
  if (h->forcing) {  // the initial condition ends with the first collision
    rc = apply_forcing(h);
    if (rc) return rc;
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->steps = 0;
  return LBM_OK;
}

static void fill_sums(const double *s, int dim, lbm_sums *out);

// bytes of one (u, rho) snapshot entry: four values of the compute type
static size_t snap_bytes(int storage) {
  switch (storage) {
    case LBM_STORE_F32: return sizeof(typename Vec4<typename Store<float>::C>::T);
    case LBM_STORE_F64: return sizeof(double4);
    default: return sizeof(float4);
  }
}

static int ensure_snapshots(lbm_handle *h) {
  const size_t cs = snap_bytes(h->storage);
  for (int b = 0; b < 2; ++b)
    if (!h->snap[b]) HIPCHK(h, hipMalloc(&h->snap[b], (size_t)h->px * h->ny * h->nz * cs));
  return LBM_OK;
}

// partials: where a check step puts its per-block sums (the handle's buffer
// unless given)
// (rho, u) of the forced cells as their step left them, into the macro
// fields k_macros wrote (compact x-fastest order)
__global__ void __launch_bounds__(256) k_forcing_macros(const double *mac, int nx, int ny, double *rho, double *u) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= 2 * ny) return;
  const int y = t >> 1, fx = (t & 1) ? 0 : nx - 1;
  const long long cc = fx + (long long)nx * y;
  rho[cc] = mac[3 * t];
  u[2 * cc] = mac[3 * t + 1];
  u[2 * cc + 1] = mac[3 * t + 2];
}

// the forcing correction on the state the last launch wrote
static int apply_forcing(lbm_handle *h) {
  const dim3 g((unsigned)((2 * h->ny + 255) / 256)), b(256);
  void *buf = h->buf[h->cur];
  switch (h->storage) {
    case LBM_STORE_F32:
      hipLaunchKernelGGL(k_forcing<float>, g, b, 0, h->stream, (float *)buf, h->plane, h->px, h->nx, h->ny, h->p_in, h->p_out, h->fmac);
      break;
    case LBM_STORE_F64:
      hipLaunchKernelGGL(k_forcing<double>, g, b, 0, h->stream, (double *)buf, h->plane, h->px, h->nx, h->ny, h->p_in,
                         h->p_out, h->fmac);
      break;
    case LBM_STORE_F16:
      hipLaunchKernelGGL(k_forcing<__half>, g, b, 0, h->stream, (__half *)buf, h->plane, h->px, h->nx, h->ny, h->p_in,
                         h->p_out, h->fmac);
      break;
    default:
      hipLaunchKernelGGL(k_forcing<__hip_bfloat16>, g, b, 0, h->stream, (__hip_bfloat16 *)buf, h->plane, h->px, h->nx,
                         h->ny, h->p_in, h->p_out, h->fmac);
  }
  HIPCHK(h, hipGetLastError());
  return LBM_OK;
}

static int launch_step(lbm_handle *h, int mode, double *partials = nullptr) {
  KArgs a = make_args(h);
  if (mode != 0) {
    const int rc = ensure_snapshots(h);
    if (rc) return rc;
    a.snap_in = h->snap[h->snap_last];
    a.snap_out = h->snap[h->snap_last ^ 1];
    a.partials = partials ? partials : h->partials;
  }
  dim3 g, b;
  launch_dims(h, g, b);
  // a periodic box without solids runs a kernel without boundary paths
  const bool periodic_box = !h->has_flags && h->per[0] && h->per[1] && (h->dim == 2 || h->per[2]);
  dispatch_qs<StepL>(h->Q, h->storage, a, g, b, h->stream, h->has_flags, mode, h->nt != 0, h->w64, periodic_box,
                     h->lds && h->bx == 256 && h->by == 1);
  HIPCHK(h, hipGetLastError());
  if (mode == 1 || mode == 3) h->snap_last ^= 1;
  h->cur ^= 1;
  h->steps++;
  if (h->forcing) return apply_forcing(h);
  return LBM_OK;
}

static int finalize_sums(lbm_handle *h, lbm_sums *out) {
  hipLaunchKernelGGL(k_finalize, dim3(NSUM), dim3(1024), 0, h->stream, (const double *)h->partials, nblocks_of(h), h->sums_d);
  HIPCHK(h, hipGetLastError());
  double s[NSUM];
  HIPCHK(h, hipMemcpyAsync(s, h->sums_d, sizeof(s), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  fill_sums(s, h->dim, out);
  if (s[10] > 0) return fail(h, LBM_DIVERGED, "non-finite values (solution diverged)");
  return LBM_OK;
}

int lbm_step(lbm_handle *h, int32_t nsteps) {
  if (!h || nsteps < 0) return fail(h, LBM_EINVAL, "lbm_step: bad argument");
  for (int i = 0; i < nsteps; ++i) {
    const int rc = launch_step(h, 0);
    if (rc) return rc;
  }
  return LBM_OK;
}

static void fill_sums(const double *s, int dim, lbm_sums *out) {
  memset(out, 0, sizeof(*out));
  out->mass = s[0];
  for (int d = 0; d < 3; ++d) out->momentum[d] = s[1 + d];
  // reference variable order u, v, (w), rho
  for (int d = 0; d < dim; ++d) out->abs_diff[d] = s[4 + d];
  out->abs_diff[dim] = s[7];
  out->diff_sq = s[8];
  out->norm_sq = s[9];
  if (!std::isfinite(s[8]) || !std::isfinite(s[9])) out->res_rel = NAN;
  else out->res_rel = s[9] > 0 ? sqrt(s[8] / s[9]) : (s[8] > 0 ? INFINITY : 0.0);
  out->nonfinite = s[10];
  out->fluid_cells = s[11];
}

int lbm_step_check(lbm_handle *h, lbm_sums *out) {
  if (!h || !out) return fail(h, LBM_EINVAL, "lbm_step_check: null argument");
  int rc = ensure_partials(h);
  if (rc) return rc;
  rc = launch_step(h, 1);
  if (rc) return rc;
  rc = launch_step(h, 2);
  if (rc) return rc;
  return finalize_sums(h, out);
}

int lbm_step_ex(lbm_handle *h, int32_t mode, lbm_sums *out) {
  if (!h || mode < 0 || mode > 3 || (mode >= 2 && !out)) return fail(h, LBM_EINVAL, "lbm_step_ex: bad argument");
  int rc = ensure_partials(h);
  if (rc) return rc;
  rc = launch_step(h, mode);
  if (rc) return rc;
  if (mode >= 2) return finalize_sums(h, out);
  return LBM_OK;
}

int lbm_step_ex_noalloc(lbm_handle *h, int32_t mode, lbm_sums *out) {
  if (!h || mode < 0 || mode > 3 || (mode >= 2 && !out)) return fail(h, LBM_EINVAL, "lbm_step_ex_noalloc: bad argument");
  if (mode == 0) return launch_step(h, 0);
  const long long nb = nblocks_of(h);
  if (mode >= 2 && nb > LBM_NOALLOC_MAX_BLOCKS)
    return fail(h, LBM_EINVAL, "lbm_step_ex_noalloc: the step launch has " + std::to_string(nb) +
                                   " blocks, above LBM_NOALLOC_MAX_BLOCKS (" + std::to_string(LBM_NOALLOC_MAX_BLOCKS) +
                                   "); use lbm_step_ex");
  if (!h->snap[0] || !h->snap[1])
    return fail(h, LBM_EINVAL, "lbm_step_ex_noalloc: no snapshot buffers yet (lbm_init_fields creates them)");
  if (mode == 1) return launch_step(h, 1);
  double *scr = nullptr;
  int rc = noalloc_scratch(h, h->device, &scr);
  if (rc) return rc;
  // the partials of this launch (NSUM arrays of nb blocks), then the sums
  double *partials = scr, *sums = scr + (long long)LBM_NOALLOC_MAX_BLOCKS * NSUM;
  rc = launch_step(h, mode, partials);
  if (rc) return rc;
  hipLaunchKernelGGL(k_finalize, dim3(NSUM), dim3(1024), 0, h->stream, (const double *)partials, nb, sums);
  HIPCHK(h, hipGetLastError());
  double s[NSUM];
  HIPCHK(h, hipMemcpyAsync(s, sums, sizeof(s), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  fill_sums(s, h->dim, out);
  if (s[10] > 0) return fail(h, LBM_DIVERGED, "non-finite values (solution diverged)");
  return LBM_OK;
}

int lbm_get_macros(lbm_handle *h, double *rho, double *u) {
  if (!h || !rho || !u) return fail(h, LBM_EINVAL, "lbm_get_macros: null argument");
  const int rc = ensure_io(h, (size_t)h->ncell * (1 + h->dim) * sizeof(double));
  if (rc) return rc;
  double *rd = static_cast<double *>(h->io), *ud = rd + h->ncell;
  KArgs a = make_args(h);
  dim3 g, b;
  launch_dims(h, g, b);
  dispatch_qs<MacroL>(h->Q, h->storage, a, g, b, h->stream, rd, ud);
  HIPCHK(h, hipGetLastError());
  if (h->forcing) {
    hipLaunchKernelGGL(k_forcing_macros, dim3((unsigned)((2 * h->ny + 255) / 256)), dim3(256), 0, h->stream,
                       (const double *)h->fmac, h->nx, h->ny, rd, ud);
    HIPCHK(h, hipGetLastError());
  }
  HIPCHK(h, hipMemcpyAsync(rho, rd, h->ncell * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipMemcpyAsync(u, ud, h->ncell * h->dim * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return LBM_OK;
}

int lbm_reduce(lbm_handle *h, lbm_sums *out) {
  if (!h || !out) return fail(h, LBM_EINVAL, "lbm_reduce: null argument");
  int rc = ensure_partials(h);
  if (rc) return rc;
  KArgs a = make_args(h);
  a.partials = h->partials;
  dim3 g, b;
  launch_dims(h, g, b);
  dispatch_qs<MacroL>(h->Q, h->storage, a, g, b, h->stream, (double *)nullptr, (double *)nullptr);
  HIPCHK(h, hipGetLastError());
  hipLaunchKernelGGL(k_finalize, dim3(NSUM), dim3(1024), 0, h->stream, (const double *)h->partials, nblocks_of(h), h->sums_d);
  HIPCHK(h, hipGetLastError());
  double s[NSUM];
  HIPCHK(h, hipMemcpyAsync(s, h->sums_d, sizeof(s), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  s[4] = s[5] = s[6] = s[7] = s[8] = 0;
  fill_sums(s, h->dim, out);
  out->res_rel = 0;
  return LBM_OK;
}

int lbm_get_populations(lbm_handle *h, double *f) {
  if (!h || !f) return fail(h, LBM_EINVAL, "lbm_get_populations: null argument");
  const size_t bytes = (size_t)h->ncell * h->Q * sizeof(double);
  DevBuf fd;
  HIPCHK(h, fd.alloc(bytes));
  KArgs a = make_args(h);
  dim3 g, b;
  launch_dims(h, g, b);
  dispatch_qs<GetL>(h->Q, h->storage, a, g, b, h->stream, static_cast<double *>(fd.p));
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipMemcpyAsync(f, fd.p, bytes, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return LBM_OK;
}

int lbm_set_populations(lbm_handle *h, const double *f) {
  if (!h || !f) return fail(h, LBM_EINVAL, "lbm_set_populations: null argument");
  const size_t bytes = (size_t)h->ncell * h->Q * sizeof(double);
  DevBuf fd;
  HIPCHK(h, fd.alloc(bytes));
  HIPCHK(h, hipMemcpyAsync(fd.p, f, bytes, hipMemcpyHostToDevice, h->stream));
  KArgs a = make_args(h);
  a.dst = h->buf[h->cur];
  dim3 g, b;
  launch_dims(h, g, b);
  dispatch_qs<SetL>(h->Q, h->storage, a, g, b, h->stream, static_cast<const double *>(fd.p));
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return LBM_OK;
}

// The new population and snapshot buffers are allocated and filled before
// the old ones are released: on any failure the handle keeps its storage
// form and state unchanged.
int lbm_set_storage(lbm_handle *h, int32_t storage) {
  if (!h) return fail(h, LBM_EINVAL, "lbm_set_storage: null handle");
  const int es = esize_of(storage);
  if (!es) return fail(h, LBM_EINVAL, "lbm_set_storage: unknown storage");
  if (storage == h->storage) return LBM_OK;
  const size_t bytes = (size_t)h->plane * h->Q * es;
  const size_t sbytes = (size_t)h->px * h->ny * h->nz * snap_bytes(storage);
  DevBuf fd, nb[2], ns[2];
  if (fd.alloc((size_t)h->ncell * h->Q * sizeof(double)) != hipSuccess || nb[0].alloc(bytes) != hipSuccess ||
      nb[1].alloc(bytes) != hipSuccess || ns[0].alloc(sbytes) != hipSuccess || ns[1].alloc(sbytes) != hipSuccess) {
    (void)hipGetLastError();
    return fail(h, LBM_ENOMEM, "lbm_set_storage: not enough device memory for the conversion (storage unchanged)");
  }
  for (int k = 0; k < 2; ++k) HIPCHK(h, hipMemsetAsync(nb[k].p, 0, bytes, h->stream));
  // current state -> FP64 f (exact for every storage form) -> new storage
  KArgs a = make_args(h);
  dim3 g, b;
  launch_dims(h, g, b);
  dispatch_qs<GetL>(h->Q, h->storage, a, g, b, h->stream, static_cast<double *>(fd.p));
  HIPCHK(h, hipGetLastError());
  a.dst = nb[h->cur].p;
  dispatch_qs<SetL>(h->Q, storage, a, g, b, h->stream, static_cast<const double *>(fd.p));
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->stream));
  for (int k = 0; k < 2; ++k) {
    (void)hipFree(h->buf[k]);
    h->buf[k] = nb[k].release();
    if (h->snap[k]) (void)hipFree(h->snap[k]);
    h->snap[k] = ns[k].release();
  }
  h->storage = storage;
  h->esize = es;
  h->w64 = (unsigned long long)h->plane * (unsigned long long)es > 0xffffffffull;
  h->snap_last = 0;
  if (storage != LBM_STORE_F32) h->lds = false;
  return LBM_OK;
}

int lbm_time_steps(lbm_handle *h, int32_t nsteps, double *ms) {
  if (!h || !ms || nsteps <= 0) return fail(h, LBM_EINVAL, "lbm_time_steps: bad argument");
  EventPair ev;
  HIPCHK(h, ev.create());
  HIPCHK(h, hipEventRecord(ev.e[0], h->stream));
  const int rc = lbm_step(h, nsteps);
  if (rc) return rc;
  HIPCHK(h, hipEventRecord(ev.e[1], h->stream));
  HIPCHK(h, hipEventSynchronize(ev.e[1]));
  float t = 0;
  HIPCHK(h, hipEventElapsedTime(&t, ev.e[0], ev.e[1]));
  *ms = t;
  return LBM_OK;
}

int lbm_time_step_groups(lbm_handle *h, const int32_t *steps, int32_t ngroups, double *ms) {
  if (!h || !steps || !ms || ngroups < 1 || ngroups > 4096) return fail(h, LBM_EINVAL, "lbm_time_step_groups: bad argument");
  for (int g = 0; g < ngroups; ++g)
    if (steps[g] < 1) return fail(h, LBM_EINVAL, "lbm_time_step_groups: every group needs at least one step");
  std::vector<EventPair> ev((ngroups + 2) / 2);
  auto event = [&](int k) { return ev[k / 2].e[k % 2]; };
  for (EventPair &p : ev) HIPCHK(h, p.create());
  HIPCHK(h, hipEventRecord(event(0), h->stream));
  for (int g = 0; g < ngroups; ++g) {
    const int rc = lbm_step(h, steps[g]);
    if (rc) return rc;
    HIPCHK(h, hipEventRecord(event(g + 1), h->stream));
  }
  HIPCHK(h, hipEventSynchronize(event(ngroups)));
  for (int g = 0; g < ngroups; ++g) {
    float t = 0;
    HIPCHK(h, hipEventElapsedTime(&t, event(g), event(g + 1)));
    ms[g] = t;
  }
  return LBM_OK;
}

int lbm_set_launch(lbm_handle *h, int32_t block_x, int32_t block_y, int32_t nontemporal) {
  if (!h) return fail(h, LBM_EINVAL, "lbm_set_launch: null handle");
  int bx = block_x > 0 ? block_x : h->bx;
  int by = block_y > 0 ? block_y : h->by;
  if (bx % 64 != 0 || bx * by != 256) return fail(h, LBM_EINVAL, "lbm_set_launch: block_x must be a multiple of 64 and block_x*block_y == 256");
  if (bx != h->bx || by != h->by) {
    // the interior-lane masks follow the waves of the step grid: new masks
    // for the new shape first, so that a failure leaves the handle as it was
    const long long nb = (long long)((h->nx + bx - 1) / bx) * ((h->ny + by - 1) / by) * h->nz;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    unsigned long long *lanes = nullptr;
    if (hipMalloc(&lanes, nb * 4 * sizeof(unsigned long long)) != hipSuccess)
      return fail(h, LBM_ENOMEM, "lbm_set_launch: hipMalloc of the interior-lane masks failed");
    if (h->lanes) (void)hipFree(h->lanes);
    h->lanes = lanes;
    h->bx = bx;
    h->by = by;
    const int rc = build_links(h);
    if (rc) return rc;
  }
  if (nontemporal >= 0) h->nt = nontemporal;
  return ensure_partials(h);
}

int lbm_set_halo(lbm_handle *h, int32_t lds) {
  if (!h || (lds != 0 && lds != 1)) return fail(h, LBM_EINVAL, "lbm_set_halo: lds must be 0 or 1");
  if (lds && (h->storage != LBM_STORE_F32 || h->w64 || h->bx != 256 || h->by != 1))
    return fail(h, LBM_EINVAL, "lbm_set_halo: the LDS x halo needs FP32 storage, 32-bit offsets and 256x1 blocks");
  h->lds = lds != 0;
  return LBM_OK;
}

int lbm_set_pressure_forcing(lbm_handle *h, int32_t on, double p_inlet, double p_outlet) {
  if (!h || (on != 0 && on != 1)) return fail(h, LBM_EINVAL, "lbm_set_pressure_forcing: on must be 0 or 1");
  if (on) {
    if (h->Q != 9) return fail(h, LBM_EINVAL, "lbm_set_pressure_forcing: D2Q9 only (the reference's forcing is 2D)");
    if (h->nx < 4) return fail(h, LBM_EINVAL, "lbm_set_pressure_forcing: nx must be at least 4");
    if (h->F[0] != 0 || h->F[1] != 0 || h->F[2] != 0)
      return fail(h, LBM_EINVAL, "lbm_set_pressure_forcing: not with a body force (the correction reads the moments "
                                 "the collision keeps)");
    if (!(p_inlet > 0) || !(p_outlet > 0)) return fail(h, LBM_EINVAL, "lbm_set_pressure_forcing: pressures must be > 0");
    if (!h->fmac) {
      HIPCHK(h, hipMalloc(&h->fmac, (size_t)2 * h->ny * 3 * sizeof(double)));
      HIPCHK(h, hipMemset(h->fmac, 0, (size_t)2 * h->ny * 3 * sizeof(double)));
    }
  }
  h->forcing = on != 0;
  h->p_in = p_inlet;
  h->p_out = p_outlet;
  return LBM_OK;
}

int lbm_set_index_width(lbm_handle *h, int32_t bits) {
  if (!h) return fail(h, LBM_EINVAL, "lbm_set_index_width: null handle");
  if (bits == 64) {
    h->w64 = true;
  } else if (bits == 32) {
    if ((unsigned long long)h->plane * (unsigned long long)h->esize > 0xffffffffull)
      return fail(h, LBM_EINVAL, "lbm_set_index_width: a population plane of this domain exceeds 4 GiB; 32-bit offsets cannot address it");
    h->w64 = false;
  } else {
    return fail(h, LBM_EINVAL, "lbm_set_index_width: bits must be 32 or 64");
  }
  return LBM_OK;
}

int lbm_info(const lbm_handle *h, int64_t *ncell, int32_t *q, int32_t *bytes_per_pop) {
  if (!h) return fail(nullptr, LBM_EINVAL, "lbm_info: null handle");
  if (ncell) *ncell = h->ncell;
  if (q) *q = h->Q;
  if (bytes_per_pop) *bytes_per_pop = h->esize;
  return LBM_OK;
}

int lbm_synchronize(lbm_handle *h) {
  if (!h) return fail(h, LBM_EINVAL, "lbm_synchronize: null handle");
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return LBM_OK;
}

int lbm_destroy(lbm_handle *h) {
  if (!h) return LBM_OK;
  if (h->stream) hipStreamSynchronize(h->stream);
  for (int b = 0; b < 2; ++b)
    if (h->buf[b]) hipFree(h->buf[b]);
  if (h->flags) hipFree(h->flags);
  if (h->fmac) hipFree(h->fmac);
  if (h->links) hipFree(h->links);
  if (h->lanes) hipFree(h->lanes);
  for (int b = 0; b < 2; ++b)
    if (h->snap[b]) hipFree(h->snap[b]);
  if (h->partials) hipFree(h->partials);
  if (h->sums_d) hipFree(h->sums_d);
  if (h->io) hipFree(h->io);
  if (h->dface) hipFree(h->dface);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
  return LBM_OK;
}

const char *lbm_last_error(const lbm_handle *h) { return h ? h->err.c_str() : g_err.c_str(); }

int lbm_copy_bench(int64_t nbytes, int32_t reps, double *ms) {
  if (nbytes < 16 || reps < 1 || !ms) return fail(nullptr, LBM_EINVAL, "lbm_copy_bench: bad argument");
  const long long n = nbytes / 16;
  DevBuf a, b;
  HIPCHK(nullptr, a.alloc(n * 16));
  HIPCHK(nullptr, b.alloc(n * 16));
  HIPCHK(nullptr, hipMemset(a.p, 0, n * 16));
  EventPair ev;
  HIPCHK(nullptr, ev.create());
  const float4 *in = static_cast<const float4 *>(a.p);
  float4 *out = static_cast<float4 *>(b.p);
  const long long grid = (n + 255) / 256;
  hipStream_t s = nullptr;  // the null stream: no stream to destroy on an error path
  hipLaunchKernelGGL(k_copy, dim3((unsigned)grid), dim3(256), 0, s, in, out, n);
  HIPCHK(nullptr, hipEventRecord(ev.e[0], s));
  for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_copy, dim3((unsigned)grid), dim3(256), 0, s, in, out, n);
  HIPCHK(nullptr, hipEventRecord(ev.e[1], s));
  HIPCHK(nullptr, hipEventSynchronize(ev.e[1]));
  float t = 0;
  HIPCHK(nullptr, hipEventElapsedTime(&t, ev.e[0], ev.e[1]));
  *ms = t / reps;
  return LBM_OK;
}

}  // extern "C"

// batched uneven domains (include/lbm_batch.h)
#include "../../include/lbm_batch.h"
#include "lbm_batch_impl.h"
