// Poisson (screened Poisson / diffusion-reaction) lattice-Boltzmann step for
// gfx950 and the C ABI of include/lbm_poisson.h.
//
// The reference solves -D lap(phi) + kappa^2-type reaction with a scalar
// BGK scheme (src/lbm/solver.cpp:542-545, 589-611): phi = sum_{i<Q-1} fold_i /
// (1 - w_rest); feq_i = w_i phi, feq_rest = (w_rest - 1) phi; f = (1 - omega)
// fold + omega feq + source * w'_i * phi for the moving populations.  Here
// one kernel launch is one step in pull form on a structured window; the
// cells of a face with a boundary object are rebuilt from their
// extrapolation cell (non-equilibrium extrapolation, bnd_dirichlet.h:367-376)
// in the same pass.  FP64 storage, unshifted populations (the reference's
// values, ~w phi).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/lbm_hip.h"
#include "../../include/lbm_poisson.h"

namespace lbmp {

template <int Q>
struct PLat;
template <>
struct PLat<3> {
  static constexpr int D = 1;
  static constexpr int cx[3] = {-1, 1, 0};
  static constexpr int cy[3] = {0, 0, 0};
  static constexpr double w[3] = {1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0};
  static constexpr double wp[3] = {0.5, 0.5, 0.0};
};
template <>
struct PLat<5> {
  static constexpr int D = 2;
  static constexpr int cx[5] = {-1, 1, 0, 0, 0};
  static constexpr int cy[5] = {0, 0, -1, 1, 0};
  static constexpr double w[5] = {1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 3.0};
  static constexpr double wp[5] = {0.25, 0.25, 0.25, 0.25, 0.0};
};
template <>
struct PLat<9> {
  static constexpr int D = 2;
  static constexpr int cx[9] = {-1, 1, 0, 0, 1, 1, -1, -1, 0};
  static constexpr int cy[9] = {0, 0, -1, 1, 1, -1, -1, 1, 0};
  static constexpr double w[9] = {1.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0, 1.0 / 36.0,
                                  1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 4.0 / 9.0};
  static constexpr double wp[9] = {0.125, 0.125, 0.125, 0.125, 0.125, 0.125, 0.125, 0.125, 0.0};
};

struct PArgs {
  const double *src;
  double *dst;
  long long plane;  // elements between population planes
  int nx, ny, px;
  const double *phi_old;   // phi of the previous step
  const double *phi_old2;  // phi of the step before that
  double *phi_new;         // phi of this step
  double omega, source;
  int kind[4], prio[4];   // boundary object per face (0: none)
  const double *val[4];   // Dirichlet values along each face
  double *partials;       // check steps: per-block sums of the change of vars
};

// main-direction vectors of the grid (cartesian::dirVec order, 2D diagonals
// 4..7 = +x+y, +x-y, -x-y, -x+y)
__device__ __forceinline__ void dir_vec(int d, int &dx, int &dy) {
  const int tx[8] = {-1, 1, 0, 0, 1, 1, -1, -1};
  const int ty[8] = {0, 0, -1, 1, 1, -1, -1, 1};
  dx = tx[d];
  dy = ty[d];
}
// cartesian::oppositeDir<NDIM> (include/common/math/cartesian.h:22-34)
__device__ __forceinline__ int opposite_dir(int d, int dim) {
  if (d < 2 * dim) return d + 1 - 2 * (d % 2);
  return d > 5 ? d - 2 : d + 2;
}

// Whether the object on face f recomputes vars of cell (x, y) from the
// post-streaming populations while it applies: layer 1 holds its
// extrapolation cells (every object, bnd_dirichlet.h:359), layer 2 the
// second cells of a Neumann object (bnd_neumann.h:50-56); along the face both
// span the cells 1 .. n-2, the corners' extrapolation cells being diagonal.
__device__ __forceinline__ bool recomputed(const PArgs &a, int f, int layer, int x, int y, int dim) {
  const bool yin = dim == 1 || (y >= 1 && y <= a.ny - 2), xin = x >= 1 && x <= a.nx - 2;
  switch (f) {
    case 0: return x == layer && yin;
    case 1: return x == a.nx - 1 - layer && yin;
    case 2: return y == layer && xin;
    default: return y == a.ny - 1 - layer && xin;
  }
}

// post-streaming populations of an interior cell (all upstream cells exist)
template <int Q>
__device__ __forceinline__ void gather(const PArgs &a, int x, int y, double (&g)[Q]) {
  using L = PLat<Q>;
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const long long s = (long long)(x - L::cx[i]) + (long long)a.px * (y - L::cy[i]);
    g[i] = __builtin_nontemporal_load(&a.src[i * a.plane + s]);  // streamed once: 2.195 -> 2.076 ms at 8192^2
  }
}

template <int Q>
__device__ __forceinline__ double phi_of(const double (&g)[Q]) {
  double s = 0;
#pragma unroll
  for (int i = 0; i < Q - 1; ++i) s += g[i];
  return (1.0 / (1.0 - PLat<Q>::w[Q - 1])) * s;
}

template <int Q, bool CHECK>
__global__ void __launch_bounds__(256) k_poisson_step(const PArgs a) {
  using L = PLat<Q>;
  const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
  double red = 0;
  if (x < a.nx) {
    const long long cell = (long long)x + (long long)a.px * y;
    // faces the cell lies on, in direction order -x +x -y +y
    bool on[4] = {x == 0, x == a.nx - 1, L::D == 2 && y == 0, L::D == 2 && y == a.ny - 1};
    int ff = -1;
#pragma unroll
    for (int f = 0; f < 2 * L::D; ++f)
      if (on[f] && a.kind[f] != 0 && (ff < 0 || a.prio[f] > a.prio[ff])) ff = f;
    double g[Q];
    if (ff < 0) {
      gather<Q>(a, x, y, g);
    } else {
      // extrapolation direction (bnd_dirichlet.h:289-312): the first missing
      // main direction; a second one makes it the corner diagonal
      int ed = -1;
#pragma unroll
      for (int d = 0; d < 2 * L::D; ++d) {
        if (!on[d]) continue;
        if (ed < 0) ed = d;
        else if (ed == 0 && d == 2) ed = 6;
        else if (ed == 0 && d == 3) ed = 7;
        else if (ed == 1 && d == 3) ed = 4;
        else if (ed == 1 && d == 2) ed = 5;
      }
      const int in = opposite_dir(ed, L::D);
      int dx, dy;
      dir_vec(in, dx, dy);
      const int ex = x + dx, ey = y + dy;
      double ge[Q];
      gather<Q>(a, ex, ey, ge);
      const double pe = phi_of<Q>(ge);
      double pb;
      if (a.kind[ff] == LBM_POISSON_DIRICHLET) {
        pb = a.val[ff][ff < 2 ? y : x];
      } else {
        // Neumann, zero gradient (bnd_neumann.h:50-58): vars(e) and the
        // post-streaming phi two cells in.  vars(e) is the previous step's
        // phi unless an object applied earlier in the same step recomputed
        // it from the post-streaming populations (then pe)
        double g2[Q];
        gather<Q>(a, ex + dx, ey + dy, g2);
        bool fresh = false;
#pragma unroll
        for (int f = 0; f < 2 * L::D; ++f)
          if (a.kind[f] && a.prio[f] < a.prio[ff])
            fresh = fresh || recomputed(a, f, 1, ex, ey, L::D) ||
                    (a.kind[f] == LBM_POISSON_NEUMANN && recomputed(a, f, 2, ex, ey, L::D));
        const long long ecell = (long long)ex + (long long)a.px * ey;
        pb = (4.0 * (fresh ? pe : a.phi_old[ecell]) - phi_of<Q>(g2) + 0.0) / 3.0;
      }
#pragma unroll
      for (int i = 0; i < Q - 1; ++i) g[i] = L::w[i] * pb + ge[i] - L::w[i] * pe;
      g[Q - 1] = (L::w[Q - 1] - 1.0) * pb + ge[Q - 1] - (L::w[Q - 1] - 1.0) * pe;
    }
    const double phi = phi_of<Q>(g);
    if (CHECK) {
      // sumAbsDiff (solver.cpp:809-815) as the check of the next iteration
      // sees it: vars holds the previous step's phi, except where a boundary
      // object rewrote it from this step's post-streaming populations (its
      // face cells and the cells it recomputes), and varsold holds the same
      // one step earlier
      bool touched = ff >= 0;
#pragma unroll
      for (int f = 0; f < 2 * L::D; ++f)
        if (a.kind[f])
          touched = touched || recomputed(a, f, 1, x, y, L::D) ||
                    (a.kind[f] == LBM_POISSON_NEUMANN && recomputed(a, f, 2, x, y, L::D));
      red = touched ? fabs(phi - a.phi_old[cell]) : fabs(a.phi_old[cell] - a.phi_old2[cell]);
    }
    a.phi_new[cell] = phi;
    const double om = a.omega;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const double feq = (i == Q - 1 ? L::w[i] - 1.0 : L::w[i]) * phi;
      double f = (1.0 - om) * g[i] + om * feq;
      if (i < Q - 1) f += a.source * L::wp[i] * phi;
      __builtin_nontemporal_store(f, &a.dst[i * a.plane + cell]);
    }
  }
  if (CHECK) {
    __shared__ double s[256];
    s[threadIdx.x] = red;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
      if ((int)threadIdx.x < k) s[threadIdx.x] += s[threadIdx.x + k];
      __syncthreads();
    }
    if (threadIdx.x == 0) a.partials[blockIdx.x + (long long)gridDim.x * blockIdx.y] = s[0];
  }
}

// initialCondition: fold = feq(phi0); then phi and the collision of the
// first step.  phi0 itself (the initial vars) also goes to vars0, the slot
// the first step reads as the step before the previous one (the first
// check's varsold)
template <int Q>
__global__ void __launch_bounds__(256) k_poisson_init(const PArgs a, const double *phi0, double *vars0) {
  using L = PLat<Q>;
  const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
  if (x >= a.nx) return;
  const long long cell = (long long)x + (long long)a.px * y;
  const double p0 = phi0[x + (long long)a.nx * y];
  vars0[cell] = p0;
  double g[Q];
#pragma unroll
  for (int i = 0; i < Q; ++i) g[i] = (i == Q - 1 ? L::w[i] - 1.0 : L::w[i]) * p0;
  const double phi = phi_of<Q>(g);
  a.phi_new[cell] = phi;
  const double om = a.omega;
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const double feq = (i == Q - 1 ? L::w[i] - 1.0 : L::w[i]) * phi;
    double f = (1.0 - om) * g[i] + om * feq;
    if (i < Q - 1) f += a.source * L::wp[i] * phi;
    a.dst[i * a.plane + cell] = f;
  }
}

// fixed-order sum of the block partials
__global__ void __launch_bounds__(1024) k_poisson_sum(const double *partials, long long n, double *out) {
  __shared__ double s[1024];
  double acc = 0;
  for (long long b = threadIdx.x; b < n; b += 1024) acc += partials[b];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int k = 512; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) s[threadIdx.x] += s[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s[0];
}

// phi (padded layout) -> compact host order
__global__ void __launch_bounds__(256) k_poisson_compact(const double *phi, int nx, int ny, int px, double *out) {
  const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
  if (x < nx) out[x + (long long)nx * y] = phi[x + (long long)px * y];
}

}  // namespace lbmp

using namespace lbmp;

struct lbm_poisson {
  int Q = 0, nx = 0, ny = 0, px = 0;
  long long plane = 0;
  double omega = 1, source = 0;
  double *buf[2] = {nullptr, nullptr};
  double *phi[3] = {nullptr, nullptr, nullptr};
  int cur = 0;   // buffer holding the current state
  int pcur = 0;  // phi slot of the current state; pcur+2, pcur+1 (mod 3) the two before
  int kind[4] = {0, 0, 0, 0}, prio[4] = {0, 0, 0, 0};
  double *val[4] = {nullptr, nullptr, nullptr, nullptr};
  double *partials = nullptr, *sum = nullptr, *io = nullptr;
  bool faces_set = false, ready = false;
  hipStream_t stream = nullptr;
  std::string err;
};

static thread_local std::string g_perr;

static int pfail(lbm_poisson *h, int code, const std::string &msg) {
  if (h) h->err = msg;
  g_perr = msg;
  return code;
}

#define PCHK(h, call)                                                                       \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess) return pfail((h), LBM_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

static PArgs pargs(const lbm_poisson *h) {
  PArgs a;
  memset(&a, 0, sizeof(a));
  a.src = h->buf[h->cur];
  a.dst = h->buf[h->cur ^ 1];
  a.plane = h->plane;
  a.nx = h->nx;
  a.ny = h->ny;
  a.px = h->px;
  a.phi_old = h->phi[h->pcur];
  a.phi_old2 = h->phi[(h->pcur + 2) % 3];
  a.phi_new = h->phi[(h->pcur + 1) % 3];
  a.omega = h->omega;
  a.source = h->source;
  for (int f = 0; f < 4; ++f) {
    a.kind[f] = h->kind[f];
    a.prio[f] = h->prio[f];
    a.val[f] = h->val[f];
  }
  a.partials = h->partials;
  return a;
}

static dim3 pgrid(const lbm_poisson *h) { return dim3((h->nx + 255) / 256, h->ny, 1); }

template <bool CHECK>
static void launch_step(lbm_poisson *h) {
  const PArgs a = pargs(h);
  const dim3 g = pgrid(h), b(256);
  if (h->Q == 3) hipLaunchKernelGGL((k_poisson_step<3, CHECK>), g, b, 0, h->stream, a);
  else if (h->Q == 5) hipLaunchKernelGGL((k_poisson_step<5, CHECK>), g, b, 0, h->stream, a);
  else hipLaunchKernelGGL((k_poisson_step<9, CHECK>), g, b, 0, h->stream, a);
  h->cur ^= 1;
  h->pcur = (h->pcur + 1) % 3;
}

extern "C" {

int lbm_poisson_create(const lbm_poisson_params *p, lbm_poisson **out) {
  if (!p || !out) return pfail(nullptr, LBM_EINVAL, "lbm_poisson_create: null argument");
  *out = nullptr;
  if (p->lattice != 3 && p->lattice != 5 && p->lattice != 9)
    return pfail(nullptr, LBM_EINVAL, "lbm_poisson_create: lattice must be 3 (D1Q3), 5 (D2Q5) or 9 (D2Q9)");
  const int dim = p->lattice == 3 ? 1 : 2;
  if (p->nx < 5 || (dim == 1 && p->ny != 1) || (dim == 2 && p->ny < 5))
    return pfail(nullptr, LBM_EINVAL, "lbm_poisson_create: extents >= 5 (ny = 1 for D1Q3)");
  if (!(p->omega > 0 && p->omega < 2)) return pfail(nullptr, LBM_EINVAL, "lbm_poisson_create: omega must be in (0,2)");
  lbm_poisson *h = new lbm_poisson();
  h->Q = p->lattice;
  h->nx = p->nx;
  h->ny = p->ny;
  h->px = (p->nx + 63) / 64 * 64;
  h->plane = (long long)h->px * h->ny;
  h->omega = p->omega;
  h->source = p->source;
  auto bail = [&](int code, const std::string &msg) {
    pfail(nullptr, code, msg);
    lbm_poisson_destroy(h);
    return code;
  };
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) return bail(LBM_EHIP, "lbm_poisson_create: stream");
  const size_t pb = (size_t)h->plane * h->Q * sizeof(double);
  for (int k = 0; k < 2; ++k) {
    if (hipMalloc(&h->buf[k], pb) != hipSuccess) return bail(LBM_ENOMEM, "lbm_poisson_create: device memory");
    if (hipMemset(h->buf[k], 0, pb) != hipSuccess) return bail(LBM_EHIP, "lbm_poisson_create: memset");
  }
  for (int k = 0; k < 3; ++k) {
    if (hipMalloc(&h->phi[k], h->plane * sizeof(double)) != hipSuccess)
      return bail(LBM_ENOMEM, "lbm_poisson_create: device memory");
    if (hipMemset(h->phi[k], 0, h->plane * sizeof(double)) != hipSuccess)
      return bail(LBM_EHIP, "lbm_poisson_create: memset");
  }
  const dim3 g = pgrid(h);
  if (hipMalloc(&h->partials, (size_t)g.x * g.y * sizeof(double)) != hipSuccess ||
      hipMalloc(&h->sum, sizeof(double)) != hipSuccess ||
      hipMalloc(&h->io, (size_t)h->nx * h->ny * sizeof(double)) != hipSuccess)
    return bail(LBM_ENOMEM, "lbm_poisson_create: device memory");
  *out = h;
  return LBM_OK;
}

int lbm_poisson_set_faces(lbm_poisson *h, const lbm_poisson_face *faces, int32_t n) {
  if (!h || !faces || n < 0) return pfail(h, LBM_EINVAL, "lbm_poisson_set_faces: bad argument");
  const int nf = h->Q == 3 ? 2 : 4;
  int kind[4] = {0, 0, 0, 0}, prio[4] = {0, 0, 0, 0};
  const double *val[4] = {nullptr, nullptr, nullptr, nullptr};
  for (int i = 0; i < n; ++i) {
    const lbm_poisson_face &f = faces[i];
    if (f.face < 0 || f.face >= nf) return pfail(h, LBM_EINVAL, "lbm_poisson_set_faces: face index out of range");
    if (kind[f.face]) return pfail(h, LBM_EINVAL, "lbm_poisson_set_faces: face described twice");
    if (f.kind != LBM_POISSON_DIRICHLET && f.kind != LBM_POISSON_NEUMANN)
      return pfail(h, LBM_EINVAL, "lbm_poisson_set_faces: kind must be LBM_POISSON_DIRICHLET or LBM_POISSON_NEUMANN");
    if (f.kind == LBM_POISSON_DIRICHLET && !f.value)
      return pfail(h, LBM_EINVAL, "lbm_poisson_set_faces: a Dirichlet face needs its values");
    kind[f.face] = f.kind;
    prio[f.face] = f.priority;
    val[f.face] = f.value;
  }
  for (int f = 0; f < nf; ++f)
    if (!kind[f]) return pfail(h, LBM_EINVAL, "lbm_poisson_set_faces: every face needs a boundary object");
  PCHK(h, hipStreamSynchronize(h->stream));
  for (int f = 0; f < nf; ++f) {
    if (h->val[f]) PCHK(h, hipFree(h->val[f]));
    h->val[f] = nullptr;
    if (kind[f] == LBM_POISSON_DIRICHLET) {
      const int len = f < 2 ? h->ny : h->nx;
      PCHK(h, hipMalloc(&h->val[f], len * sizeof(double)));
      PCHK(h, hipMemcpy(h->val[f], val[f], len * sizeof(double), hipMemcpyHostToDevice));
    }
    h->kind[f] = kind[f];
    h->prio[f] = prio[f];
  }
  h->faces_set = true;
  return LBM_OK;
}

int lbm_poisson_init(lbm_poisson *h, const double *phi0) {
  if (!h || !phi0) return pfail(h, LBM_EINVAL, "lbm_poisson_init: null argument");
  if (!h->faces_set) return pfail(h, LBM_EINVAL, "lbm_poisson_init: lbm_poisson_set_faces first");
  PCHK(h, hipMemcpyAsync(h->io, phi0, (size_t)h->nx * h->ny * sizeof(double), hipMemcpyHostToDevice, h->stream));
  h->cur = 0;
  h->pcur = 0;
  PArgs a = pargs(h);
  a.dst = h->buf[0];
  a.phi_new = h->phi[0];
  const dim3 g = pgrid(h), b(256);
  if (h->Q == 3) hipLaunchKernelGGL(k_poisson_init<3>, g, b, 0, h->stream, a, (const double *)h->io, h->phi[2]);
  else if (h->Q == 5) hipLaunchKernelGGL(k_poisson_init<5>, g, b, 0, h->stream, a, (const double *)h->io, h->phi[2]);
  else hipLaunchKernelGGL(k_poisson_init<9>, g, b, 0, h->stream, a, (const double *)h->io, h->phi[2]);
  PCHK(h, hipGetLastError());
  PCHK(h, hipStreamSynchronize(h->stream));
  h->ready = true;
  return LBM_OK;
}

int lbm_poisson_step(lbm_poisson *h, int32_t nsteps) {
  if (!h || nsteps < 0) return pfail(h, LBM_EINVAL, "lbm_poisson_step: bad argument");
  if (!h->ready) return pfail(h, LBM_EINVAL, "lbm_poisson_step: lbm_poisson_init first");
  for (int i = 0; i < nsteps; ++i) launch_step<false>(h);
  PCHK(h, hipGetLastError());
  return LBM_OK;
}

int lbm_poisson_step_check(lbm_poisson *h, double *abs_diff) {
  if (!h || !abs_diff) return pfail(h, LBM_EINVAL, "lbm_poisson_step_check: null argument");
  if (!h->ready) return pfail(h, LBM_EINVAL, "lbm_poisson_step_check: lbm_poisson_init first");
  launch_step<true>(h);
  const dim3 g = pgrid(h);
  hipLaunchKernelGGL(k_poisson_sum, dim3(1), dim3(1024), 0, h->stream, (const double *)h->partials,
                     (long long)g.x * g.y, h->sum);
  PCHK(h, hipGetLastError());
  PCHK(h, hipMemcpyAsync(abs_diff, h->sum, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  PCHK(h, hipStreamSynchronize(h->stream));
  if (!std::isfinite(*abs_diff)) return pfail(h, LBM_DIVERGED, "non-finite values (solution diverged)");
  return LBM_OK;
}

int lbm_poisson_get(lbm_poisson *h, double *phi) {
  if (!h || !phi) return pfail(h, LBM_EINVAL, "lbm_poisson_get: null argument");
  hipLaunchKernelGGL(k_poisson_compact, pgrid(h), dim3(256), 0, h->stream, (const double *)h->phi[h->pcur], h->nx, h->ny,
                     h->px, h->io);
  PCHK(h, hipGetLastError());
  PCHK(h, hipMemcpyAsync(phi, h->io, (size_t)h->nx * h->ny * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  PCHK(h, hipStreamSynchronize(h->stream));
  return LBM_OK;
}

int lbm_poisson_destroy(lbm_poisson *h) {
  if (!h) return LBM_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  for (int k = 0; k < 2; ++k)
    if (h->buf[k]) (void)hipFree(h->buf[k]);
  for (int k = 0; k < 3; ++k)
    if (h->phi[k]) (void)hipFree(h->phi[k]);
  for (int f = 0; f < 4; ++f)
    if (h->val[f]) (void)hipFree(h->val[f]);
  if (h->partials) (void)hipFree(h->partials);
  if (h->sum) (void)hipFree(h->sum);
  if (h->io) (void)hipFree(h->io);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return LBM_OK;
}

const char *lbm_poisson_last_error(const lbm_poisson *h) { return h ? h->err.c_str() : g_perr.c_str(); }

}  // extern "C"
