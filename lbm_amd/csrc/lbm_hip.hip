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
// keeps FP32 and 16-bit storage well conditioned; collisions run in FP32
// (FP64 for FP64 storage).  Checked steps reduce mass, momentum and the
// one-step change of (u, rho) in FP64 with a fixed-order two-stage tree.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/lbm_hip.h"
#include "lbm_lattice.h"

namespace lbm {

constexpr int NSUM = 12;  // mass, mom[3], absdiff[4], diff_sq, norm_sq, nonfinite, fluid_cells

struct FaceDev {
  int kind;
  int prio;
  double u[3];
  double rho;
};

struct KArgs {
  const void *src;
  void *dst;
  const uint8_t *flags;  // padded layout, may be null
  long long plane;       // elements between population planes
  int nx, ny, nz, px;
  int per[3];
  FaceDev face[6];
  double omega;
  double F[3];
  void *snap_in;
  void *snap_out;
  double *partials;
};

template <typename C>
struct Vec4;
template <>
struct Vec4<float> {
  using T = float4;
};
template <>
struct Vec4<double> {
  using T = double4;
};

// --------------------------------------------------------------------------
// wave64 + block reduction of NSUM doubles; block result to partials[block]
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <int NT>
__device__ __forceinline__ void block_reduce_store(double (&v)[NSUM], double *partials) {
  __shared__ double red[NT / 64][NSUM];
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
  const int lane = tid & 63, wid = tid >> 6;
#pragma unroll
  for (int k = 0; k < NSUM; ++k) v[k] = wave_sum(v[k]);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NSUM; ++k) red[wid][k] = v[k];
  }
  __syncthreads();
  const int nw = (blockDim.x * blockDim.y) >> 6;
  if (tid < NSUM) {
    double s = 0;
    for (int w2 = 0; w2 < nw; ++w2) s += red[w2][tid];
    // structure of arrays: partials[k * nblocks + block] (coalesced second stage)
    const long long nb = (long long)gridDim.x * gridDim.y * gridDim.z;
    const long long b = blockIdx.x + (long long)gridDim.x * (blockIdx.y + (long long)gridDim.y * blockIdx.z);
    partials[(long long)tid * nb + b] = s;
  }
}

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

// --------------------------------------------------------------------------
// Collision in shifted variables.  h_i = f_i - w_i, rho = 1 + sum h,
// rho*u = sum c h (+F/2, Guo), heq_i = w_i (drho + rho (3cu + 4.5cu^2 - 1.5u^2)),
// h' = h + omega (heq - h) + (1 - omega/2) w_i [3(c_i - u).F + 9 (c_i.u)(c_i.F)].
// Same algebra as (1-w) fold + w feq of src/lbm/solver.cpp:601-613 with
// eq::defaultEq (src/lbm/equilibrium_func.h:52-54) after subtracting w_i.
template <int Q, typename C>
struct Moments {
  C drho, rho, ux, uy, uz;
};

template <int Q, typename C>
__device__ __forceinline__ Moments<Q, C> moments(const C (&h)[Q], C Fx, C Fy, C Fz) {
  using L = Lat<Q>;
  C drho = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
  for (int j = 0; j < Q; ++j) drho += h[j];
#pragma unroll
  for (int j = 0; j < Q - 1; ++j) {
    if (L::cx[j] == 1) jx += h[j];
    if (L::cx[j] == -1) jx -= h[j];
    if (L::cy[j] == 1) jy += h[j];
    if (L::cy[j] == -1) jy -= h[j];
    if (L::cz[j] == 1) jz += h[j];
    if (L::cz[j] == -1) jz -= h[j];
  }
  Moments<Q, C> m;
  m.drho = drho;
  m.rho = C(1) + drho;
  const C inv = C(1) / m.rho;
  m.ux = (jx + C(0.5) * Fx) * inv;
  m.uy = (jy + C(0.5) * Fy) * inv;
  m.uz = (jz + C(0.5) * Fz) * inv;
  return m;
}

template <int Q, typename ST, bool NT>
__device__ __forceinline__ void collide_store(typename Store<ST>::C (&h)[Q], const Moments<Q, typename Store<ST>::C> &m,
                                              typename Store<ST>::C omega, typename Store<ST>::C Fx,
                                              typename Store<ST>::C Fy, typename Store<ST>::C Fz, bool force,
                                              ST *dst, long long plane, long long cell) {
  using L = Lat<Q>;
  using C = typename Store<ST>::C;
  const C usq = m.ux * m.ux + m.uy * m.uy + m.uz * m.uz;
  const C uF = m.ux * Fx + m.uy * Fy + m.uz * Fz;
  const C fpre = C(1) - C(0.5) * omega;
#pragma unroll
  for (int j = 0; j < Q; ++j) {
    C cu = 0, cF = 0;
    if (L::cx[j] == 1) { cu += m.ux; cF += Fx; }
    if (L::cx[j] == -1) { cu -= m.ux; cF -= Fx; }
    if (L::cy[j] == 1) { cu += m.uy; cF += Fy; }
    if (L::cy[j] == -1) { cu -= m.uy; cF -= Fy; }
    if (L::cz[j] == 1) { cu += m.uz; cF += Fz; }
    if (L::cz[j] == -1) { cu -= m.uz; cF -= Fz; }
    const C wj = C(L::w[j]);
    const C heq = wj * (m.drho + m.rho * (C(3) * cu + C(4.5) * cu * cu - C(1.5) * usq));
    C out = h[j] + omega * (heq - h[j]);
    if (force) out += fpre * wj * (C(3) * (cF - uF) + C(9) * cu * cF);
    if (NT) Store<ST>::st_nt(dst + j * plane + cell, out);
    else Store<ST>::st(dst + j * plane + cell, out);
  }
}

// Zou-He face (Hecht & Harting 2010 form) on shifted populations; axis and
// outward sign of the face normal; velocity face prescribes u, pressure
// face prescribes rho with zero tangential velocity.
template <int Q, typename C>
__device__ void zou_he(C (&h)[Q], int axis, int sign, bool is_vel, C u0, C u1, C u2, C rho_p) {
  using L = Lat<Q>;
  auto cax = [](int j, int a) { return a == 0 ? L::cx[j] : (a == 1 ? L::cy[j] : L::cz[j]); };
  C s0 = 0, sout = 0;
#pragma unroll
  for (int j = 0; j < Q; ++j) {
    const int cn = cax(j, axis) * sign;
    if (cn == 0) s0 += h[j];
    else if (cn > 0) sout += h[j];
  }
  C u[3] = {0, 0, 0};
  C rho;
  if (is_vel) {
    u[0] = u0; u[1] = u1; u[2] = u2;
    const C un = (axis == 0 ? u0 : (axis == 1 ? u1 : u2)) * C(sign);
    rho = (C(1) + s0 + C(2) * sout) / (C(1) + un);
  } else {
    rho = rho_p;
    const C un = (C(1) + s0 + C(2) * sout) / rho - C(1);
    if (axis == 0) u[0] = un * C(sign);
    if (axis == 1) u[1] = un * C(sign);
    if (axis == 2) u[2] = un * C(sign);
  }
  C N[3] = {0, 0, 0};
#pragma unroll
  for (int t = 0; t < Lat<Q>::D; ++t) {
    if (t == axis) continue;
    C mt = 0;
#pragma unroll
    for (int j = 0; j < Q; ++j)
      if (cax(j, axis) == 0) mt += C(cax(j, t)) * h[j];
    N[t] = C(0.5) * mt - rho * u[t] / C(3);
  }
  C hn[Q];
#pragma unroll
  for (int j = 0; j < Q; ++j) hn[j] = h[j];
#pragma unroll
  for (int j = 0; j < Q; ++j) {
    const int cn = cax(j, axis) * sign;
    if (!(cn < 0)) continue;
    C cu = 0, corr = 0;
#pragma unroll
    for (int d = 0; d < Lat<Q>::D; ++d) cu += C(cax(j, d)) * u[d];
#pragma unroll
    for (int t = 0; t < Lat<Q>::D; ++t)
      if (t != axis) corr += C(cax(j, t)) * N[t];
    hn[j] = h[L::opp[j]] + C(6.0 * L::w[j]) * rho * cu - corr;
  }
#pragma unroll
  for (int j = 0; j < Q; ++j) h[j] = hn[j];
}

// velocity u* of the stored post-collision state at a cell (used for the
// anti-bounce-back extrapolation: collision conserves rho and rho*u, so the
// moments of S equal those of the pre-collision populations plus F).
template <int Q, typename ST>
__device__ void cell_velocity(const ST *src, long long plane, long long cell, double Fx, double Fy, double Fz,
                              typename Store<ST>::C (&u)[3]) {
  using C = typename Store<ST>::C;
  C h[Q];
#pragma unroll
  for (int j = 0; j < Q; ++j) h[j] = Store<ST>::ld(src + j * plane + cell);
  const Moments<Q, C> m = moments<Q, C>(h, C(-Fx), C(-Fy), C(-Fz));  // (m_S - F/2)/rho
  u[0] = m.ux; u[1] = m.uy; u[2] = m.uz;
}

// --------------------------------------------------------------------------
// The fused step.  MODE 0: plain; 1: also write (u, rho) snapshot;
// 2: reduce against the snapshot written by the previous step; 3: both
// (reduce against snap_in, write snap_out).
#ifndef LBM_STEP_WAVES_PER_EU
#define LBM_STEP_WAVES_PER_EU 0
#endif
#if LBM_STEP_WAVES_PER_EU > 0
#define LBM_STEP_ATTR __attribute__((amdgpu_waves_per_eu(LBM_STEP_WAVES_PER_EU)))
#else
#define LBM_STEP_ATTR
#endif
template <int Q, typename ST, bool FLAGS, int MODE, bool NT>
__global__ void __launch_bounds__(256) LBM_STEP_ATTR k_step(const KArgs a) {
  using L = Lat<Q>;
  using C = typename Store<ST>::C;
  const ST *__restrict__ src = static_cast<const ST *>(a.src);
  ST *__restrict__ dst = static_cast<ST *>(a.dst);
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  const int z = blockIdx.z;
  const long long plane = a.plane;
  const long long cell = (long long)x + (long long)a.px * ((long long)y + (long long)a.ny * z);
  bool active = x < a.nx && y < a.ny;
  if (FLAGS && active) active = a.flags[cell] == LBM_CELL_FLUID;

  double red[NSUM];
  if (MODE >= 2) {
#pragma unroll
    for (int k = 0; k < NSUM; ++k) red[k] = 0;
  }

  if (active) {
    // neighbour coordinates per axis with periodic wrap; v* = exists
    int xm = x - 1, xp = x + 1, ym = y - 1, yp = y + 1, zm = z - 1, zp = z + 1;
    bool vxm = true, vxp = true, vym = true, vyp = true, vzm = true, vzp = true;
    if (x == 0) { if (a.per[0]) xm = a.nx - 1; else vxm = false; }
    if (x == a.nx - 1) { if (a.per[0]) xp = 0; else vxp = false; }
    if (y == 0) { if (a.per[1]) ym = a.ny - 1; else vym = false; }
    if (y == a.ny - 1) { if (a.per[1]) yp = 0; else vyp = false; }
    if (Q == 19) {
      if (z == 0) { if (a.per[2]) zm = a.nz - 1; else vzm = false; }
      if (z == a.nz - 1) { if (a.per[2]) zp = 0; else vzp = false; }
    }
    C h[Q];
    unsigned missing = 0;
#pragma unroll
    for (int j = 0; j < Q; ++j) {
      const int cx = L::cx[j], cy = L::cy[j], cz = L::cz[j];
      const int sx = cx > 0 ? xm : (cx < 0 ? xp : x);
      const int sy = cy > 0 ? ym : (cy < 0 ? yp : y);
      const int sz = cz > 0 ? zm : (cz < 0 ? zp : z);
      bool valid = (cx > 0 ? vxm : (cx < 0 ? vxp : true)) && (cy > 0 ? vym : (cy < 0 ? vyp : true)) &&
                   (cz > 0 ? vzm : (cz < 0 ? vzp : true));
      const long long s = (long long)sx + (long long)a.px * ((long long)sy + (long long)a.ny * sz);
      // A population streams iff its upstream cell is fluid.  This is the
      // symmetric rule; the reference composes diagonal neighbours from main
      // ones (src/cartesiangrid.h:466-486), which next to staircase obstacles
      // makes the neighbour relation one-sided and leaves some populations
      // unwritten for a step.
      if (FLAGS && valid && j != Q - 1) valid = a.flags[s] == LBM_CELL_FLUID;
      const long long addr = valid ? (long long)j * plane + s : (long long)L::opp[j] * plane + cell;
      h[j] = NT ? Store<ST>::ld_nt(src + addr) : Store<ST>::ld(src + addr);
      if (!valid) missing |= 1u << j;
    }

    const C Fx = C(a.F[0]), Fy = C(a.F[1]), Fz = C(a.F[2]);
    const bool force = a.F[0] != 0.0 || a.F[1] != 0.0 || a.F[2] != 0.0;

    if (missing) {
      // boundary substitution for populations whose upstream cell is missing
      C uext[3] = {0, 0, 0};
      int abb_face = -1;
#pragma unroll
      for (int j = 0; j < Q - 1; ++j) {
        if (!((missing >> j) & 1u)) continue;
        const int cx = L::cx[j], cy = L::cy[j], cz = L::cz[j];
        const bool ix = (cx > 0 && !vxm) || (cx < 0 && !vxp);
        const bool iy = (cy > 0 && !vym) || (cy < 0 && !vyp);
        const bool iz = (cz > 0 && !vzm) || (cz < 0 && !vzp);
        int f = -1;
        // crossed faces: the higher priority wins, ties to the lower axis
        if (ix) f = cx > 0 ? 0 : 1;
        if (iy) { const int g = cy > 0 ? 2 : 3; if (f < 0 || a.face[g].prio > a.face[f].prio) f = g; }
        if (iz) { const int g = cz > 0 ? 4 : 5; if (f < 0 || a.face[g].prio > a.face[f].prio) f = g; }
        if (f < 0) continue;  // solid obstacle: plain bounce-back, h[j] already holds S(x, opp j)
        const FaceDev &fd = a.face[f];
        const C wj = C(L::w[j]);
        if (fd.kind == LBM_FACE_WALL) {
          // Ladd moving wall, rho0 = 1 (bnd_dirichlet.h:109-116)
          const C cu = C(cx) * C(fd.u[0]) + C(cy) * C(fd.u[1]) + C(cz) * C(fd.u[2]);
          h[j] += C(6) * wj * cu;
        } else if (fd.kind == LBM_FACE_PRESSURE_ABB) {
          if (abb_face != f) {
            // u extrapolated from the two inward cells (bnd_pressure.h:70-81)
            const int ax = f >> 1, out = (f & 1) ? 1 : -1;
            const int dx = ax == 0 ? -out : 0, dy = ax == 1 ? -out : 0, dz = ax == 2 ? -out : 0;
            const long long n1 = (long long)(x + dx) + (long long)a.px * ((long long)(y + dy) + (long long)a.ny * (z + dz));
            const long long n2 = (long long)(x + 2 * dx) + (long long)a.px * ((long long)(y + 2 * dy) + (long long)a.ny * (z + 2 * dz));
            C u1[3], u2[3];
            cell_velocity<Q, ST>(src, plane, n1, a.F[0], a.F[1], a.F[2], u1);
            cell_velocity<Q, ST>(src, plane, n2, a.F[0], a.F[1], a.F[2], u2);
            for (int d = 0; d < 3; ++d) uext[d] = C(1.5) * u1[d] - C(0.5) * u2[d];
            abb_face = f;
          }
          const C cu = C(cx) * uext[0] + C(cy) * uext[1] + C(cz) * uext[2];
          const C vsq = uext[0] * uext[0] + uext[1] * uext[1] + uext[2] * uext[2];
          // f = -f_post + 2 symmEq(p, u) in shifted form
          h[j] = -h[j] + wj * (C(2) * C(fd.rho) * (C(1) + C(4.5) * cu * cu - C(1.5) * vsq) - C(2));
        }
        // Zou-He faces: overwritten below
      }
      // Zou-He on the first Zou-He face the cell lies on
#pragma unroll 1
      for (int f = 0; f < 2 * L::D; ++f) {
        const int k = a.face[f].kind;
        if (k != LBM_FACE_VELOCITY_ZH && k != LBM_FACE_PRESSURE_ZH) continue;
        const int ax = f >> 1;
        const int coord = ax == 0 ? x : (ax == 1 ? y : z);
        const int n = ax == 0 ? a.nx : (ax == 1 ? a.ny : a.nz);
        if (coord != ((f & 1) ? n - 1 : 0)) continue;
        zou_he<Q, C>(h, ax, (f & 1) ? 1 : -1, k == LBM_FACE_VELOCITY_ZH, C(a.face[f].u[0]), C(a.face[f].u[1]),
                     C(a.face[f].u[2]), C(a.face[f].rho));
        break;
      }
    }

    const Moments<Q, C> m = moments<Q, C>(h, Fx, Fy, Fz);
    using V4 = typename Vec4<C>::T;
    if (MODE == 1 || MODE == 3) {
      V4 v;
      v.x = m.ux; v.y = m.uy; v.z = m.uz; v.w = m.rho;
      static_cast<V4 *>(a.snap_out)[cell] = v;
    }
    if (MODE >= 2) {
      const V4 o = static_cast<const V4 *>(a.snap_in)[cell];
      const double ux = m.ux, uy = m.uy, uz = m.uz, rho = m.rho;
      const double dux = ux - (double)o.x, duy = uy - (double)o.y, duz = uz - (double)o.z, drh = rho - (double)o.w;
      red[0] = rho;
      red[1] = rho * ux;
      red[2] = rho * uy;
      red[3] = rho * uz;
      red[4] = fabs(dux);
      red[5] = fabs(duy);
      red[6] = fabs(duz);
      red[7] = fabs(drh);
      red[8] = dux * dux + duy * duy + duz * duz;
      red[9] = ux * ux + uy * uy + uz * uz;
      red[10] = (isfinite(rho) && isfinite(ux) && isfinite(uy) && isfinite(uz)) ? 0.0 : 1.0;
      red[11] = 1.0;
    }
    collide_store<Q, ST, NT>(h, m, C(a.omega), Fx, Fy, Fz, force, dst, plane, cell);
  }
  if (MODE >= 2) block_reduce_store<256>(red, a.partials);
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
  C h[Q];
#pragma unroll
  for (int j = 0; j < Q; ++j) {
    const C cu = C(L::cx[j]) * u0 + C(L::cy[j]) * u1 + C(L::cz[j]) * u2;
    h[j] = C(L::w[j]) * ((r - C(1)) + r * (C(3) * cu + C(4.5) * cu * cu - C(1.5) * usq));
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
    C h[Q];
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
  if (a.partials) block_reduce_store<256>(red, a.partials);
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
  bool has_flags = false;
  FaceDev face[6];
  int per[3] = {1, 1, 1};
  double omega = 1.0, F[3] = {0, 0, 0};
  void *snap[2] = {nullptr, nullptr};
  int snap_last = 0;  // buffer holding the most recent snapshot
  double *partials = nullptr;
  long long partials_cap = 0;
  double *sums_d = nullptr;
  hipStream_t stream = nullptr;
  int bx = 256, by = 1, nt = 1;  // nontemporal loads/stores measured fastest on gfx950
  long long steps = 0;
  std::string err;
};

static thread_local std::string g_err;

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

static KArgs make_args(lbm_handle *h) {
  KArgs a;
  memset(&a, 0, sizeof(a));
  a.src = h->buf[h->cur];
  a.dst = h->buf[h->cur ^ 1];
  a.flags = h->has_flags ? h->flags : nullptr;
  a.plane = h->plane;
  a.nx = h->nx; a.ny = h->ny; a.nz = h->nz; a.px = h->px;
  for (int i = 0; i < 3; ++i) a.per[i] = h->per[i];
  for (int f = 0; f < 6; ++f) a.face[f] = h->face[f];
  a.omega = h->omega;
  for (int i = 0; i < 3; ++i) a.F[i] = h->F[i];
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
struct StepL {
  static void run(const KArgs &a, dim3 g, dim3 b, hipStream_t s, bool flags, int mode, bool nt) {
    if (flags) {
      if (mode == 0) {
        if (nt) hipLaunchKernelGGL((k_step<Q, ST, true, 0, true>), g, b, 0, s, a);
        else hipLaunchKernelGGL((k_step<Q, ST, true, 0, false>), g, b, 0, s, a);
      } else if (mode == 1) hipLaunchKernelGGL((k_step<Q, ST, true, 1, false>), g, b, 0, s, a);
      else if (mode == 2) hipLaunchKernelGGL((k_step<Q, ST, true, 2, false>), g, b, 0, s, a);
      else hipLaunchKernelGGL((k_step<Q, ST, true, 3, false>), g, b, 0, s, a);
    } else {
      if (mode == 0) {
        if (nt) hipLaunchKernelGGL((k_step<Q, ST, false, 0, true>), g, b, 0, s, a);
        else hipLaunchKernelGGL((k_step<Q, ST, false, 0, false>), g, b, 0, s, a);
      } else if (mode == 1) hipLaunchKernelGGL((k_step<Q, ST, false, 1, false>), g, b, 0, s, a);
      else if (mode == 2) hipLaunchKernelGGL((k_step<Q, ST, false, 2, false>), g, b, 0, s, a);
      else hipLaunchKernelGGL((k_step<Q, ST, false, 3, false>), g, b, 0, s, a);
    }
  }
};
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
  h->omega = p->omega;
  for (int i = 0; i < 3; ++i) h->F[i] = p->force[i];
  for (int f = 0; f < 6; ++f) { memset(&h->face[f], 0, sizeof(FaceDev)); h->face[f].kind = LBM_FACE_PERIODIC; h->face[f].rho = 1.0; }
  h->bx = h->px >= 256 ? 256 : (h->px >= 128 ? 128 : 64);
  h->by = 256 / h->bx;
  int rc = LBM_OK;
  auto bail = [&](int code) { lbm_destroy(h); return code; };
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    fail(nullptr, LBM_EHIP, "lbm_create: hipStreamCreate failed");
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
    h->has_flags = true;
  }
  rc = ensure_partials(h);
  if (rc) return bail(rc);
  *out = h;
  return LBM_OK;
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
    if (d.kind < LBM_FACE_PERIODIC || d.kind > LBM_FACE_PRESSURE_ZH) return fail(h, LBM_EINVAL, "lbm_set_bc: unknown face kind");
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
    const int na = ax == 0 ? h->nx : (ax == 1 ? h->ny : h->nz);
    for (int s = 0; s < 2; ++s) {
      const int k = face[2 * ax + s].kind;
      if (k == LBM_FACE_PRESSURE_ABB && na < 3)
        return fail(h, LBM_EINVAL, "lbm_set_bc: pressure face needs at least 3 cells along its normal");
    }
  }
  memcpy(h->face, face, sizeof(face));
  memcpy(h->per, per, sizeof(per));
  return LBM_OK;
}

int lbm_init_fields(lbm_handle *h, const double *rho, const double *u) {
  if (!h || !rho || !u) return fail(h, LBM_EINVAL, "lbm_init_fields: null argument");
  double *rd = nullptr, *ud = nullptr;
  HIPCHK(h, hipMalloc(&rd, h->ncell * sizeof(double)));
  HIPCHK(h, hipMalloc(&ud, h->ncell * h->dim * sizeof(double)));
  HIPCHK(h, hipMemcpyAsync(rd, rho, h->ncell * sizeof(double), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(ud, u, h->ncell * h->dim * sizeof(double), hipMemcpyHostToDevice, h->stream));
  int rc = ensure_snapshots(h);
  if (rc) return rc;
  KArgs a = make_args(h);
  a.dst = h->buf[h->cur];
  h->snap_last = 0;
  a.snap_out = h->snap[0];
  dim3 g, b;
  launch_dims(h, g, b);
  dispatch_qs<InitL>(h->Q, h->storage, a, g, b, h->stream, (const double *)rd, (const double *)ud);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->stream));
  HIPCHK(h, hipFree(rd));
  HIPCHK(h, hipFree(ud));
  h->steps = 0;
  return LBM_OK;
}

static void fill_sums(const double *s, int dim, lbm_sums *out);

static int ensure_snapshots(lbm_handle *h) {
  const size_t cs = h->storage == LBM_STORE_F64 ? sizeof(double4) : sizeof(float4);
  for (int b = 0; b < 2; ++b)
    if (!h->snap[b]) HIPCHK(h, hipMalloc(&h->snap[b], (size_t)h->px * h->ny * h->nz * cs));
  return LBM_OK;
}

static int launch_step(lbm_handle *h, int mode) {
  KArgs a = make_args(h);
  if (mode != 0) {
    const int rc = ensure_snapshots(h);
    if (rc) return rc;
    a.snap_in = h->snap[h->snap_last];
    a.snap_out = h->snap[h->snap_last ^ 1];
    a.partials = h->partials;
  }
  dim3 g, b;
  launch_dims(h, g, b);
  dispatch_qs<StepL>(h->Q, h->storage, a, g, b, h->stream, h->has_flags, mode, h->nt != 0);
  HIPCHK(h, hipGetLastError());
  if (mode == 1 || mode == 3) h->snap_last ^= 1;
  h->cur ^= 1;
  h->steps++;
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
  out->res_rel = s[9] > 0 ? sqrt(s[8] / s[9]) : (s[8] > 0 ? INFINITY : 0.0);
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

int lbm_get_macros(lbm_handle *h, double *rho, double *u) {
  if (!h || !rho || !u) return fail(h, LBM_EINVAL, "lbm_get_macros: null argument");
  double *rd = nullptr, *ud = nullptr;
  HIPCHK(h, hipMalloc(&rd, h->ncell * sizeof(double)));
  HIPCHK(h, hipMalloc(&ud, h->ncell * h->dim * sizeof(double)));
  KArgs a = make_args(h);
  dim3 g, b;
  launch_dims(h, g, b);
  dispatch_qs<MacroL>(h->Q, h->storage, a, g, b, h->stream, rd, ud);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipMemcpyAsync(rho, rd, h->ncell * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipMemcpyAsync(u, ud, h->ncell * h->dim * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  HIPCHK(h, hipFree(rd));
  HIPCHK(h, hipFree(ud));
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
  double *fd = nullptr;
  const size_t bytes = (size_t)h->ncell * h->Q * sizeof(double);
  HIPCHK(h, hipMalloc(&fd, bytes));
  KArgs a = make_args(h);
  dim3 g, b;
  launch_dims(h, g, b);
  dispatch_qs<GetL>(h->Q, h->storage, a, g, b, h->stream, fd);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipMemcpyAsync(f, fd, bytes, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  HIPCHK(h, hipFree(fd));
  return LBM_OK;
}

int lbm_set_populations(lbm_handle *h, const double *f) {
  if (!h || !f) return fail(h, LBM_EINVAL, "lbm_set_populations: null argument");
  double *fd = nullptr;
  const size_t bytes = (size_t)h->ncell * h->Q * sizeof(double);
  HIPCHK(h, hipMalloc(&fd, bytes));
  HIPCHK(h, hipMemcpyAsync(fd, f, bytes, hipMemcpyHostToDevice, h->stream));
  KArgs a = make_args(h);
  a.dst = h->buf[h->cur];
  dim3 g, b;
  launch_dims(h, g, b);
  dispatch_qs<SetL>(h->Q, h->storage, a, g, b, h->stream, (const double *)fd);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->stream));
  HIPCHK(h, hipFree(fd));
  return LBM_OK;
}

int lbm_time_steps(lbm_handle *h, int32_t nsteps, double *ms) {
  if (!h || !ms || nsteps <= 0) return fail(h, LBM_EINVAL, "lbm_time_steps: bad argument");
  hipEvent_t e0, e1;
  HIPCHK(h, hipEventCreate(&e0));
  HIPCHK(h, hipEventCreate(&e1));
  HIPCHK(h, hipEventRecord(e0, h->stream));
  int rc = lbm_step(h, nsteps);
  if (rc) return rc;
  HIPCHK(h, hipEventRecord(e1, h->stream));
  HIPCHK(h, hipEventSynchronize(e1));
  float t = 0;
  HIPCHK(h, hipEventElapsedTime(&t, e0, e1));
  *ms = t;
  HIPCHK(h, hipEventDestroy(e0));
  HIPCHK(h, hipEventDestroy(e1));
  return LBM_OK;
}

int lbm_set_launch(lbm_handle *h, int32_t block_x, int32_t block_y, int32_t nontemporal) {
  if (!h) return fail(h, LBM_EINVAL, "lbm_set_launch: null handle");
  int bx = block_x > 0 ? block_x : h->bx;
  int by = block_y > 0 ? block_y : h->by;
  if (bx % 64 != 0 || bx * by != 256) return fail(h, LBM_EINVAL, "lbm_set_launch: block_x must be a multiple of 64 and block_x*block_y == 256");
  h->bx = bx;
  h->by = by;
  if (nontemporal >= 0) h->nt = nontemporal;
  return ensure_partials(h);
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
  for (int b = 0; b < 2; ++b)
    if (h->snap[b]) hipFree(h->snap[b]);
  if (h->partials) hipFree(h->partials);
  if (h->sums_d) hipFree(h->sums_d);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
  return LBM_OK;
}

const char *lbm_last_error(const lbm_handle *h) { return h ? h->err.c_str() : g_err.c_str(); }

int lbm_copy_bench(int64_t nbytes, int32_t reps, double *ms) {
  if (nbytes < 16 || reps < 1 || !ms) return fail(nullptr, LBM_EINVAL, "lbm_copy_bench: bad argument");
  const long long n = nbytes / 16;
  float4 *a = nullptr, *b = nullptr;
  HIPCHK(nullptr, hipMalloc(&a, n * 16));
  HIPCHK(nullptr, hipMalloc(&b, n * 16));
  HIPCHK(nullptr, hipMemset(a, 0, n * 16));
  hipStream_t s;
  HIPCHK(nullptr, hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  const long long grid = (n + 255) / 256;
  hipLaunchKernelGGL(k_copy, dim3((unsigned)grid), dim3(256), 0, s, a, b, n);
  hipEvent_t e0, e1;
  HIPCHK(nullptr, hipEventCreate(&e0));
  HIPCHK(nullptr, hipEventCreate(&e1));
  HIPCHK(nullptr, hipEventRecord(e0, s));
  for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_copy, dim3((unsigned)grid), dim3(256), 0, s, a, b, n);
  HIPCHK(nullptr, hipEventRecord(e1, s));
  HIPCHK(nullptr, hipEventSynchronize(e1));
  float t = 0;
  HIPCHK(nullptr, hipEventElapsedTime(&t, e0, e1));
  *ms = t / reps;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipStreamDestroy(s);
  hipFree(a);
  hipFree(b);
  return LBM_OK;
}

}  // extern "C"
