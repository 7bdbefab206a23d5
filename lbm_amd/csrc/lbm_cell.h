// Device building blocks of the lattice-Boltzmann step: reductions, moments,
// BGK collision in shifted variables, Zou-He faces, and the fused per-cell
// pull update used by every step kernel.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "../../include/lbm_hip.h"
#include "lbm_lattice.h"

namespace lbm {

// cell index type: 64-bit, or 32-bit unsigned when plane * sizeof(ST) < 2^32
template <bool W64>
using Idx = typename std::conditional<W64, long long, unsigned>::type;

constexpr int NSUM = 12;  // mass, mom[3], absdiff[4], diff_sq, norm_sq, nonfinite, fluid_cells

struct FaceDev {
  int kind;
  int prio;
  double u[3];
  double rho;
};

// Solid-link mask of a cell (padded layout): bit j (j < Q-1) is set when the
// upstream cell x - c_j (periodic wrap applied) is solid; bit 31 when the
// cell itself is solid.  One coalesced 4-byte load replaces the Q-1 flag
// loads of the neighbours in the step kernels.
constexpr uint32_t LINK_SELF_SOLID = 0x80000000u;

struct KArgs {
  const void *src;
  void *dst;
  const uint8_t *flags;    // padded layout, may be null
  const uint32_t *links;   // solid-link masks (padded layout), null without solids
  long long plane;       // elements between population planes
  int nx, ny, nz, px;
  int per[3];
  // the six faces, in device memory: indexed at run time in the boundary
  // paths (a table among the kernel arguments would then be copied to
  // scratch memory by every thread)
  const FaceDev *face;
  double omega;
  double F[3];
  void *snap_in;
  void *snap_out;
  double *partials;
  // two-launch plain steps: bit i of lanes[w] marks lane i of wave w of the
  // step grid as an interior cell (no face, wrap or solid within one cell)
  const unsigned long long *lanes;
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
__device__ __forceinline__ void block_reduce_store(double (&v)[NSUM], double *partials, long long b, long long nb) {
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
    // structure of arrays: partials[k * nb + b] (coalesced second stage)
    partials[(long long)tid * nb + b] = s;
  }
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

template <int Q, typename C, typename R>
__device__ __forceinline__ Moments<Q, C> moments(const R (&h)[Q], C Fx, C Fy, C Fz) {
  using L = Lat<Q>;
  C drho = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
  for (int j = 0; j < Q; ++j) drho += C(h[j]);
#pragma unroll
  for (int j = 0; j < Q - 1; ++j) {
    if (L::cx[j] == 1) jx += C(h[j]);
    if (L::cx[j] == -1) jx -= C(h[j]);
    if (L::cy[j] == 1) jy += C(h[j]);
    if (L::cy[j] == -1) jy -= C(h[j]);
    if (L::cz[j] == 1) jz += C(h[j]);
    if (L::cz[j] == -1) jz -= C(h[j]);
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

// Population j of the cell at index `cell` of a plane-major buffer.  With
// W64 false the per-plane byte offset is a 32-bit unsigned value added to the
// plane's base address, which is uniform over the wave: the load or store
// then takes one 32-bit VGPR offset plus an SGPR base instead of a 64-bit
// VGPR address per access (the host selects it when plane * sizeof(ST) fits
// in 32 bits).
template <typename ST, bool W64>
__device__ __forceinline__ const ST *pop_ptr(const ST *base, long long plane, int j, Idx<W64> cell) {
  if (W64) return base + (long long)j * plane + (long long)cell;
  return reinterpret_cast<const ST *>(reinterpret_cast<const char *>(base + (long long)j * plane) +
                                      (unsigned)((unsigned)cell * (unsigned)sizeof(ST)));
}
template <typename ST, bool W64>
__device__ __forceinline__ ST *pop_ptr(ST *base, long long plane, int j, Idx<W64> cell) {
  return const_cast<ST *>(pop_ptr<ST, W64>(static_cast<const ST *>(base), plane, j, cell));
}

template <int Q, typename ST, bool NT, bool W64 = true>
__device__ __forceinline__ void collide_store(const typename Store<ST>::R (&h)[Q], const Moments<Q, typename Store<ST>::C> &m,
                                              typename Store<ST>::C omega, typename Store<ST>::C Fx,
                                              typename Store<ST>::C Fy, typename Store<ST>::C Fz, bool force,
                                              ST *dst, long long plane, Idx<W64> cell) {
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
    const C hj = C(h[j]);
    C out = hj + omega * (heq - hj);
    if (force) out += fpre * wj * (C(3) * (cF - uF) + C(9) * cu * cF);
    ST *p = pop_ptr<ST, W64>(dst, plane, j, cell);
    if (NT) Store<ST>::st_nt(p, out);
    else Store<ST>::st(p, out);
  }
}

// Zou-He face (Hecht & Harting 2010 form) on shifted populations; axis and
// outward sign of the face normal; velocity face prescribes u, pressure
// face prescribes rho with zero tangential velocity.
template <int Q, typename C>
__device__ __forceinline__ void zou_he(C (&h)[Q], int axis, int sign, bool is_vel, C u0, C u1, C u2, C rho_p) {
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

// density and velocity u* of the stored post-collision state at a cell,
// accumulated load by load (no population array held in registers)
template <int Q, typename ST>
__device__ __forceinline__ void cell_rho_u(const ST *src, long long plane, long long cell, double Fx, double Fy, double Fz,
                           typename Store<ST>::C &rho, typename Store<ST>::C (&u)[3]) {
  using L = Lat<Q>;
  using C = typename Store<ST>::C;
  C d = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
  for (int j = 0; j < Q; ++j) {
    const C v = C(Store<ST>::ld(src + j * plane + cell));
    d += v;
    if (L::cx[j]) jx += C(L::cx[j]) * v;
    if (L::cy[j]) jy += C(L::cy[j]) * v;
    if (L::cz[j]) jz += C(L::cz[j]) * v;
  }
  rho = C(1) + d;
  const C inv = C(1) / rho;
  u[0] = (jx - C(0.5) * C(Fx)) * inv;
  u[1] = (jy - C(0.5) * C(Fy)) * inv;
  u[2] = (jz - C(0.5) * C(Fz)) * inv;
}

// velocity u* of the stored post-collision state at a cell (used for the
// anti-bounce-back extrapolation: collision conserves rho and rho*u, so the
// moments of S equal those of the pre-collision populations plus F).
template <int Q, typename ST>
__device__ __forceinline__ void cell_velocity(const ST *src, long long plane, long long cell, double Fx, double Fy, double Fz,
                              typename Store<ST>::C (&u)[3]) {
  using C = typename Store<ST>::C;
  typename Store<ST>::R h[Q];
#pragma unroll
  for (int j = 0; j < Q; ++j) h[j] = Store<ST>::ld(src + j * plane + cell);
  const Moments<Q, C> m = moments<Q, C>(h, C(-Fx), C(-Fy), C(-Fz));  // (m_S - F/2)/rho
  u[0] = m.ux; u[1] = m.uy; u[2] = m.uz;
}

// Non-equilibrium extrapolation faces (Guo, Zheng & Shi 2002): all
// populations of a face cell are rebuilt from its inward neighbour n1,
//   f_i = f_i(n1) + feq_i(rho_b, u_b) - feq_i(rho1, u1),
// with (rho_b, u_b) = (rho_face, u1) on a pressure face and (rho1, u_face)
// on a velocity face; in shifted form h_i = h_i(n1) + w_i (rho_b P_i(u_b) -
// rho1 P_i(u1)), P_i(u) = 1 + 3 c.u + 4.5 (c.u)^2 - 1.5 u^2.  The
// populations of n1 are its post-streaming set: gathered from its upstream
// cells, bounced back where one is missing (walls with the Ladd term,
// solids plain; faces next to a NEEM face are walls or periodic, checked on
// the host).  With a solid n1 the cell keeps its bounce-back values.
template <int Q, typename ST, typename G>
__device__ __forceinline__ void neem_correct(const G &a, int f, typename Store<ST>::R (&h)[Q]);

template <int Q, typename ST, bool FLAGS, typename G>
__device__ __forceinline__ void neem_face(const G &a, int x, int y, int z, int f, typename Store<ST>::R (&h)[Q]) {
  using L = Lat<Q>;
  using C = typename Store<ST>::C;
  using R = typename Store<ST>::R;
  const ST *src = static_cast<const ST *>(a.src);
  const long long plane = a.plane;
  const int ax = f >> 1, in = (f & 1) ? -1 : 1;
  const int x1 = x + (ax == 0 ? in : 0), y1 = y + (ax == 1 ? in : 0), z1 = z + (ax == 2 ? in : 0);
  const long long n1 = (long long)x1 + (long long)a.px * ((long long)y1 + (long long)a.ny * z1);
  const uint32_t lm1 = FLAGS ? a.links[n1] : 0u;
  if (lm1 & LINK_SELF_SOLID) return;
  // unrolled so that the flag and population loads of all directions are
  // in flight together (face cells are one lane of their wave on x faces)
  const int ext[3] = {a.nx, a.ny, a.nz};
  int face_of[Q];
#pragma unroll
  for (int j = 0; j < Q; ++j) {
    const int cj[3] = {L::cx[j], L::cy[j], L::cz[j]};
    int s[3] = {x1 - cj[0], y1 - cj[1], z1 - cj[2]};
    int face = -1;
    bool ok = true;
#pragma unroll
    for (int d = 0; d < L::D; ++d) {
      if (cj[d] == 0) continue;
      if (s[d] >= 0 && s[d] < ext[d]) continue;
      if (a.per[d]) { s[d] = s[d] < 0 ? s[d] + ext[d] : s[d] - ext[d]; continue; }
      ok = false;
      const int g = 2 * d + (cj[d] > 0 ? 0 : 1);
      if (face < 0 || a.face[g].prio > a.face[face].prio) face = g;
    }
    const long long t = (long long)s[0] + (long long)a.px * ((long long)s[1] + (long long)a.ny * s[2]);
    if (ok && FLAGS && j != Q - 1) ok = !((lm1 >> j) & 1u);
    const long long addr = ok ? (long long)j * plane + t : (long long)L::opp[j] * plane + n1;
    h[j] = Store<ST>::ld(src + addr);
    face_of[j] = ok ? -2 : face;  // -2: streamed, -1: solid (plain bounce-back)
  }
#pragma unroll
  for (int j = 0; j < Q - 1; ++j) {
    const int face = face_of[j];
    if (face >= 0 && a.face[face].kind == LBM_FACE_WALL)
      h[j] = R(C(h[j]) + C(6) * C(L::w[j]) *
                             (C(L::cx[j]) * C(a.face[face].u[0]) + C(L::cy[j]) * C(a.face[face].u[1]) +
                              C(L::cz[j]) * C(a.face[face].u[2])));
  }
  neem_correct<Q, ST>(a, f, h);
}

// The NEEM rebuild of a face cell from its inward neighbour's post-streaming
// populations, already in h.
template <int Q, typename ST, typename G>
__device__ __forceinline__ void neem_correct(const G &a, int f, typename Store<ST>::R (&h)[Q]) {
  using L = Lat<Q>;
  using C = typename Store<ST>::C;
  using R = typename Store<ST>::R;
  const Moments<Q, C> m = moments<Q, C>(h, C(a.F[0]), C(a.F[1]), C(a.F[2]));
  const bool vel = a.face[f].kind == LBM_FACE_VELOCITY_NEEM;
  const C rb = vel ? m.rho : C(a.face[f].rho);
  const C bx = vel ? C(a.face[f].u[0]) : m.ux, by = vel ? C(a.face[f].u[1]) : m.uy,
          bz = vel ? C(a.face[f].u[2]) : m.uz;
  const C usq1 = m.ux * m.ux + m.uy * m.uy + m.uz * m.uz;
  const C usqb = bx * bx + by * by + bz * bz;
#pragma unroll
  for (int j = 0; j < Q; ++j) {
    const C cu1 = C(L::cx[j]) * m.ux + C(L::cy[j]) * m.uy + C(L::cz[j]) * m.uz;
    const C cub = C(L::cx[j]) * bx + C(L::cy[j]) * by + C(L::cz[j]) * bz;
    const C pb = C(1) + C(3) * cub + C(4.5) * cub * cub - C(1.5) * usqb;
    const C p1 = C(1) + C(3) * cu1 + C(4.5) * cu1 * cu1 - C(1.5) * usq1;
    h[j] = R(C(h[j]) + C(L::w[j]) * (rb * pb - m.rho * p1));
  }
}

__device__ __forceinline__ bool full_cell_kind(int k) {
  return k == LBM_FACE_VELOCITY_ZH || k == LBM_FACE_PRESSURE_ZH || k == LBM_FACE_PRESSURE_NEEM ||
         k == LBM_FACE_VELOCITY_NEEM || k == LBM_FACE_WALL_EQ || k == LBM_FACE_WALL_NEEM || k == LBM_FACE_WALL_NEBB ||
         k == LBM_FACE_WALL_NEBB0;
}
__device__ __forceinline__ bool wet_kind(int k) {
  return k == LBM_FACE_WALL_EQ || k == LBM_FACE_WALL_NEEM || k == LBM_FACE_WALL_NEBB || k == LBM_FACE_WALL_NEBB0;
}

// Post-streaming population j of the cell (x1, y1, z1) (index n1, link mask
// lm1): pulled from its upstream cell; across a wall face bounced back with
// the Ladd term; across the periodic-pressure face pg the value that face
// supplies, feq(p', u_c) + S(c, j) - feq(rho_c, u_c) with the moments
// (rc, uc) of the cell pc straight across; from a solid or across any other
// face plain bounce-back S(n1, opp j).
template <int Q, typename ST, bool FLAGS, typename G>
__device__ __forceinline__ typename Store<ST>::C post_stream_pop(const G &a, int x1, int y1, int z1, long long n1,
                                                                 uint32_t lm1, int j, int pg, long long pc,
                                                                 typename Store<ST>::C rc,
                                                                 const typename Store<ST>::C (&uc)[3]) {
  using L = Lat<Q>;
  using C = typename Store<ST>::C;
  const ST *src = static_cast<const ST *>(a.src);
  const long long plane = a.plane;
  const int ext[3] = {a.nx, a.ny, a.nz};
  const int cj[3] = {L::cx[j], L::cy[j], L::cz[j]};
  int s[3] = {x1 - cj[0], y1 - cj[1], z1 - cj[2]};
  int face = -1;
  bool ok = true;
#pragma unroll
  for (int d = 0; d < L::D; ++d) {
    if (cj[d] == 0) continue;
    if (s[d] >= 0 && s[d] < ext[d]) continue;
    if (a.per[d]) { s[d] = s[d] < 0 ? s[d] + ext[d] : s[d] - ext[d]; continue; }
    ok = false;
    const int g = 2 * d + (cj[d] > 0 ? 0 : 1);
    if (face < 0 || a.face[g].prio > a.face[face].prio) face = g;
  }
  if (ok && FLAGS && j != Q - 1 && ((lm1 >> j) & 1u)) ok = false;
  const C wj = C(L::w[j]);
  long long addr = ok ? (long long)j * plane + (long long)s[0] + (long long)a.px * ((long long)s[1] + (long long)a.ny * s[2])
                      : (long long)L::opp[j] * plane + n1;
  if (!ok && face >= 0 && face == pg) addr = (long long)j * plane + pc;
  C v = C(Store<ST>::ld(src + addr));
  if (ok) return v;
  if (face >= 0 && face == pg) {
    const C cu = C(cj[0]) * uc[0] + C(cj[1]) * uc[1] + C(cj[2]) * uc[2];
    const C usq = uc[0] * uc[0] + uc[1] * uc[1] + uc[2] * uc[2];
    return v + wj * (C(a.face[face ^ 1].rho) - rc) * (C(1) + C(3) * cu + C(4.5) * cu * cu - C(1.5) * usq);
  }
  if (face >= 0 && a.face[face].kind == LBM_FACE_WALL)
    v += C(6) * wj * (C(cj[0]) * C(a.face[face].u[0]) + C(cj[1]) * C(a.face[face].u[1]) + C(cj[2]) * C(a.face[face].u[2]));
  return v;
}

// Wet-node walls (src/lbm/bnd/bnd_wall.h:100-479) on the cell's face f.
// Density from the populations known at the cell (LBMBnd_wallWetnode,
// bnd_wetnode.h:20-55, and calcDensity_limited, moments.h:120-153): a
// population leaving through the wall counts twice, one parallel to it once
// (twice when the cell's downstream neighbour along it is missing), each only
// when it streamed in or entered across a periodic-pressure face (ppm);
// when the constants do not add up to Q-1 (a corner) the density is the
// cell's density before the step.  With a wall velocity u_w the density is
// divided by 1 + n.u_w (n the outward normal).  Then
//   equilibrium: h_i = heq_i(rho, u_w);
//   NEEM:        h_i = heq_i(rho, u_w) + h_i(n1) - heq_i(rho1, u1), n1 the
//                inward neighbour with its post-streaming populations and
//                their moments (no force term, as calcVelocity);
//   NEBB (D2Q9): the inward normal population = the outward one + s 2/3 rho
//                u_w,n, the two inward diagonals from the outward ones, the
//                tangential difference and u_w (Zou & He 1997), the other
//                populations unchanged.
// Populations in `hold` (upstream face of higher priority) keep their value.
template <int Q, typename ST, bool FLAGS, bool W64, typename G>
__device__ __forceinline__ void wet_wall(const G &a, int x, int y, int z, int f, typename Store<ST>::R (&h)[Q], unsigned missing,
                         unsigned ppm, unsigned hold, Idx<W64> cell) {
  using L = Lat<Q>;
  using C = typename Store<ST>::C;
  using R = typename Store<ST>::R;
  const ST *src = static_cast<const ST *>(a.src);
  const long long plane = a.plane;
  const int ax = f >> 1, s = (f & 1) ? 1 : -1;
  const int k = a.face[f].kind;
  auto cax = [](int j, int d) { return d == 0 ? L::cx[j] : (d == 1 ? L::cy[j] : L::cz[j]); };
  if (k == LBM_FACE_WALL_NEBB0) {
    if constexpr (Q == 9) {
      // the two diagonals entering through the wall: their initial value
      // (0 in shifted form at rho 1, u 0), or on a periodic-pressure face
      // the value that face supplies (it writes every population of its
      // cells before streaming)
      int pg = -1;
      if (ax != 0 && (x == 0 || x == a.nx - 1)) {
        const int g = x == 0 ? 0 : 1;
        if (a.face[g].kind == LBM_FACE_PERIODIC_PRESSURE) pg = g;
      }
      if (ax != 1 && (y == 0 || y == a.ny - 1)) {
        const int g = y == 0 ? 2 : 3;
        if (a.face[g].kind == LBM_FACE_PERIODIC_PRESSURE) pg = g;
      }
      C prc = 1, puc[3] = {0, 0, 0};
      long long pc = 0;
      if (pg >= 0) {
        const int ox = (pg >> 1) == 0 ? ((pg & 1) ? 0 : a.nx - 1) : x;
        const int oy = (pg >> 1) == 1 ? ((pg & 1) ? 0 : a.ny - 1) : y;
        pc = (long long)ox + (long long)a.px * (long long)oy;
        cell_rho_u<Q, ST>(src, plane, pc, a.F[0], a.F[1], a.F[2], prc, puc);
      }
      const C pusq = puc[0] * puc[0] + puc[1] * puc[1];
#pragma unroll
      for (int j = 4; j < 8; ++j) {
        if (cax(j, ax) * s >= 0 || ((hold >> j) & 1u)) continue;
        C v = 0;
        if (pg >= 0) {
          const C cu = C(L::cx[j]) * puc[0] + C(L::cy[j]) * puc[1];
          v = C(Store<ST>::ld(src + (long long)j * plane + pc)) +
              C(L::w[j]) * (C(a.face[pg ^ 1].rho) - prc) * (C(1) + C(3) * cu + C(4.5) * cu * cu - C(1.5) * pusq);
        }
        h[j] = R(v);
      }
      // every main population whose neighbour is missing reflected
#pragma unroll
      for (int d = 0; d < 4; ++d)
        if (((missing >> L::opp[d]) & 1u) && !((hold >> L::opp[d]) & 1u)) h[L::opp[d]] = h[d];
      // the diagonal formulas of the branch the outward normal selects
      C n[9];
#pragma unroll
      for (int j = 0; j < 9; ++j) n[j] = C(h[j]);
      if (ax == 0 && s < 0) {
        n[6] = C(h[4]) + C(0.5) * (C(h[3]) - C(h[2]));
        n[7] = C(h[5]) - C(0.5) * (C(h[3]) - C(h[2]));
      } else if (ax == 0) {
        n[4] = C(h[6]) - C(0.5) * (C(h[3]) - C(h[2]));
        n[5] = C(h[7]) + C(0.5) * (C(h[3]) - C(h[2]));
      } else if (s < 0) {
        n[5] = C(h[7]) - C(0.5) * (C(h[1]) - C(h[0]));
        n[6] = C(h[4]) + C(0.5) * (C(h[1]) - C(h[0]));
      } else {
        n[7] = C(h[5]) + C(0.5) * (C(h[1]) - C(h[0]));
        n[4] = C(h[6]) - C(0.5) * (C(h[1]) - C(h[0]));
      }
#pragma unroll
      for (int j = 4; j < 8; ++j)
        if (!((hold >> j) & 1u)) h[j] = R(n[j]);
    }
    return;
  }
  C dr = 0, ws = 0;
  int cnt = 0;
#pragma unroll
  for (int j = 0; j < Q - 1; ++j) {
    const int cn = cax(j, ax) * s;
    if (cn < 0) continue;
    if (((missing >> j) & 1u) && !((ppm >> j) & 1u)) continue;
    int kk = 2;
    if (cn == 0) {
      const int o = L::opp[j];
      kk = (((missing >> o) & 1u) && !((ppm >> o) & 1u)) ? 2 : 1;
    }
    dr += C(kk) * C(h[j]);
    ws += C(kk) * C(L::w[j]);
    cnt += kk;
  }
  C drho;
  if (cnt == Q - 1) {
    drho = dr + C(h[Q - 1]) + (ws + C(L::w[Q - 1]) - C(1));
  } else {
    drho = 0;
#pragma unroll
    for (int j = 0; j < Q; ++j) drho += C(Store<ST>::ld(pop_ptr<ST, W64>(src, plane, j, cell)));
  }
  const C uw[3] = {C(a.face[f].u[0]), C(a.face[f].u[1]), C(a.face[f].u[2])};
  const C uwn = ax == 0 ? uw[0] : (ax == 1 ? uw[1] : uw[2]);  // no run-time array index
  C rho = C(1) + drho;
  const C un = C(s) * uwn;
  if (un != C(0)) {
    rho = rho * (C(1) / (C(1) + un));
    drho = rho - C(1);
  }
  const C usq = uw[0] * uw[0] + uw[1] * uw[1] + uw[2] * uw[2];
  if (k == LBM_FACE_WALL_NEBB) {
    if constexpr (Q == 9) {
      const int out = 2 * ax + (s > 0 ? 1 : 0), in = L::opp[out];
      if (!((hold >> in) & 1u)) h[in] = R(C(h[out]) + C(s) * C(2.0 / 3.0) * rho * uwn);
      const C sixth = C(1.0 / 6.0);
      C n[9];
#pragma unroll
      for (int j = 0; j < 9; ++j) n[j] = C(h[j]);
      if (ax == 0 && s > 0) {
        n[6] = C(h[4]) + C(0.5) * (C(h[3]) - C(h[2])) - C(0.5) * rho * uw[1] - sixth * rho * uw[0];
        n[7] = C(h[5]) - C(0.5) * (C(h[3]) - C(h[2])) + C(0.5) * rho * uw[1] - sixth * rho * uw[0];
      } else if (ax == 0) {
        n[4] = C(h[6]) - C(0.5) * (C(h[3]) - C(h[2])) + C(0.5) * rho * uw[1] + sixth * rho * uw[0];
        n[5] = C(h[7]) + C(0.5) * (C(h[3]) - C(h[2])) - C(0.5) * rho * uw[1] + sixth * rho * uw[0];
      } else if (s > 0) {
        n[5] = C(h[7]) - C(0.5) * (C(h[1]) - C(h[0])) + C(0.5) * rho * uw[0] - sixth * rho * uw[1];
        n[6] = C(h[4]) + C(0.5) * (C(h[1]) - C(h[0])) - C(0.5) * rho * uw[0] - sixth * rho * uw[1];
      } else {
        n[7] = C(h[5]) + C(0.5) * (C(h[1]) - C(h[0])) - C(0.5) * rho * uw[0] + sixth * rho * uw[1];
        n[4] = C(h[6]) - C(0.5) * (C(h[1]) - C(h[0])) + C(0.5) * rho * uw[0] + sixth * rho * uw[1];
      }
#pragma unroll
      for (int j = 4; j < 8; ++j)
        if (!((hold >> j) & 1u)) h[j] = R(n[j]);
    }
    return;
  }
  // NEEM: moments of the inward neighbour's post-streaming populations
  bool neem = k == LBM_FACE_WALL_NEEM;
  int x1 = x, y1 = y, z1 = z;
  if (ax == 0) x1 -= s;
  if (ax == 1) y1 -= s;
  if (ax == 2) z1 -= s;
  const long long n1 = (long long)x1 + (long long)a.px * ((long long)y1 + (long long)a.ny * z1);
  const uint32_t lm1 = (FLAGS && neem) ? a.links[n1] : 0u;
  if (lm1 & LINK_SELF_SOLID) neem = false;
  // the periodic-pressure face n1 lies on, if any, and the moments of the
  // cell straight across from n1 on it
  int pg = -1;
  long long pc = 0;
  C rc = 1, uc[3] = {0, 0, 0};
  if (neem) {
    const int c1[3] = {x1, y1, z1}, ext[3] = {a.nx, a.ny, a.nz};
    int cx = x1, cy = y1, cz = z1;
#pragma unroll
    for (int g = 0; g < 2 * L::D; ++g) {
      const int ag = g >> 1;
      if (pg >= 0 || ag == ax || a.face[g].kind != LBM_FACE_PERIODIC_PRESSURE) continue;
      if (c1[ag] != ((g & 1) ? ext[ag] - 1 : 0)) continue;
      const int o = (g & 1) ? 0 : ext[ag] - 1;
      if (ag == 0) cx = o;
      if (ag == 1) cy = o;
      if (ag == 2) cz = o;
      pg = g;
    }
    if (pg >= 0) {
      pc = (long long)cx + (long long)a.px * ((long long)cy + (long long)a.ny * cz);
      cell_rho_u<Q, ST>(src, plane, pc, a.F[0], a.F[1], a.F[2], rc, uc);
    }
  }
  C d1 = 0, j1[3] = {0, 0, 0};
  if (neem) {
#pragma unroll
    for (int j = 0; j < Q; ++j) {
      const C v = post_stream_pop<Q, ST, FLAGS>(a, x1, y1, z1, n1, lm1, j, pg, pc, rc, uc);
      d1 += v;
      j1[0] += C(L::cx[j]) * v;
      j1[1] += C(L::cy[j]) * v;
      j1[2] += C(L::cz[j]) * v;
    }
  }
  const C rho1 = C(1) + d1;
  const C u1[3] = {j1[0] / rho1, j1[1] / rho1, j1[2] / rho1};
  const C usq1 = u1[0] * u1[0] + u1[1] * u1[1] + u1[2] * u1[2];
#pragma unroll
  for (int j = 0; j < Q; ++j) {
    if ((hold >> j) & 1u) continue;
    const C wj = C(L::w[j]);
    const C cu = C(L::cx[j]) * uw[0] + C(L::cy[j]) * uw[1] + C(L::cz[j]) * uw[2];
    C v = wj * (drho + rho * (C(3) * cu + C(4.5) * cu * cu - C(1.5) * usq));
    if (neem) {
      const C c1 = C(L::cx[j]) * u1[0] + C(L::cy[j]) * u1[1] + C(L::cz[j]) * u1[2];
      v += post_stream_pop<Q, ST, FLAGS>(a, x1, y1, z1, n1, lm1, j, pg, pc, rc, uc) -
           wj * (d1 + rho1 * (C(3) * c1 + C(4.5) * c1 * c1 - C(1.5) * usq1));
    }
    h[j] = R(v);
  }
}

// Index of population j among the populations with c_x != 0 (the ones an
// x-row tile pulls from a shifted position), in population order; -1 for
// c_x == 0.  NXS: how many there are (6 for D2Q9, 10 for D3Q19).
template <int Q>
__host__ __device__ constexpr int xshift_index(int j) {
  int k = 0;
  for (int i = 0; i < j; ++i) k += Lat<Q>::cx[i] != 0;
  return Lat<Q>::cx[j] != 0 ? k : -1;
}
template <int Q>
__host__ __device__ constexpr int count_xshift() {
  int k = 0;
  for (int i = 0; i < Q; ++i) k += Lat<Q>::cx[i] != 0;
  return k;
}
template <int Q>
constexpr int NXS = count_xshift<Q>();
static_assert(NXS<9> == 6 && NXS<19> == 10, "x-shifted populations per lattice");

// Row width of the LDS x-halo stage: a 256-cell row plus one cell each side.
constexpr int LDS_ROW = 256 + 2;

// --------------------------------------------------------------------------
// The fused step.  MODE 0: plain; 1: also write (u, rho) snapshot;
// 2: reduce against the snapshot written by the previous step; 3: both
// (reduce against snap_in, write snap_out).
// Per-cell update shared by the single-domain and batched kernels.  G is a
// view of one structured domain with members src, dst, links, plane, nx, ny,
// nz, px, per[3], face[6], omega, F[3], snap_in, snap_out (the kernel
// arguments for one domain, or a domain of a batch).  red must be zeroed by
// the caller; MODE >= 2 accumulates this cell's sums into it.
// TOPO 0: any domain.  The boundary paths are compiled out when the caller
// guarantees that no population of the cell crosses a face or comes from a
// solid: TOPO 1, every neighbour is a fluid cell inside the domain (no
// wrap); TOPO 2, every axis is periodic and there are no solids (wraps
// only).
// LDSX: the populations with c_x != 0 come from the block's LDS stage xs
// (k_step_lds: row of plane j at the source (y, z), cells x0 - 1 .. x0 + 256
// in slots 0 .. 257) instead of shifted global loads, and those with
// c_x == 0 from pre[j], loaded by the caller.
template <int Q, typename ST, bool FLAGS, int MODE, bool NT, bool W64, int TOPO = 0, bool LDSX = false, typename G>
__device__ __forceinline__ void cell_update(const G &a, int x, int y, int z, double (&red)[NSUM],
                                            const typename Store<ST>::R *xs = nullptr,
                                            const typename Store<ST>::R *pre = nullptr) {
  using L = Lat<Q>;
  using C = typename Store<ST>::C;
  using I = Idx<W64>;
  const ST *__restrict__ src = static_cast<const ST *>(a.src);
  ST *__restrict__ dst = static_cast<ST *>(a.dst);
  const long long plane = a.plane;
  const I cell = I(x) + I(a.px) * (I(y) + I(a.ny) * I(z));
  bool active = x < a.nx && y < a.ny;
  const uint32_t lm = (FLAGS && active) ? a.links[cell] : 0u;
  if (lm & LINK_SELF_SOLID) active = false;

  using R = typename Store<ST>::R;
  R h[Q];
  unsigned missing = 0;
  // ff: the face whose rule replaces the whole cell (Zou-He, NEEM, wet-node
  // walls), the one of highest priority the cell lies on, ties to the lower
  // face index; hold: populations whose upstream face outranks a wet-node
  // wall ff; ppm: populations entering across a periodic-pressure face
  int ff = -1;
  unsigned hold = 0, ppm = 0;
  if (active) {
    // neighbour coordinates per axis with periodic wrap; v* = exists
    int xm = x - 1, xp = x + 1, ym = y - 1, yp = y + 1, zm = z - 1, zp = z + 1;
    bool vxm = true, vxp = true, vym = true, vyp = true, vzm = true, vzp = true;
    if (TOPO == 2) {
      if (x == 0) xm = a.nx - 1;
      if (x == a.nx - 1) xp = 0;
      if (y == 0) ym = a.ny - 1;
      if (y == a.ny - 1) yp = 0;
      if (Q == 19) {
        if (z == 0) zm = a.nz - 1;
        if (z == a.nz - 1) zp = 0;
      }
    } else if (TOPO == 0) {
      if (x == 0) { if (a.per[0]) xm = a.nx - 1; else vxm = false; }
      if (x == a.nx - 1) { if (a.per[0]) xp = 0; else vxp = false; }
      if (y == 0) { if (a.per[1]) ym = a.ny - 1; else vym = false; }
      if (y == a.ny - 1) { if (a.per[1]) yp = 0; else vyp = false; }
      if (Q == 19) {
        if (z == 0) { if (a.per[2]) zm = a.nz - 1; else vzm = false; }
        if (z == a.nz - 1) { if (a.per[2]) zp = 0; else vzp = false; }
      }
    }
#pragma unroll
    for (int j = 0; j < Q; ++j) {
      const int cx = L::cx[j], cy = L::cy[j], cz = L::cz[j];
      const int sx = cx > 0 ? xm : (cx < 0 ? xp : x);
      const int sy = cy > 0 ? ym : (cy < 0 ? yp : y);
      const int sz = cz > 0 ? zm : (cz < 0 ? zp : z);
      bool valid = (cx > 0 ? vxm : (cx < 0 ? vxp : true)) && (cy > 0 ? vym : (cy < 0 ? vyp : true)) &&
                   (cz > 0 ? vzm : (cz < 0 ? vzp : true));
      const I s = I(sx) + I(a.px) * (I(sy) + I(a.ny) * I(sz));
      if (LDSX && cx != 0) {
        // staged: slot (x - x0) + 1 - c_x of the row of plane j
        h[j] = xs[xshift_index<Q>(j) * LDS_ROW + (int)threadIdx.x + 1 - cx];
      } else if (LDSX) {
        // c_x == 0: loaded by the caller before the stage's barrier (the
        // same source, or the cell's own entry when the source is missing)
        h[j] = pre[j];
      } else {
        // a population whose upstream cell is missing (across a face, or a
        // solid by the link bits) loads its bounce-back source S(x, opp j)
        // at once, so the boundary rules below start from it without a
        // second round of loads (the boundary waves of the 512x128x128
        // sphere channel waited for that round)
        const bool miss = !valid || (FLAGS && j < Q - 1 && ((lm >> j) & 1u));
        const ST *p = pop_ptr<ST, W64>(src, plane, miss ? L::opp[j] : j, miss ? cell : s);
        h[j] = NT ? Store<ST>::ld_nt(p) : Store<ST>::ld(p);
      }
      if (!valid) missing |= 1u << j;
    }
    // A population streams iff its upstream cell is fluid.  This is the
    // symmetric rule; the reference composes diagonal neighbours from main
    // ones (src/cartesiangrid.h:466-486), which next to staircase obstacles
    // makes the neighbour relation one-sided and leaves some populations
    // unwritten for a step.  Populations pulled from a solid cell (link
    // bits) are replaced below like the ones across a face.
    if (FLAGS) missing |= lm & ((1u << (Q - 1)) - 1u);

    if (missing) {
      // bit f: the cell lies on face f
      const unsigned on = (vxm ? 0u : 1u) | (vxp ? 0u : 2u) | (vym ? 0u : 4u) | (vyp ? 0u : 8u) | (vzm ? 0u : 16u) |
                          (vzp ? 0u : 32u);
      // (kind and priority of ff tracked here: the face table is indexed
      // with compile-time indices only)
      int ffprio = 0, ffkind = -1;
#pragma unroll
      for (int f = 0; f < 2 * L::D; ++f)
        if (((on >> f) & 1u) && full_cell_kind(a.face[f].kind) && (ff < 0 || a.face[f].prio > ffprio)) {
          ff = f;
          ffprio = a.face[f].prio;
          ffkind = a.face[f].kind;
        }
      const bool wet = wet_kind(ffkind);
      // bounce-back source S(x, opp j) of every missing population (the
      // gather loaded it already, except from the LDS stage)
      if (LDSX)
#pragma unroll
        for (int j = 0; j < Q - 1; ++j)
          if ((missing >> j) & 1u) h[j] = Store<ST>::ld(pop_ptr<ST, W64>(src, plane, L::opp[j], cell));
      // boundary substitution for populations whose upstream cell is missing.
      // First pass: the face each one crosses (the higher priority wins,
      // ties to the lower axis), walls applied at once; anti-bounce-back
      // and periodic-pressure faces read other cells and go in passes of
      // their own, so that their state is live only there.
      unsigned abbm = 0;
#pragma unroll
      for (int j = 0; j < Q - 1; ++j) {
        if (!((missing >> j) & 1u)) continue;
        const int cx = L::cx[j], cy = L::cy[j], cz = L::cz[j];
        const bool ix = (cx > 0 && !vxm) || (cx < 0 && !vxp);
        const bool iy = (cy > 0 && !vym) || (cy < 0 && !vyp);
        const bool iz = (cz > 0 && !vzm) || (cz < 0 && !vzp);
        int f = -1;
        if (ix) f = cx > 0 ? 0 : 1;
        if (iy) { const int g = cy > 0 ? 2 : 3; if (f < 0 || a.face[g].prio > a.face[f].prio) f = g; }
        if (iz) { const int g = cz > 0 ? 4 : 5; if (f < 0 || a.face[g].prio > a.face[f].prio) f = g; }
        if (f < 0) continue;  // solid obstacle: plain bounce-back, h[j] already holds S(x, opp j)
        const int fk = a.face[f].kind;
        if (wet && a.face[f].prio > ffprio) hold |= 1u << j;
        if (fk == LBM_FACE_WALL) {
          // Ladd moving wall, rho0 = 1 (bnd_dirichlet.h:109-116)
          const C cu = C(cx) * C(a.face[f].u[0]) + C(cy) * C(a.face[f].u[1]) + C(cz) * C(a.face[f].u[2]);
          h[j] = R(C(h[j]) + C(6) * C(L::w[j]) * cu);
        } else if (fk == LBM_FACE_PRESSURE_ABB) {
          abbm |= 1u << j;
        } else if (fk == LBM_FACE_PERIODIC_PRESSURE) {
          ppm |= 1u << j;
        }
        // Zou-He, NEEM and wet-node faces: whole-cell rule below
      }
      // anti-bounce-back, per face the cell lies on: u extrapolated from the
      // two inward cells (bnd_pressure.h:70-81), f = -f_post + 2 symmEq(p, u)
      if (abbm) {
#pragma unroll 1
        for (int f = 0; f < 2 * L::D; ++f) {
          if (!((on >> f) & 1u) || a.face[f].kind != LBM_FACE_PRESSURE_ABB) continue;
          const int ax = f >> 1, out = (f & 1) ? 1 : -1;
          const int dx = ax == 0 ? -out : 0, dy = ax == 1 ? -out : 0, dz = ax == 2 ? -out : 0;
          const long long n1 = (long long)(x + dx) + (long long)a.px * ((long long)(y + dy) + (long long)a.ny * (z + dz));
          const long long n2 = (long long)(x + 2 * dx) + (long long)a.px * ((long long)(y + 2 * dy) + (long long)a.ny * (z + 2 * dz));
          C u1[3], u2[3], uext[3];
          cell_velocity<Q, ST>(src, plane, n1, a.F[0], a.F[1], a.F[2], u1);
          cell_velocity<Q, ST>(src, plane, n2, a.F[0], a.F[1], a.F[2], u2);
          for (int d = 0; d < 3; ++d) uext[d] = C(1.5) * u1[d] - C(0.5) * u2[d];
          const C vsq = uext[0] * uext[0] + uext[1] * uext[1] + uext[2] * uext[2];
          const C p2 = C(2) * C(a.face[f].rho);
#pragma unroll
          for (int j = 0; j < Q - 1; ++j) {
            if (!((abbm >> j) & 1u)) continue;
            if (!(ax == 0 ? L::cx[j] == -out : (ax == 1 ? L::cy[j] == -out : L::cz[j] == -out))) continue;
            // owned by this face unless a face of higher priority won it
            const int cx = L::cx[j], cy = L::cy[j], cz = L::cz[j];
            const bool ix = (cx > 0 && !vxm) || (cx < 0 && !vxp);
            const bool iy = (cy > 0 && !vym) || (cy < 0 && !vyp);
            const bool iz = (cz > 0 && !vzm) || (cz < 0 && !vzp);
            int g = -1;
            if (ix) g = cx > 0 ? 0 : 1;
            if (iy) { const int t = cy > 0 ? 2 : 3; if (g < 0 || a.face[t].prio > a.face[g].prio) g = t; }
            if (iz) { const int t = cz > 0 ? 4 : 5; if (g < 0 || a.face[t].prio > a.face[g].prio) g = t; }
            if (g != f) continue;
            const C cu = C(cx) * uext[0] + C(cy) * uext[1] + C(cz) * uext[2];
            h[j] = R(-C(h[j]) + C(L::w[j]) * (p2 * (C(1) + C(4.5) * cu * cu - C(1.5) * vsq) - C(2)));
          }
        }
      }
      // periodic pressure, per face: the cell straight across on the opposite
      // face supplies feq(p', u_c) + S(c, j) - feq(rho_c, u_c)
      // (bnd_periodic.h:102-108)
      if (ppm) {
#pragma unroll 1
        for (int f = 0; f < 2 * L::D; ++f) {
          if (!((on >> f) & 1u) || a.face[f].kind != LBM_FACE_PERIODIC_PRESSURE) continue;
          const int ax = f >> 1, out = (f & 1) ? 1 : -1;
          const int ox = ax == 0 ? ((f & 1) ? 0 : a.nx - 1) : x;
          const int oy = ax == 1 ? ((f & 1) ? 0 : a.ny - 1) : y;
          const int oz = ax == 2 ? ((f & 1) ? 0 : a.nz - 1) : z;
          const long long pc = (long long)ox + (long long)a.px * ((long long)oy + (long long)a.ny * oz);
          C prho, pu[3];
          cell_rho_u<Q, ST>(src, plane, pc, a.F[0], a.F[1], a.F[2], prho, pu);
          const C usq = pu[0] * pu[0] + pu[1] * pu[1] + pu[2] * pu[2];
          const C dp = C(a.face[f ^ 1].rho) - prho;
#pragma unroll
          for (int j = 0; j < Q - 1; ++j) {
            if (!((ppm >> j) & 1u)) continue;
            if (!(ax == 0 ? L::cx[j] == -out : (ax == 1 ? L::cy[j] == -out : L::cz[j] == -out))) continue;
            const int cx = L::cx[j], cy = L::cy[j], cz = L::cz[j];
            const bool ix = (cx > 0 && !vxm) || (cx < 0 && !vxp);
            const bool iy = (cy > 0 && !vym) || (cy < 0 && !vyp);
            const bool iz = (cz > 0 && !vzm) || (cz < 0 && !vzp);
            int g = -1;
            if (ix) g = cx > 0 ? 0 : 1;
            if (iy) { const int t = cy > 0 ? 2 : 3; if (g < 0 || a.face[t].prio > a.face[g].prio) g = t; }
            if (iz) { const int t = cz > 0 ? 4 : 5; if (g < 0 || a.face[t].prio > a.face[g].prio) g = t; }
            if (g != f) continue;
            const C cu = C(cx) * pu[0] + C(cy) * pu[1] + C(cz) * pu[2];
            const C sj = C(Store<ST>::ld(src + (long long)j * plane + pc));
            h[j] = R(sj + C(L::w[j]) * dp * (C(1) + C(3) * cu + C(4.5) * cu * cu - C(1.5) * usq));
          }
        }
      }
    }
  }

  // NEEM faces across x: the inward neighbour of a face cell is the adjacent
  // lane of the same wave (a wave holds contiguous x of one row), so its
  // post-streaming populations, complete after its own bounce-back, move
  // across lanes instead of being gathered again.  The source lane is
  // checked by its cell index; where it is not in the wave the face cell
  // gathers (neem_face).
  bool from_lane = false;
  if (TOPO == 0 && (a.face[0].kind == LBM_FACE_PRESSURE_NEEM || a.face[0].kind == LBM_FACE_VELOCITY_NEEM ||
                    a.face[1].kind == LBM_FACE_PRESSURE_NEEM || a.face[1].kind == LBM_FACE_VELOCITY_NEEM)) {
    const bool lo = x == 0 && (a.face[0].kind == LBM_FACE_PRESSURE_NEEM || a.face[0].kind == LBM_FACE_VELOCITY_NEEM);
    const bool hi = x == a.nx - 1 &&
                    (a.face[1].kind == LBM_FACE_PRESSURE_NEEM || a.face[1].kind == LBM_FACE_VELOCITY_NEEM);
    const bool face_lane = active && (lo || hi);
    if (__ballot(face_lane)) {  // wave-uniform
      const int lane = (int)__lane_id();
      const int dir = lo ? 1 : -1;
      const int sl = face_lane ? min(63, max(0, lane + dir)) : lane;
      const unsigned key = active ? (unsigned)cell : 0xffffffffu;
      const unsigned got = __shfl(key, sl, 64);
      from_lane = face_lane && sl != lane && got == (unsigned)(dir > 0 ? cell + I(1) : cell - I(1));
#pragma unroll
      for (int j = 0; j < Q; ++j) {
        const R t = __shfl(h[j], sl, 64);
        if (from_lane) h[j] = t;
      }
    }
  }

  if (active) {
    const C Fx = C(a.F[0]), Fy = C(a.F[1]), Fz = C(a.F[2]);
    const bool force = a.F[0] != 0.0 || a.F[1] != 0.0 || a.F[2] != 0.0;
    if (ff >= 0) {
      // the whole-cell rule of face ff
      const int f = ff, k = a.face[f].kind, ax = f >> 1;
      if (k == LBM_FACE_PRESSURE_NEEM || k == LBM_FACE_VELOCITY_NEEM) {
        if (from_lane && ax == 0) neem_correct<Q, ST>(a, f, h);
        else neem_face<Q, ST, FLAGS>(a, x, y, z, f, h);
      } else if (k == LBM_FACE_VELOCITY_ZH || k == LBM_FACE_PRESSURE_ZH) {
        zou_he<Q, R>(h, ax, (f & 1) ? 1 : -1, k == LBM_FACE_VELOCITY_ZH, R(a.face[f].u[0]), R(a.face[f].u[1]),
                     R(a.face[f].u[2]), R(a.face[f].rho));
      } else {
        wet_wall<Q, ST, FLAGS, W64>(a, x, y, z, f, h, missing, ppm, hold, cell);
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
    collide_store<Q, ST, NT, W64>(h, m, C(a.omega), Fx, Fy, Fz, force, dst, plane, cell);
  }
}

}  // namespace lbm
