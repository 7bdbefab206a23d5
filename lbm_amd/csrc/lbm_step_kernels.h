// Step kernels (single domain and batched) and their launchers, included
// by the per-storage translation units lbm_step_<storage>.hip.
#pragma once
#include "lbm_launch.h"

namespace lbm {

// Occupancy target of the step kernel (waves per SIMD).  D3Q19 with FP32
// storage (FP64 arithmetic) compiles to 129-131 VGPRs unconstrained, i.e. 3
// waves; held to 128 it keeps 4 waves without spilling (32-bit offsets) and
// runs 3-4 % faster.  FP64 storage spills under that limit and is left free.
template <int Q, typename ST, int TOPO>
struct StepWaves {
  static constexpr int value = 1;
};
template <>
struct StepWaves<19, float, 0> {
  static constexpr int value = 4;  // 5 waves: 96 VGPRs + 128 B of spill, 0.455 against 0.441 ms at 256^3
};
template <>
struct StepWaves<19, float, 3> {
  static constexpr int value = 4;
};
// 16-bit storage (FP32 arithmetic), general kernel: 104 VGPRs unconstrained
// (4 waves); held to 6 waves it spills 92-96 bytes per lane in its boundary
// paths and still runs the 512x128x128 sphere channel 21-26 % faster (FP16
// 273 -> 216 us, BF16 297 -> 240 us per step, identical results); 5 waves
// gave 234 us (FP16); at 8 the register allocator gives up.
template <>
struct StepWaves<19, __half, 0> {
  static constexpr int value = 6;
};
template <>
struct StepWaves<19, __hip_bfloat16, 0> {
  static constexpr int value = 6;
};
template <>
struct StepWaves<19, __half, 3> {
  static constexpr int value = 6;
};
template <>
struct StepWaves<19, __hip_bfloat16, 3> {
  static constexpr int value = 6;
};


// TOPO 3 (domains with faces or solids, a.lanes set): a path per wave -- a
// wave whose cells are all interior (every bit of a.lanes[wave]: no face,
// wrap or solid within one cell) runs without the boundary code and the
// flag loads (TOPO 1), the others run the general path (TOPO 0).  Splitting
// the same classes into two launches was measured slower on the sphere
// channel (the boundary launch's blocks hold one working wave in four, and
// the halo lines the two classes share are fetched twice).
template <int Q, typename ST, bool FLAGS, int MODE, bool NT, bool W64, int TOPO>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(StepWaves<Q, ST, TOPO>::value))) k_step(const KArgs a) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  const int z = blockIdx.z;
  double red[NSUM];
#pragma unroll
  for (int k = 0; k < NSUM; ++k) red[k] = 0;
  if constexpr (TOPO == 3) {
    const int t = (int)(threadIdx.y * blockDim.x + threadIdx.x);
    const long long wave =
        ((long long)(blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x) * (blockDim.x * blockDim.y / 64) +
        (t >> 6);
    if (a.lanes[wave] == ~0ull) cell_update<Q, ST, false, MODE, NT, W64, 1>(a, x, y, z, red);
    else cell_update<Q, ST, FLAGS, MODE, NT, W64, 0>(a, x, y, z, red);
  } else {
    cell_update<Q, ST, FLAGS, MODE, NT, W64, TOPO>(a, x, y, z, red);
  }
  if (MODE >= 2) {
    const long long nb = (long long)gridDim.x * gridDim.y * gridDim.z;
    const long long b = blockIdx.x + (long long)gridDim.x * (blockIdx.y + (long long)gridDim.y * blockIdx.z);
    block_reduce_store<256>(red, a.partials, b, nb);
  }
}

// LDS-staged x halo.  The block is one 256-cell x row (blockDim 256 x 1).
// For every population with c_x != 0 the block first loads the source row of
// plane j into LDS -- cells x0 .. x0 + 255 one aligned element per lane, and
// the two halo cells x0 - 1 and x0 + 256 (wrapped on a periodic x axis) --
// then each cell reads its shifted value there.  Every population element
// is still read once from memory; the ten (D3Q19) shifted loads per cell,
// each touching three cache lines per wave, become aligned loads of two
// lines plus two single-element halo loads per row and plane.
template <int Q, typename ST, bool FLAGS, int MODE, bool NT, bool W64, int TOPO>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(StepWaves<Q, ST, TOPO>::value)))
k_step_lds(const KArgs a) {
  using L = Lat<Q>;
  using R = typename Store<ST>::R;
  using I = Idx<W64>;
  __shared__ R xs[NXS<Q> * LDS_ROW];
  const int t = threadIdx.x;
  const int x0 = blockIdx.x * 256, x = x0 + t;
  const int y = blockIdx.y, z = blockIdx.z;
  const ST *src = static_cast<const ST *>(a.src);
  const long long plane = a.plane;
  const int last = min(x0 + 256, a.nx) - 1;  // last cell of this row
  const bool perx = TOPO == 2 || a.per[0];
  // Source row (y - c_y, z - c_z) of every staged population; a row across
  // a non-periodic face is not staged (the populations it would give are
  // replaced by the boundary rules) -- its lanes load their own row instead,
  // so that every load is issued, unconditionally and all before any LDS
  // store, as in the direct kernel.
  I rb[NXS<Q>];
  bool ok[NXS<Q>];
#pragma unroll
  for (int j = 0; j < Q; ++j) {
    const int cx = L::cx[j], cy = L::cy[j], cz = L::cz[j];
    if (cx == 0) continue;
    const int k = xshift_index<Q>(j);
    int sy = y - cy, sz = z - cz;
    ok[k] = true;
    if (sy < 0 || sy >= a.ny) {
      if (TOPO == 2 || a.per[1]) sy = sy < 0 ? sy + a.ny : sy - a.ny;
      else { ok[k] = false; sy = y; }
    }
    if (sz < 0 || sz >= a.nz) {
      if (TOPO == 2 || a.per[2]) sz = sz < 0 ? sz + a.nz : sz - a.nz;
      else { ok[k] = false; sz = z; }
    }
    rb[k] = I(a.px) * (I(sy) + I(a.ny) * I(sz));
  }
  // the row itself: lane t holds cell x0 + t (lanes past the last cell load
  // the last cell; their slots are never read)
  const int xc = min(x, last);
  R v[NXS<Q>];
#pragma unroll
  for (int j = 0; j < Q; ++j) {
    if (L::cx[j] == 0) continue;
    const int k = xshift_index<Q>(j);
    const ST *p = pop_ptr<ST, W64>(src, plane, j, rb[k] + I(xc));
    v[k] = NT ? Store<ST>::ld_nt(p) : Store<ST>::ld(p);
  }
  // the halo cells: lane k < NXS of the last wave loads the left halo of
  // staged population k, lane 32 + k the right one (wrapped on a periodic
  // x axis, else absent)
  R hv = 0;
  int hslot = -1;
  const int hl = t - (256 - 64), hk = hl & 31;
  if (hl >= 0 && hk < NXS<Q>) {
    int xh = hl < 32 ? x0 - 1 : last + 1;
    if (xh < 0) xh = perx ? a.nx - 1 : -1;
    if (xh >= a.nx) xh = perx ? 0 : -1;
    if (xh >= 0) {
      int jj = 0;  // population of staged index hk
#pragma unroll
      for (int j = 0; j < Q; ++j)
        if (xshift_index<Q>(j) == hk) jj = j;
      I r = 0;
#pragma unroll
      for (int k = 0; k < NXS<Q>; ++k)
        if (k == hk) r = rb[k];
      hv = Store<ST>::ld(pop_ptr<ST, W64>(src, plane, jj, r + I(xh)));
      hslot = hk * LDS_ROW + (hl < 32 ? 0 : last - x0 + 2);
    }
  }
  // the populations with c_x == 0: straight loads from the (y - c_y, z - c_z)
  // row, issued before the barrier with the stage's loads so that all Q
  // loads of a cell are in flight together; a source across a non-periodic
  // face loads the cell's own entry (replaced by the boundary rules)
  R pre[Q];
  if (x <= last) {
#pragma unroll
    for (int j = 0; j < Q; ++j) {
      const int cx = L::cx[j], cy = L::cy[j], cz = L::cz[j];
      if (cx != 0) continue;
      int sy = y - cy, sz = z - cz;
      bool valid = true;
      if (sy < 0 || sy >= a.ny) {
        if (TOPO == 2 || a.per[1]) sy = sy < 0 ? sy + a.ny : sy - a.ny;
        else valid = false;
      }
      if (sz < 0 || sz >= a.nz) {
        if (TOPO == 2 || a.per[2]) sz = sz < 0 ? sz + a.nz : sz - a.nz;
        else valid = false;
      }
      const I src_cell = valid ? I(x) + I(a.px) * (I(sy) + I(a.ny) * I(sz)) : I(x) + I(a.px) * (I(y) + I(a.ny) * I(z));
      const ST *p = pop_ptr<ST, W64>(src, plane, j, src_cell);
      pre[j] = NT ? Store<ST>::ld_nt(p) : Store<ST>::ld(p);
    }
  }
#pragma unroll
  for (int k = 0; k < NXS<Q>; ++k)
    if (ok[k]) xs[k * LDS_ROW + t + 1] = v[k];
  if (hslot >= 0) xs[hslot] = hv;
  __syncthreads();
  double red[NSUM];
#pragma unroll
  for (int k = 0; k < NSUM; ++k) red[k] = 0;
  cell_update<Q, ST, FLAGS, MODE, NT, W64, TOPO, true>(a, x, y, z, red, xs, pre);
  if (MODE >= 2) {
    const long long nb = (long long)gridDim.x * gridDim.y * gridDim.z;
    const long long b = blockIdx.x + (long long)gridDim.x * (blockIdx.y + (long long)gridDim.y * blockIdx.z);
    block_reduce_store<256>(red, a.partials, b, nb);
  }
}

template <int Q, typename ST, bool FLAGS, bool W64, int TOPO>
static void launch_step_lds(const KArgs &a, dim3 g, hipStream_t s, int mode, bool nt) {
  const dim3 b(256, 1, 1);
  if (mode == 0) {
    if (nt) hipLaunchKernelGGL((k_step_lds<Q, ST, FLAGS, 0, true, W64, TOPO>), g, b, 0, s, a);
    else hipLaunchKernelGGL((k_step_lds<Q, ST, FLAGS, 0, false, W64, TOPO>), g, b, 0, s, a);
  } else if (mode == 1) hipLaunchKernelGGL((k_step_lds<Q, ST, FLAGS, 1, false, W64, TOPO>), g, b, 0, s, a);
  else if (mode == 2) hipLaunchKernelGGL((k_step_lds<Q, ST, FLAGS, 2, false, W64, TOPO>), g, b, 0, s, a);
  else hipLaunchKernelGGL((k_step_lds<Q, ST, FLAGS, 3, false, W64, TOPO>), g, b, 0, s, a);
}

template <int Q, typename ST, bool FLAGS, bool W64, int TOPO>
static void launch_step_kernel(const KArgs &a, dim3 g, dim3 b, hipStream_t s, int mode, bool nt) {
  if (mode == 0) {
    if (nt) hipLaunchKernelGGL((k_step<Q, ST, FLAGS, 0, true, W64, TOPO>), g, b, 0, s, a);
    else hipLaunchKernelGGL((k_step<Q, ST, FLAGS, 0, false, W64, TOPO>), g, b, 0, s, a);
  } else if (mode == 1) hipLaunchKernelGGL((k_step<Q, ST, FLAGS, 1, false, W64, TOPO>), g, b, 0, s, a);
  else if (mode == 2) hipLaunchKernelGGL((k_step<Q, ST, FLAGS, 2, false, W64, TOPO>), g, b, 0, s, a);
  else hipLaunchKernelGGL((k_step<Q, ST, FLAGS, 3, false, W64, TOPO>), g, b, 0, s, a);
}
// The kernel without boundary paths (TOPO 2) is used for 16-bit storage
// only: measured on the 256^3 D3Q19 periodic box it is 9.7 % faster for
// FP16 (71.3 vs 65.0 GLUPS) and 3.6 % slower for FP32 (37.3 vs 38.7).
template <int Q, typename ST>
void StepL<Q, ST>::run(const KArgs &a, dim3 g, dim3 b, hipStream_t s, bool flags, int mode, bool nt, bool w64,
                       bool periodic_box, bool lds) {
  if constexpr (std::is_same<ST, float>::value) {
    // LDS x halo: FP32 storage, 32-bit offsets, 256-cell rows (checked by the host)
    if (lds && !w64) {
      if (flags) launch_step_lds<Q, ST, true, false, 0>(a, g, s, mode, nt);
      else launch_step_lds<Q, ST, false, false, 0>(a, g, s, mode, nt);
      return;
    }
  }
  if constexpr (sizeof(ST) == 2) {
    if (periodic_box && !flags) {
      if (w64) launch_step_kernel<Q, ST, false, true, 2>(a, g, b, s, mode, nt);
      else launch_step_kernel<Q, ST, false, false, 2>(a, g, b, s, mode, nt);
      return;
    }
  }
  if (a.lanes && !periodic_box) {
    if (w64) {
      if (flags) launch_step_kernel<Q, ST, true, true, 3>(a, g, b, s, mode, nt);
      else launch_step_kernel<Q, ST, false, true, 3>(a, g, b, s, mode, nt);
    } else {
      if (flags) launch_step_kernel<Q, ST, true, false, 3>(a, g, b, s, mode, nt);
      else launch_step_kernel<Q, ST, false, false, 3>(a, g, b, s, mode, nt);
    }
    return;
  }
  if (w64) {
    if (flags) launch_step_kernel<Q, ST, true, true, 0>(a, g, b, s, mode, nt);
    else launch_step_kernel<Q, ST, false, true, 0>(a, g, b, s, mode, nt);
  } else {
    if (flags) launch_step_kernel<Q, ST, true, false, 0>(a, g, b, s, mode, nt);
    else launch_step_kernel<Q, ST, false, false, 0>(a, g, b, s, mode, nt);
  }
}

// Occupancy target of the batched step (waves per SIMD).  The tiles of a
// batch are small, so the bytes in flight per CU come from occupancy: the
// D2Q9 FP32 boundary-tile kernel (87 VGPRs, 5 waves unconstrained) runs
// 10 % faster held to 8 waves despite the spill this costs in its boundary
// paths.  PATH 1: tiles whose waves are all interior (no boundary code);
// 2: any tile, each wave taking the path of its class.
template <int Q, typename ST, int PATH>
struct BatchWaves {
  static constexpr int value = 1;
};
#ifndef LBM_BATCH_WAVES
#define LBM_BATCH_WAVES 8
#endif
template <>
struct BatchWaves<9, float, 2> {
  static constexpr int value = LBM_BATCH_WAVES;
};

template <int Q, typename ST, int MODE, bool NT, int PATH>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(BatchWaves<Q, ST, PATH>::value))) k_batch_step(const BatchArgs a) {
  using C = typename Store<ST>::C;
  using V4 = typename Vec4<C>::T;
  // Workgroups go to the 8 XCDs in turn, and each XCD has its own L2.  With
  // xcd_chunks the blocks of one XCD walk one contiguous range of the tile
  // list (domain-major, rows in order), so the tiles that share cache lines
  // (the x and y neighbours, through the shifted loads) run through the
  // same L2 instead of fetching those lines once per XCD.
  constexpr int NXCD = 8;
  int k = blockIdx.x, k_end = a.nids, k_step = gridDim.x;
  if (a.xcd_chunks && gridDim.x >= NXCD) {  // every XCD then has a block
    const int xcd = (int)(blockIdx.x % NXCD);
    const int chunk = (a.nids + NXCD - 1) / NXCD;
    k = xcd * chunk + (int)(blockIdx.x / NXCD);
    k_end = min(a.nids, (xcd + 1) * chunk);
    k_step = ((int)gridDim.x - xcd + NXCD - 1) / NXCD;  // blocks on this XCD
  }
  for (; k < k_end; k += k_step) {
    const int t = a.ids[k];
    const BatchTile tl = a.tiles[t];
    if (!a.active[tl.d]) continue;  // uniform per block
    const DomainDev &d = a.doms[tl.d];
    BatchView v;
    v.src = static_cast<const ST *>(a.src) + d.pop_off;
    v.dst = static_cast<ST *>(a.dst) + d.pop_off;
    v.links = a.links + d.cell_off;
    v.plane = d.plane;
    v.nx = d.nx;
    v.ny = d.ny;
    v.nz = d.nz;
    v.px = d.px;
    v.per[0] = d.per[0];
    v.per[1] = d.per[1];
    v.per[2] = d.per[2];
    v.face = d.face;
    v.omega = d.omega;
    v.F[0] = v.F[1] = v.F[2] = 0.0;
    v.snap_in = a.snap_in ? static_cast<V4 *>(a.snap_in) + d.cell_off : nullptr;
    v.snap_out = a.snap_out ? static_cast<V4 *>(a.snap_out) + d.cell_off : nullptr;
    // lane i: cell s + i of the tile's rectangle; lanes past n stay idle
    // (x = nx fails cell_update's bounds test)
    int x = d.nx, y = 0;
    if ((int)threadIdx.x < tl.n) {
      const int j = tl.s + (int)threadIdx.x;
      const int r = j / tl.w;
      x = tl.xs + (j - r * tl.w);
      y = tl.ys + r;
    }
    const int z = tl.z;
    double red[NSUM];
#pragma unroll
    for (int k = 0; k < NSUM; ++k) red[k] = 0;
    // the path of this wave (uniform per wave): no boundary code (no solid
    // lane, no face or solid within one cell), the boundary paths with or
    // without the flag loads, or none (no fluid lane)
    const int wcls = (tl.cls >> (2 * ((int)threadIdx.x >> 6))) & 3;
    if (PATH == 1 || wcls == 0) cell_update<Q, ST, false, MODE, NT, false, 1>(v, x, y, z, red);
    else if (wcls == 2) cell_update<Q, ST, true, MODE, NT, false>(v, x, y, z, red);
    else if (wcls == 1) cell_update<Q, ST, false, MODE, NT, false>(v, x, y, z, red);
    if (MODE >= 2) {
      block_reduce_store<256>(red, a.partials, t, a.ntiles);
      __syncthreads();  // the LDS of the reduction is reused by the next tile
    }
  }
}

template <int Q, typename ST, int PATH>
static void launch_batch_kernel(const BatchArgs &a, int grid, hipStream_t s, int mode) {
  if (mode == 0) hipLaunchKernelGGL((k_batch_step<Q, ST, 0, true, PATH>), dim3(grid), dim3(256), 0, s, a);
  else if (mode == 1) hipLaunchKernelGGL((k_batch_step<Q, ST, 1, false, PATH>), dim3(grid), dim3(256), 0, s, a);
  else hipLaunchKernelGGL((k_batch_step<Q, ST, 2, false, PATH>), dim3(grid), dim3(256), 0, s, a);
}
template <int Q, typename ST>
void BatchStepL<Q, ST>::run(const BatchArgs &a, int grid, hipStream_t s, int mode, int path) {
  if (path == 1) launch_batch_kernel<Q, ST, 1>(a, grid, s, mode);
  else launch_batch_kernel<Q, ST, 2>(a, grid, s, mode);
}

}  // namespace lbm

#define LBM_INSTANTIATE_STEPS(ST)              \
  template struct lbm::StepL<9, ST>;           \
  template struct lbm::StepL<19, ST>;          \
  template struct lbm::BatchStepL<9, ST>;      \
  template struct lbm::BatchStepL<19, ST>;
