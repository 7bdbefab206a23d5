// Launchers of the step kernels.  The kernels are compiled in one
// translation unit per storage type (lbm_step_<storage>.hip), which define
// and explicitly instantiate these launchers; the API translation unit only
// calls them.
#pragma once
#include <hip/hip_runtime.h>

#include "lbm_cell.h"

namespace lbm {

// Batched domains: per-domain table in device memory (see lbm_batch_impl.h).
struct DomainDev {
  long long pop_off;   // elements: start of plane 0 of this domain
  long long cell_off;  // cells: start of this domain's flags / snapshot block
  long long plane;     // elements per plane
  int nx, ny, nz, px;
  int per[3];
  int tile_begin, tile_end;
  FaceDev face[6];
  double omega;
};

// One tile of a batch: up to 256 cells of domain d, the linear range
// [s, s + n) of the row-major rectangle of width w whose first cell is
// (xs, ys) in plane z; lane i updates cell s + i.  Bits 2w, 2w + 1 of cls:
// the path of wave w, 0: no face, wrap or solid within one cell of its
// cells (no boundary code); 1: boundary paths without flag loads; 2:
// boundary paths and flag loads (a solid within one cell, or a solid
// lane); 3: no fluid lane (no update).  A 2D tile of
// tx x (256 / tx) cells is the rectangle of width tx at (x0, y0) with
// s = 0, n = 256 (lanes outside the domain are idle).
struct BatchTile {
  int d, xs, w, ys, s, n, z, cls;
};

// G for cell_update: one domain of a batch
struct BatchView {
  const void *src;
  void *dst;
  const uint32_t *links;
  long long plane;
  int nx, ny, nz, px;
  int per[3];
  const FaceDev *face;
  double omega;
  double F[3];
  void *snap_in;
  void *snap_out;
};

struct BatchArgs {
  const void *src;  // buffer base pointers of this step
  void *dst;
  const uint32_t *links;  // solid-link masks of all domains (cell_off layout)
  const DomainDev *doms;
  const BatchTile *tiles;
  int ntiles;
  int xcd_chunks;  // 1: each XCD walks a contiguous range of tiles
  const int *ids;  // the tiles this launch covers (indices into tiles)
  int nids;
  const int *active;
  void *snap_in;
  void *snap_out;
  double *partials;
};

template <int Q, typename ST>
struct StepL {
  // mode: 0 plain, 1 snapshot, 2 check, 3 check + snapshot; nt: nontemporal
  // loads/stores of the plain step; w64: 64-bit population offsets
  // lds: the LDS-staged x-halo kernel (FP32 storage, 32-bit offsets, 256x1 blocks)
  static void run(const KArgs &a, dim3 g, dim3 b, hipStream_t s, bool flags, int mode, bool nt, bool w64,
                  bool periodic_box, bool lds);
};

template <int Q, typename ST>
struct BatchStepL {
  // mode: 0 plain, 1 snapshot, 2 check; path 1: every wave of the tiles in
  // a.ids is interior (compiled without the boundary paths); 2: any tiles,
  // each wave branching on its class
  static void run(const BatchArgs &a, int grid, hipStream_t s, int mode, int path);
};

}  // namespace lbm
