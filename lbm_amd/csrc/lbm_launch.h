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
  const int4 *tiles;
  int ntiles;
  int tx;  // tile width in cells (tile = tx x (256/tx))
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
  // mode: 0 plain, 1 snapshot, 2 check; interior: the tiles in a.ids have
  // no boundary within one cell (compiled without the boundary paths)
  static void run(const BatchArgs &a, int grid, hipStream_t s, int mode, bool interior);
};

}  // namespace lbm
