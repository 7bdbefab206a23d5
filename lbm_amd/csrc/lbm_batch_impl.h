// Batched uneven domains (include/lbm_batch.h).  Included at the end of
// lbm_hip.hip so that it shares the single-domain set-up kernels; the batch
// step kernel lives with the other step kernels (lbm_step_kernels.h) and the
// per-domain table (DomainDev, BatchArgs) in lbm_launch.h.
//
// Layout: every domain owns Q population planes of its own padded size in
// two shared buffers (A-B), a flag block and a snapshot block; a table of
// DomainDev in device memory holds offsets and parameters.  Work is a flat
// list of tiles (domain, x0, y0, z), each tile = 256 cells; the step kernel
// loops over tiles with a grid-stride loop, so one launch is either
// persistent (a few blocks per CU) or one block per tile -- the autotuner
// picks.  The check step reduces per tile into structure-of-arrays partials
// and k_domain_finalize sums each domain's contiguous tile range in order
// (deterministic), compares with the tolerance and retires the domain.

namespace lbm {

// One thread per domain: sum its tiles' partials in order, update the
// residual, retire converged or diverged domains.
__global__ void k_domain_finalize(const DomainDev *doms, int ndom, const double *partials, long long ntiles,
                                  int *active, double *res, double *mass, long long *conv_cycle, const long long *cycle,
                                  double tol) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= ndom || !active[d]) return;
  double diff = 0, norm = 0, m = 0, bad = 0;
  for (int t = doms[d].tile_begin; t < doms[d].tile_end; ++t) {
    m += partials[0 * ntiles + t];
    diff += partials[8 * ntiles + t];
    norm += partials[9 * ntiles + t];
    bad += partials[10 * ntiles + t];
  }
  const double r = norm > 0 ? sqrt(diff / norm) : (diff > 0 ? INFINITY : 0.0);
  res[d] = r;
  mass[d] = m;
  if (bad > 0 || !isfinite(r)) {
    active[d] = 0;
    conv_cycle[d] = -2;
  } else if (r < tol) {
    active[d] = 0;
    conv_cycle[d] = *cycle;
  }
}

__global__ void k_tick(long long *cycle) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *cycle += 1;
}

__global__ void k_fill_int(int *p, int n, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_fill_ll(long long *p, int n, long long v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace lbm

struct lbm_batch {
  int Q = 0, dim = 0, storage = 0, esize = 0;
  int ndom = 0;
  std::vector<DomainDev> doms;
  std::vector<int> ncells;  // compact cells per domain
  std::vector<std::vector<uint8_t>> solid;  // host copy of each domain's solid mask (compact)
  std::vector<int> active_h;                // host view of the active flags at the last tiling
  long long total_pop = 0, total_cells = 0, compact_cells = 0;
  void *buf[2] = {nullptr, nullptr};
  void *snap[2] = {nullptr, nullptr};
  uint8_t *flags = nullptr;
  uint32_t *links = nullptr;  // solid-link masks, same layout as flags
  DomainDev *doms_d = nullptr;
  BatchTile *tiles_d = nullptr;
  long long ntiles = 0;
  std::vector<std::vector<uint8_t>> cclass;       // every domain's cell classes (tiles)
  std::vector<std::vector<BatchTile>> dom_tiles;  // every domain's tiles for tiling cache_tx
  int cache_tx = -1;
  int *ids_d = nullptr;  // launch lists: interior tiles, boundary tiles, every tile
  int n_int = 0, n_bnd = 0;
  int tx = 64, blocks = 0;  // blocks == 0: one block per tile
  int xcd_chunks = 1;       // LBM_BATCH_XCD=0 turns the per-XCD tile ranges off
  int one_launch = 1;       // 1: one launch over every tile; 0: interior tiles, then the others
  int *active_d = nullptr;
  double *res_d = nullptr, *mass_d = nullptr, *partials = nullptr;
  long long *conv_d = nullptr, *cycle_d = nullptr;
  hipStream_t stream = nullptr;
  int cur = 0;
  hipGraphExec_t graph = nullptr;
  int graph_interval = 0;
  double graph_tol = -1;
  std::string err;
};

static int bfail(lbm_batch *b, int code, const std::string &msg) {
  if (b) b->err = msg;
  g_err = msg;
  return code;
}

#define BHIPCHK(b, call)                                                                     \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      return bfail((b), LBM_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));       \
  } while (0)

// The instantiated graph captures the tile list, the tile ids and the
// partial sums among its kernel arguments, and HIP keeps a reference to
// each argument's allocation in the captured kernel nodes; the graph must
// therefore be destroyed before any of them is freed, never after.
static void batch_drop_graph(lbm_batch *b) {
  if (b->graph) {
    (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
  }
}

// Tiles of one domain.  Every cell is classified first: solid (never
// updated), 0 interior (no face, wrap or solid within one cell: the path
// without boundary code), 1 boundary, 2 boundary with a solid within one
// cell (flag loads; solids across periodic faces count).  A wave takes the
// path of the highest class among its fluid cells (tile field cls, two bits
// per wave), class 2 when it also holds a solid cell, and 3 (no update)
// when it holds none of the domain's fluid cells.  tx > 0: 2D tiles of tx x (256 / tx) cells in one plane (tiles
// of solid cells only are left out).  tx == 0: rows: each row is cut into
// runs of one class, runs that repeat over consecutive rows become
// rectangles, and each rectangle is cut into linear pieces of 256 cells.
constexpr uint8_t CELL_SOLID = 255;
static void cell_classes(const DomainDev &D, const std::vector<uint8_t> &sol, std::vector<uint8_t> &cc) {
  const int nx = D.nx, ny = D.ny, nz = D.nz;
  // periodic axes of the lattice (a 2D domain's z faces carry the periodic
  // default but never wrap)
  const bool per3[3] = {D.per[0] != 0, D.per[1] != 0, nz > 1 && D.per[2] != 0};
  auto at = [&](int x, int y, int z) { return (size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * z); };
  // near[c]: a solid cell within one cell of c, across periodic faces too
  // (box dilation, axis by axis)
  std::vector<uint8_t> near;
  if (!sol.empty()) {
    near.resize(sol.size());
    for (size_t c = 0; c < sol.size(); ++c) near[c] = sol[c] != 0;
    std::vector<uint8_t> tmp(near.size());
    const int n3[3] = {nx, ny, nz};
    const size_t st3[3] = {1, (size_t)nx, (size_t)nx * ny};
    for (int ax = 0; ax < 3; ++ax) {
      if (n3[ax] == 1) continue;
      const size_t st = st3[ax];
      for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
          for (int x = 0; x < nx; ++x) {
            const int k = ax == 0 ? x : (ax == 1 ? y : z);
            const size_t c = at(x, y, z);
            const size_t span = (size_t)(n3[ax] - 1) * st;
            uint8_t v = near[c];
            if (k > 0) v |= near[c - st];
            else if (per3[ax]) v |= near[c + span];
            if (k < n3[ax] - 1) v |= near[c + st];
            else if (per3[ax]) v |= near[c - span];
            tmp[c] = v;
          }
      near.swap(tmp);
    }
  }
  cc.assign((size_t)nx * ny * nz, 0);
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        const size_t c = at(x, y, z);
        uint8_t k;
        if (!sol.empty() && sol[c]) k = CELL_SOLID;
        else if (!sol.empty() && near[c]) k = 2;
        else k = (x >= 1 && x <= nx - 2 && y >= 1 && y <= ny - 2 && (nz == 1 || (z >= 1 && z <= nz - 2))) ? 0 : 1;
        cc[c] = k;
      }
}

static void domain_tiles(const DomainDev &D, const std::vector<uint8_t> &cc, int d, int tx,
                         std::vector<BatchTile> &out) {
  const int nx = D.nx, ny = D.ny, nz = D.nz;
  constexpr uint8_t SOLID = CELL_SOLID;
  auto at = [&](int x, int y, int z) { return (size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * z); };
  // one 2D tile: the per-wave paths of its lanes (lane i: cell
  // (x0 + i % w, y0 + i / w)), left out when it holds no fluid cell
  auto tile = [&](int x0, int w, int y0, int z) {
    int mask = 0;
    bool fluid = false;
    for (int wv = 0; wv < 4; ++wv) {
      // the path without boundary code never sees a solid lane: it would
      // update the cell as fluid and, on a domain edge, pull from outside
      // the domain
      int wc = 0;
      bool wfluid = false, wsolid = false;
      for (int i = 64 * wv; i < 64 * wv + 64; ++i) {
        const int x = x0 + i % w, y = y0 + i / w;
        if (x >= nx || y >= ny) continue;
        const uint8_t k = cc[at(x, y, z)];
        if (k == SOLID) {
          wsolid = true;
          continue;
        }
        wfluid = true;
        wc = std::max(wc, (int)k);
      }
      if (!wfluid) wc = 3;      // nothing to update: the wave skips the cell update
      else if (wsolid) wc = 2;  // flag loads deactivate the solid lanes
      fluid = fluid || wfluid;
      mask |= wc << (2 * wv);
    }
    if (fluid) out.push_back(BatchTile{d, x0, w, y0, 0, 256, z, mask});
  };
  if (tx == 1) {
    // banded: 16x16 tiles over the first 16 columns and from the last
    // 64-wide tile on, 64x4 tiles between, so that a wave holding a face
    // cell of an x face holds 16 columns instead of 64 (16 % of the fluid
    // cells of the 2048 step channels in waves with a boundary cell
    // instead of 33 %); tiles go by 16-row bands, column by column
    std::vector<std::pair<int, int>> cols;
    int x = std::min(16, nx);
    cols.push_back({0, 16});
    while (x + 64 <= nx - 16) {
      cols.push_back({x, 64});
      x += 64;
    }
    for (; x < nx; x += 16) cols.push_back({x, 16});
    for (int z = 0; z < nz; ++z)
      for (int b0 = 0; b0 < ny; b0 += 16)
        for (const auto &c : cols)
          for (int y0 = b0; y0 < std::min(ny, b0 + 16); y0 += 256 / c.second) tile(c.first, c.second, y0, z);
    return;
  }
  if (tx > 0) {
    const int ty = 256 / tx;
    for (int z = 0; z < nz; ++z)
      for (int y0 = 0; y0 < ny; y0 += ty)
        for (int x0 = 0; x0 < nx; x0 += tx) tile(x0, tx, y0, z);
    return;
  }
  struct Run {
    int xa, xb, cls, ya;
  };
  struct Rect {
    int xa, xb, ya, yb, cls;
  };
  std::vector<Run> open, next, row;
  std::vector<Rect> rects;
  for (int z = 0; z < nz; ++z) {
    open.clear();
    rects.clear();
    for (int y = 0; y <= ny; ++y) {
      row.clear();
      if (y < ny)
        for (int x = 0; x < nx;) {
          auto cls = [&](int xx) {
            const uint8_t k = cc[at(xx, y, z)];
            return k == SOLID ? -1 : (int)k;
          };
          const int k = cls(x);
          int xe = x + 1;
          while (xe < nx && cls(xe) == k) ++xe;
          if (k >= 0) row.push_back(Run{x, xe, k, y});
          x = xe;
        }
      // continue the open rectangles whose run repeats in this row (both
      // lists are ordered by xa), close the others
      next.clear();
      size_t i = 0;
      for (const Run &r : row) {
        while (i < open.size() && open[i].xa < r.xa) {
          rects.push_back(Rect{open[i].xa, open[i].xb, open[i].ya, y, open[i].cls});
          ++i;
        }
        if (i < open.size() && open[i].xa == r.xa && open[i].xb == r.xb && open[i].cls == r.cls) {
          next.push_back(open[i]);
          ++i;
        } else {
          next.push_back(r);
        }
      }
      for (; i < open.size(); ++i) rects.push_back(Rect{open[i].xa, open[i].xb, open[i].ya, y, open[i].cls});
      open.swap(next);
    }
    std::sort(rects.begin(), rects.end(),
              [](const Rect &a, const Rect &b) { return a.ya != b.ya ? a.ya < b.ya : a.xa < b.xa; });
    for (const Rect &r : rects) {
      const int w = r.xb - r.xa, total = w * (r.yb - r.ya);
      for (int s0 = 0; s0 < total; s0 += 256)
        out.push_back(BatchTile{d, r.xa, w, r.ya, s0, std::min(256, total - s0), z, r.cls * 0x55});
    }
  }
}

// Tiles of every still-active domain (each domain's list is built once per
// tiling and kept).  The new buffers are allocated and filled before the
// old ones are freed, so a failed rebuild leaves the previous tiling in
// place (without a graph; the next run captures a new one).
static int batch_build_tiles(lbm_batch *b, int tx) {
  BHIPCHK(b, hipStreamSynchronize(b->stream));
  batch_drop_graph(b);
  if ((int)b->cclass.size() != b->ndom) {
    b->cclass.assign(b->ndom, {});
    for (int d = 0; d < b->ndom; ++d) cell_classes(b->doms[d], b->solid[d], b->cclass[d]);
  }
  if (b->cache_tx != tx || (int)b->dom_tiles.size() != b->ndom) {
    b->dom_tiles.assign(b->ndom, {});
    for (int d = 0; d < b->ndom; ++d) domain_tiles(b->doms[d], b->cclass[d], d, tx, b->dom_tiles[d]);
    b->cache_tx = tx;
  }
  std::vector<BatchTile> tiles;
  std::vector<DomainDev> doms = b->doms;
  // interior tiles run in a kernel compiled without the boundary paths
  std::vector<int> ids_int, ids_bnd;
  for (int d = 0; d < b->ndom; ++d) {
    DomainDev &D = doms[d];
    D.tile_begin = (int)tiles.size();
    if (b->active_h.empty() || b->active_h[d])
      for (const BatchTile &t : b->dom_tiles[d]) {
        (t.cls == 0 ? ids_int : ids_bnd).push_back((int)tiles.size());  // every wave interior, or not
        tiles.push_back(t);
      }
    D.tile_end = (int)tiles.size();
  }
  if (tiles.empty()) {
    ids_bnd.push_back(0);
    tiles.push_back(BatchTile{0, 0, 1, 0, 0, 0, 0, 1});  // no lane active
  }
  // launch lists: the interior tiles, the boundary tiles, then every tile in
  // list order (the single-launch form)
  std::vector<int> ids(ids_int);
  ids.insert(ids.end(), ids_bnd.begin(), ids_bnd.end());
  for (int t = 0; t < (int)tiles.size(); ++t) ids.push_back(t);
  BatchTile *tiles_d = nullptr;
  int *ids_d = nullptr;
  double *partials = nullptr;
  auto release = [&](int code, const std::string &msg) {
    if (tiles_d) (void)hipFree(tiles_d);
    if (ids_d) (void)hipFree(ids_d);
    if (partials) (void)hipFree(partials);
    return bfail(b, code, msg);
  };
  if (hipMalloc(&tiles_d, tiles.size() * sizeof(BatchTile)) != hipSuccess ||
      hipMalloc(&ids_d, ids.size() * sizeof(int)) != hipSuccess ||
      hipMalloc(&partials, tiles.size() * NSUM * sizeof(double)) != hipSuccess)
    return release(LBM_ENOMEM, "batch tiling: device allocation failed");
  if (hipMemcpy(tiles_d, tiles.data(), tiles.size() * sizeof(BatchTile), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(ids_d, ids.data(), ids.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->doms_d, doms.data(), doms.size() * sizeof(DomainDev), hipMemcpyHostToDevice) != hipSuccess)
    return release(LBM_EHIP, "batch tiling: upload failed");
  if (b->tiles_d) (void)hipFree(b->tiles_d);
  if (b->ids_d) (void)hipFree(b->ids_d);
  if (b->partials) (void)hipFree(b->partials);
  b->tiles_d = tiles_d;
  b->ids_d = ids_d;
  b->partials = partials;
  b->doms = std::move(doms);
  b->ntiles = (long long)tiles.size();
  b->n_int = (int)ids_int.size();
  b->n_bnd = (int)ids_bnd.size();
  b->tx = tx;
  return LBM_OK;
}

static BatchArgs batch_args(lbm_batch *b, int parity, int mode) {
  BatchArgs a;
  memset(&a, 0, sizeof(a));
  a.src = b->buf[parity];
  a.dst = b->buf[parity ^ 1];
  a.links = b->links;
  a.doms = b->doms_d;
  a.tiles = b->tiles_d;
  a.ntiles = (int)b->ntiles;
  a.active = b->active_d;
  a.xcd_chunks = b->xcd_chunks;
  if (mode >= 1) {
    // snapshot written by the step before the check, read by the check
    a.snap_in = b->snap[0];
    a.snap_out = b->snap[0];
    a.partials = b->partials;
  }
  return a;
}

// blocks of one of the two launches: one per tile, or the persistent grid
// shared in proportion to the tiles (at least one block per XCD)
static int batch_grid(const lbm_batch *b, int n) {
  if (b->blocks <= 0) return n;
  const long long g = (long long)b->blocks * n / std::max<long long>(1, b->ntiles);
  return (int)std::max<long long>(std::min<long long>(8, n), g);
}

static void batch_launch(lbm_batch *b, int parity, int mode) {
  BatchArgs a = batch_args(b, parity, mode);
  if (b->one_launch) {
    // every tile in list order (domain-major, rows in order), so a tile and
    // its neighbours of either class share the L2 of their XCD
    a.ids = b->ids_d + b->n_int + b->n_bnd;
    a.nids = (int)b->ntiles;
    dispatch_qs<BatchStepL>(b->Q, b->storage, a, batch_grid(b, (int)b->ntiles), b->stream, mode, 2);
    return;
  }
  if (b->n_int > 0) {
    a.ids = b->ids_d;
    a.nids = b->n_int;
    dispatch_qs<BatchStepL>(b->Q, b->storage, a, batch_grid(b, b->n_int), b->stream, mode, 1);
  }
  if (b->n_bnd > 0) {
    a.ids = b->ids_d + b->n_int;
    a.nids = b->n_bnd;
    dispatch_qs<BatchStepL>(b->Q, b->storage, a, batch_grid(b, b->n_bnd), b->stream, mode, 2);
  }
}

extern "C" {

int lbm_batch_create(int32_t lattice, int32_t storage, const lbm_domain_desc *doms, int32_t ndom, lbm_batch **out) {
  if (!doms || !out || ndom <= 0) return bfail(nullptr, LBM_EINVAL, "lbm_batch_create: bad argument");
  *out = nullptr;
  if (lattice != LBM_D2Q9 && lattice != LBM_D3Q19) return bfail(nullptr, LBM_EINVAL, "lbm_batch_create: lattice 9 or 19");
  const int es = esize_of(storage);
  if (!es) return bfail(nullptr, LBM_EINVAL, "lbm_batch_create: unknown storage");
  lbm_batch *b = new lbm_batch();
  if (const char *e = getenv("LBM_BATCH_XCD")) b->xcd_chunks = atoi(e) != 0;
  b->Q = lattice;
  b->dim = lattice == LBM_D2Q9 ? 2 : 3;
  b->storage = storage;
  b->esize = es;
  b->ndom = ndom;
  auto bail = [&](int code) {
    lbm_batch_destroy(b);
    return code;
  };
  std::vector<uint8_t> flags;
  long long pop = 0, cells = 0, compact = 0;
  for (int d = 0; d < ndom; ++d) {
    const lbm_domain_desc &dd = doms[d];
    if (dd.nx < 3 || dd.ny < 3 || dd.nz < 1 || (b->dim == 2 && dd.nz != 1) || (b->dim == 3 && dd.nz < 3)) {
      bfail(nullptr, LBM_EINVAL, "lbm_batch_create: domain " + std::to_string(d) + " has an invalid size");
      return bail(LBM_EINVAL);
    }
    if (!(dd.omega > 0 && dd.omega < 2)) {
      bfail(nullptr, LBM_EINVAL, "lbm_batch_create: omega must be in (0,2)");
      return bail(LBM_EINVAL);
    }
    DomainDev D;
    memset(&D, 0, sizeof(D));
    D.nx = dd.nx;
    D.ny = dd.ny;
    D.nz = dd.nz;
    // rows start on 128-byte lines: 16 (FP64), 32 (FP32) or 64 (16-bit) cells
    const int line = 128 / es;
    D.px = (dd.nx + line - 1) / line * line;
    D.plane = (long long)D.px * D.ny * D.nz;
    if ((unsigned long long)D.plane * (unsigned long long)es > 0xffffffffull) {
      bfail(nullptr, LBM_EINVAL, "lbm_batch_create: domain " + std::to_string(d) +
                                     " is too large for a batch (a population plane above 4 GiB); run it alone");
      return bail(LBM_EINVAL);
    }
    D.pop_off = pop;
    D.cell_off = cells;
    D.omega = dd.omega;
    for (int f = 0; f < 6; ++f) {
      D.face[f].kind = LBM_FACE_PERIODIC;
      D.face[f].rho = 1.0;
    }
    for (int i = 0; i < dd.nfaces; ++i) {
      const lbm_bc_desc &f = dd.faces[i];
      if (f.face < 0 || f.face >= 2 * b->dim || f.kind < 0 || f.kind > LBM_FACE_PERIODIC_PRESSURE) {
        bfail(nullptr, LBM_EINVAL, "lbm_batch_create: bad face in domain " + std::to_string(d));
        return bail(LBM_EINVAL);
      }
      D.face[f.face].kind = f.kind;
      D.face[f.face].prio = f.priority;
      for (int k = 0; k < 3; ++k) D.face[f.face].u[k] = f.value[k];
      D.face[f.face].rho = f.rho;
    }
    for (int ax = 0; ax < 3; ++ax) {
      const bool p0 = D.face[2 * ax].kind == LBM_FACE_PERIODIC, p1 = D.face[2 * ax + 1].kind == LBM_FACE_PERIODIC;
      const bool q0 = D.face[2 * ax].kind == LBM_FACE_PERIODIC_PRESSURE, q1 = D.face[2 * ax + 1].kind == LBM_FACE_PERIODIC_PRESSURE;
      if (ax < b->dim && (p0 != p1 || q0 != q1)) {
        bfail(nullptr, LBM_EINVAL, "lbm_batch_create: periodic face without partner");
        return bail(LBM_EINVAL);
      }
      D.per[ax] = p0 ? 1 : 0;
    }
    if (!neem_neighbours_ok(D.face, b->dim)) {
      bfail(nullptr, LBM_EINVAL, "lbm_batch_create: the faces next to a NEEM face must be walls or periodic");
      return bail(LBM_EINVAL);
    }
    const long long nc = D.plane;
    flags.resize(cells + nc, LBM_CELL_SOLID);
    for (int z = 0; z < D.nz; ++z)
      for (int y = 0; y < D.ny; ++y)
        for (int x = 0; x < D.nx; ++x) {
          const uint8_t v = dd.solid ? dd.solid[x + (long long)D.nx * (y + (long long)D.ny * z)] : LBM_CELL_FLUID;
          flags[cells + x + (long long)D.px * (y + (long long)D.ny * z)] = v ? LBM_CELL_SOLID : LBM_CELL_FLUID;
        }
    b->doms.push_back(D);
    b->ncells.push_back(D.nx * D.ny * D.nz);
    std::vector<uint8_t> sol;
    if (dd.solid) sol.assign(dd.solid, dd.solid + (size_t)D.nx * D.ny * D.nz);
    b->solid.push_back(std::move(sol));
    pop += (long long)b->Q * D.plane;
    cells += nc;
    compact += (long long)D.nx * D.ny * D.nz;
  }
  b->total_pop = pop;
  b->total_cells = cells;
  b->compact_cells = compact;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) return bail(bfail(nullptr, LBM_EHIP, "stream"));
  for (int k = 0; k < 2; ++k) {
    if (hipMalloc(&b->buf[k], (size_t)pop * es) != hipSuccess) return bail(bfail(nullptr, LBM_ENOMEM, "population buffers"));
    if (hipMemset(b->buf[k], 0, (size_t)pop * es) != hipSuccess) return bail(bfail(nullptr, LBM_EHIP, "memset"));
  }
  const size_t cs = snap_bytes(storage);
  if (hipMalloc(&b->snap[0], (size_t)cells * cs) != hipSuccess) return bail(bfail(nullptr, LBM_ENOMEM, "snapshot"));
  if (hipMalloc(&b->flags, flags.size()) != hipSuccess) return bail(bfail(nullptr, LBM_ENOMEM, "flags"));
  if (hipMemcpy(b->flags, flags.data(), flags.size(), hipMemcpyHostToDevice) != hipSuccess) return bail(bfail(nullptr, LBM_EHIP, "flags"));
  if (hipMalloc(&b->links, flags.size() * sizeof(uint32_t)) != hipSuccess) return bail(bfail(nullptr, LBM_ENOMEM, "link masks"));
  for (int d = 0; d < ndom; ++d) {
    const DomainDev &D = b->doms[d];
    if (b->Q == 9) launch_links<9>(b->flags + D.cell_off, b->links + D.cell_off, D.nx, D.ny, D.nz, D.px, D.per, b->stream);
    else launch_links<19>(b->flags + D.cell_off, b->links + D.cell_off, D.nx, D.ny, D.nz, D.px, D.per, b->stream);
  }
  if (hipGetLastError() != hipSuccess) return bail(bfail(nullptr, LBM_EHIP, "link masks"));
  if (hipMalloc(&b->doms_d, ndom * sizeof(DomainDev)) != hipSuccess) return bail(bfail(nullptr, LBM_ENOMEM, "domain table"));
  if (hipMalloc(&b->active_d, ndom * sizeof(int)) != hipSuccess) return bail(bfail(nullptr, LBM_ENOMEM, "active"));
  if (hipMalloc(&b->res_d, ndom * sizeof(double)) != hipSuccess) return bail(bfail(nullptr, LBM_ENOMEM, "res"));
  if (hipMalloc(&b->mass_d, ndom * sizeof(double)) != hipSuccess) return bail(bfail(nullptr, LBM_ENOMEM, "mass"));
  if (hipMalloc(&b->conv_d, ndom * sizeof(long long)) != hipSuccess) return bail(bfail(nullptr, LBM_ENOMEM, "conv"));
  if (hipMalloc(&b->cycle_d, sizeof(long long)) != hipSuccess) return bail(bfail(nullptr, LBM_ENOMEM, "cycle"));
  hipLaunchKernelGGL(k_fill_int, dim3((ndom + 255) / 256), dim3(256), 0, b->stream, b->active_d, ndom, 1);
  if (hipStreamSynchronize(b->stream) != hipSuccess) return bail(bfail(nullptr, LBM_EHIP, "init"));
  int rc = batch_build_tiles(b, 64);
  if (rc) return bail(rc);
  *out = b;
  return LBM_OK;
}

int lbm_batch_set_tiling(lbm_batch *b, int32_t tile_x, int32_t blocks) {
  if (!b || !(tile_x == 0 || tile_x == 1 || tile_x == 16 || tile_x == 32 || tile_x == 64 || tile_x == 128 ||
              tile_x == 256) || blocks < 0)
    return bfail(b, LBM_EINVAL, "lbm_batch_set_tiling: tile_x in {0,1,16,32,64,128,256}, blocks >= 0");
  b->blocks = blocks;
  return batch_build_tiles(b, tile_x);
}

int lbm_batch_set_launch(lbm_batch *b, int32_t one_launch) {
  if (!b || one_launch < 0 || one_launch > 1) return bfail(b, LBM_EINVAL, "lbm_batch_set_launch: one_launch 0 or 1");
  BHIPCHK(b, hipStreamSynchronize(b->stream));
  batch_drop_graph(b);  // the next run captures the new form
  b->one_launch = one_launch;
  return LBM_OK;
}

int lbm_batch_tune(lbm_batch *b, int32_t steps_per_trial, char *choice, int32_t choice_len) {
  if (!b || steps_per_trial < 1) return bfail(b, LBM_EINVAL, "lbm_batch_tune: bad argument");
  int dev = 0, ncu = 256;
  if (hipGetDevice(&dev) == hipSuccess) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess) ncu = prop.multiProcessorCount;
  }
  EventPair ev;
  BHIPCHK(b, ev.create());
  hipEvent_t e0 = ev.e[0], e1 = ev.e[1];
  struct Cand {
    int tx, blocks, one;
    double ms;
  };
  // time steps_per_trial steps of one candidate as the run executes them:
  // captured in a graph, replayed once untimed, then timed
  const int nst = steps_per_trial + (steps_per_trial & 1);
  auto trial = [&](const Cand &c, double &ms) -> int {
    if (b->tx != c.tx || b->cache_tx != c.tx) {
      const int rc = batch_build_tiles(b, c.tx);
      if (rc) return rc;
    }
    b->blocks = c.blocks;
    b->one_launch = c.one;
    hipGraph_t g;
    BHIPCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeGlobal));
    for (int s = 0; s < nst; ++s) batch_launch(b, (b->cur + s) & 1, 0);
    BHIPCHK(b, hipStreamEndCapture(b->stream, &g));
    hipGraphExec_t ge = nullptr;
    const hipError_t ie = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    BHIPCHK(b, ie);
    hipError_t e = hipGraphLaunch(ge, b->stream);
    if (e == hipSuccess) e = hipEventRecord(e0, b->stream);
    if (e == hipSuccess) e = hipGraphLaunch(ge, b->stream);
    if (e == hipSuccess) e = hipEventRecord(e1, b->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    if (e != hipSuccess) (void)hipStreamSynchronize(b->stream);
    (void)hipGraphExecDestroy(ge);  // after the synchronisation: nothing runs it any more
    BHIPCHK(b, e);
    ms = t / nst;
    return LBM_OK;
  };
  const int txs[7] = {0, 1, 16, 32, 64, 128, 256};
  const int per_cu[4] = {0, 2, 4, 16};
  std::vector<Cand> cands;
  for (int tx : txs)
    for (int one = 1; one >= 0; --one)
      for (int g : per_cu) cands.push_back(Cand{tx, g * ncu, one, 0.0});
  // untimed steps first, so that no candidate is timed on an idle GPU
  double ms = 0;
  for (int r = 0; r < 2; ++r) {
    const int rc = trial(Cand{64, 0, 1, 0.0}, ms);
    if (rc) return rc;
  }
  for (Cand &c : cands) {
    const int rc = trial(c, c.ms);
    if (rc) return rc;
  }
  // the four fastest are timed twice more, in turn, and ranked by the sum
  // of their three timings (one sweep is noisy at the 5 % level that
  // separates the best tilings)
  std::sort(cands.begin(), cands.end(), [](const Cand &a, const Cand &c) { return a.ms < c.ms; });
  cands.resize(std::min<size_t>(4, cands.size()));
  for (int r = 0; r < 2; ++r)
    for (Cand &c : cands) {
      const int rc = trial(c, ms);
      if (rc) return rc;
      c.ms += ms;
    }
  const Cand best = *std::min_element(cands.begin(), cands.end(),
                                      [](const Cand &a, const Cand &c) { return a.ms < c.ms; });
  b->blocks = best.blocks;
  b->one_launch = best.one;
  int rc = batch_build_tiles(b, best.tx);
  if (rc) return rc;
  if (choice && choice_len > 0)
    snprintf(choice, choice_len, "%d,%d,%d,%.4f", best.tx, best.blocks, best.one, best.ms / 3);
  return LBM_OK;
}

int lbm_batch_init_fields(lbm_batch *b, const double *rho, const double *u) {
  if (!b || !rho || !u) return bfail(b, LBM_EINVAL, "lbm_batch_init_fields: null argument");
  DevBuf rbuf, ubuf;
  BHIPCHK(b, rbuf.alloc(b->compact_cells * sizeof(double)));
  BHIPCHK(b, ubuf.alloc(b->compact_cells * b->dim * sizeof(double)));
  double *rd = static_cast<double *>(rbuf.p), *ud = static_cast<double *>(ubuf.p);
  BHIPCHK(b, hipMemcpy(rd, rho, b->compact_cells * sizeof(double), hipMemcpyHostToDevice));
  BHIPCHK(b, hipMemcpy(ud, u, b->compact_cells * b->dim * sizeof(double), hipMemcpyHostToDevice));
  long long coff = 0;
  b->cur = 0;
  if (!b->active_h.empty()) {
    b->active_h.clear();
    const int rc = batch_build_tiles(b, b->tx);
    if (rc) return rc;
  }
  for (int d = 0; d < b->ndom; ++d) {
    const DomainDev &D = b->doms[d];
    KArgs a;
    memset(&a, 0, sizeof(a));
    a.dst = static_cast<char *>(b->buf[0]) + D.pop_off * b->esize;
    a.flags = b->flags + D.cell_off;
    a.plane = D.plane;
    a.nx = D.nx; a.ny = D.ny; a.nz = D.nz; a.px = D.px;
    a.omega = D.omega;
    a.snap_out = nullptr;
    dim3 g((D.nx + 63) / 64, (D.ny + 3) / 4, D.nz), bl(64, 4, 1);
    dispatch_qs<InitL>(b->Q, b->storage, a, g, bl, b->stream, (const double *)(rd + coff),
                       (const double *)(ud + coff * b->dim));
    coff += b->ncells[d];
  }
  BHIPCHK(b, hipGetLastError());
  hipLaunchKernelGGL(k_fill_int, dim3((b->ndom + 255) / 256), dim3(256), 0, b->stream, b->active_d, b->ndom, 1);
  hipLaunchKernelGGL(k_fill_ll, dim3((b->ndom + 255) / 256), dim3(256), 0, b->stream, b->conv_d, b->ndom, -1LL);
  hipLaunchKernelGGL(k_fill_ll, dim3(1), dim3(64), 0, b->stream, b->cycle_d, 1, 0LL);
  BHIPCHK(b, hipStreamSynchronize(b->stream));
  return LBM_OK;
}

static int batch_capture(lbm_batch *b, int interval, double tol) {
  batch_drop_graph(b);
  hipGraph_t g;
  BHIPCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeGlobal));
  for (int s = 0; s < interval; ++s) {
    const int mode = s == interval - 2 ? 1 : (s == interval - 1 ? 2 : 0);
    batch_launch(b, (b->cur + s) & 1, mode);
  }
  hipLaunchKernelGGL(k_domain_finalize, dim3((b->ndom + 255) / 256), dim3(256), 0, b->stream, (const DomainDev *)b->doms_d,
                     b->ndom, (const double *)b->partials, b->ntiles, b->active_d, b->res_d, b->mass_d, b->conv_d,
                     (const long long *)b->cycle_d, tol);
  hipLaunchKernelGGL(k_tick, dim3(1), dim3(64), 0, b->stream, b->cycle_d);
  BHIPCHK(b, hipStreamEndCapture(b->stream, &g));
  BHIPCHK(b, hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0));
  BHIPCHK(b, hipGraphDestroy(g));
  b->graph_interval = interval;
  b->graph_tol = tol;
  return LBM_OK;
}

int lbm_batch_run(lbm_batch *b, int32_t interval, double tol, int64_t max_steps, lbm_domain_result *results, double *ms) {
  if (!b || interval < 2 || interval % 2 || max_steps < 0) return bfail(b, LBM_EINVAL, "lbm_batch_run: interval must be even >= 2");
  if (!b->graph || b->graph_interval != interval || b->graph_tol != tol) {
    const int rc = batch_capture(b, interval, tol);
    if (rc) return rc;
  }
  std::vector<int> active(b->ndom);
  EventPair ev;
  BHIPCHK(b, ev.create());
  hipEvent_t e0 = ev.e[0], e1 = ev.e[1];
  BHIPCHK(b, hipEventRecord(e0, b->stream));
  long long done = 0;
  int tiled_active = 0;
  for (int d = 0; d < b->ndom; ++d) tiled_active += b->active_h.empty() ? 1 : b->active_h[d];
  while (done + interval <= max_steps) {
    BHIPCHK(b, hipGraphLaunch(b->graph, b->stream));
    done += interval;
    BHIPCHK(b, hipMemcpyAsync(active.data(), b->active_d, b->ndom * sizeof(int), hipMemcpyDeviceToHost, b->stream));
    BHIPCHK(b, hipStreamSynchronize(b->stream));
    int nact = 0;
    for (int v : active) nact += v != 0;
    if (!nact) break;
    // drop the tiles of retired domains once a quarter of them is idle
    if (nact * 4 <= tiled_active * 3) {
      b->active_h = active;
      int rc = batch_build_tiles(b, b->tx);
      if (rc) return rc;
      rc = batch_capture(b, interval, tol);
      if (rc) return rc;
      tiled_active = nact;
    }
  }
  BHIPCHK(b, hipEventRecord(e1, b->stream));
  BHIPCHK(b, hipEventSynchronize(e1));
  float t = 0;
  BHIPCHK(b, hipEventElapsedTime(&t, e0, e1));
  if (ms) *ms = t;
  if (results) {
    std::vector<double> res(b->ndom), mass(b->ndom);
    std::vector<long long> conv(b->ndom);
    long long cyc = 0;
    BHIPCHK(b, hipMemcpy(res.data(), b->res_d, b->ndom * sizeof(double), hipMemcpyDeviceToHost));
    BHIPCHK(b, hipMemcpy(mass.data(), b->mass_d, b->ndom * sizeof(double), hipMemcpyDeviceToHost));
    BHIPCHK(b, hipMemcpy(conv.data(), b->conv_d, b->ndom * sizeof(long long), hipMemcpyDeviceToHost));
    BHIPCHK(b, hipMemcpy(&cyc, b->cycle_d, sizeof(long long), hipMemcpyDeviceToHost));
    for (int d = 0; d < b->ndom; ++d) {
      results[d].steps = conv[d] >= 0 ? (conv[d] + 1) * (long long)interval : conv[d];
      results[d].res_rel = res[d];
      results[d].mass = mass[d];
    }
  }
  return LBM_OK;
}

int lbm_batch_get_macros(lbm_batch *b, int32_t dom, double *rho, double *u) {
  if (!b || dom < 0 || dom >= b->ndom || !rho || !u) return bfail(b, LBM_EINVAL, "lbm_batch_get_macros: bad argument");
  const DomainDev &D = b->doms[dom];
  const long long n = (long long)D.nx * D.ny * D.nz;
  DevBuf rbuf, ubuf;
  BHIPCHK(b, rbuf.alloc(n * sizeof(double)));
  BHIPCHK(b, ubuf.alloc(n * b->dim * sizeof(double)));
  double *rd = static_cast<double *>(rbuf.p), *ud = static_cast<double *>(ubuf.p);
  KArgs a;
  memset(&a, 0, sizeof(a));
  a.src = static_cast<char *>(b->buf[b->cur]) + D.pop_off * b->esize;
  a.flags = b->flags + D.cell_off;
  a.plane = D.plane;
  a.nx = D.nx; a.ny = D.ny; a.nz = D.nz; a.px = D.px;
  dim3 g((D.nx + 63) / 64, (D.ny + 3) / 4, D.nz), bl(64, 4, 1);
  dispatch_qs<MacroL>(b->Q, b->storage, a, g, bl, b->stream, rd, ud);
  BHIPCHK(b, hipGetLastError());
  BHIPCHK(b, hipMemcpyAsync(rho, rd, n * sizeof(double), hipMemcpyDeviceToHost, b->stream));
  BHIPCHK(b, hipMemcpyAsync(u, ud, n * b->dim * sizeof(double), hipMemcpyDeviceToHost, b->stream));
  BHIPCHK(b, hipStreamSynchronize(b->stream));
  return LBM_OK;
}

int lbm_batch_get_populations(lbm_batch *b, int32_t dom, double *f) {
  if (!b || dom < 0 || dom >= b->ndom || !f) return bfail(b, LBM_EINVAL, "lbm_batch_get_populations: bad argument");
  const DomainDev &D = b->doms[dom];
  const long long n = (long long)D.nx * D.ny * D.nz;
  DevBuf fbuf;
  BHIPCHK(b, fbuf.alloc(n * b->Q * sizeof(double)));
  double *fd = static_cast<double *>(fbuf.p);
  KArgs a;
  memset(&a, 0, sizeof(a));
  a.src = static_cast<char *>(b->buf[b->cur]) + D.pop_off * b->esize;
  a.plane = D.plane;
  a.nx = D.nx; a.ny = D.ny; a.nz = D.nz; a.px = D.px;
  dim3 g((D.nx + 63) / 64, (D.ny + 3) / 4, D.nz), bl(64, 4, 1);
  dispatch_qs<GetL>(b->Q, b->storage, a, g, bl, b->stream, fd);
  BHIPCHK(b, hipGetLastError());
  BHIPCHK(b, hipMemcpyAsync(f, fd, n * b->Q * sizeof(double), hipMemcpyDeviceToHost, b->stream));
  BHIPCHK(b, hipStreamSynchronize(b->stream));
  return LBM_OK;
}

int lbm_batch_info(const lbm_batch *b, int64_t *ncell, int64_t *ntiles, int32_t *tile_x, int32_t *blocks) {
  if (!b) return bfail(nullptr, LBM_EINVAL, "lbm_batch_info: null");
  if (ncell) *ncell = b->compact_cells;
  if (ntiles) *ntiles = b->ntiles;
  if (tile_x) *tile_x = b->tx;
  if (blocks) *blocks = b->blocks;
  return LBM_OK;
}

const char *lbm_batch_last_error(const lbm_batch *b) { return b ? b->err.c_str() : g_err.c_str(); }

int lbm_batch_destroy(lbm_batch *b) {
  if (!b) return LBM_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  for (int k = 0; k < 2; ++k) {
    if (b->buf[k]) (void)hipFree(b->buf[k]);
    if (b->snap[k]) (void)hipFree(b->snap[k]);
  }
  void *ptrs[] = {b->flags, b->links, b->doms_d, b->tiles_d, b->ids_d, b->active_d, b->res_d, b->mass_d, b->partials, b->conv_d, b->cycle_d};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return LBM_OK;
}

}  // extern "C"
