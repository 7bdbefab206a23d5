/*
 * lbm_batch.h -- batches of many small, uneven lattice-Boltzmann domains in
 * one launch per step (one block per tile or a persistent grid, as the
 * autotuner chooses), K steps plus the convergence check captured in one
 * hipGraph (part of liblbm_hip.so).
 *
 * No reference counterpart: SFCMM/LBM runs one domain per process
 * (src/lbm/solver.cpp:176-214).  Each domain here is a structured box with its
 * own size, relaxation rate, boundary faces and solid cells; every domain
 * follows exactly the single-domain rules of lbm_hip.h, so a domain of a batch
 * and the same domain run alone through lbm_step give identical populations.
 * Converged domains (relative residual below the tolerance) stop being
 * updated.  Same status codes as lbm_hip.h.
 */
#ifndef LBM_BATCH_H
#define LBM_BATCH_H
#include <stdint.h>

#include "lbm_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lbm_batch lbm_batch;

typedef struct {
  int32_t nx, ny, nz;       /* nz = 1 for D2Q9 */
  double omega;
  const uint8_t *solid;     /* nx*ny*nz flags (LBM_CELL_*), may be NULL */
  lbm_bc_desc faces[6];
  int32_t nfaces;
} lbm_domain_desc;

/* Per-domain result of lbm_batch_run. */
typedef struct {
  int64_t steps;   /* steps until the residual fell below tol, -1 if not reached, -2 if diverged */
  double res_rel;  /* last relative residual ||u^n - u^(n-1)|| / ||u^n|| */
  double mass;     /* last checked mass */
} lbm_domain_result;

int lbm_batch_create(int32_t lattice, int32_t storage, const lbm_domain_desc *doms, int32_t ndom, lbm_batch **out);

/* Autotuner: times every candidate tiling (2D tiles of 256 cells and width
 * 16/32/64/128/256; tile_x 1: banded, 16x16 tiles along both x faces and
 * 64x4 tiles between; tile_x 0: row runs, 256 consecutive cells of a
 * rectangle of cells of one class -- interior, boundary, next to a solid;
 * each wave of 64 cells takes the boundary paths only when one of its cells
 * needs them; persistent grids of 2..16 blocks per CU or one block per tile)
 * and both launch forms (lbm_batch_set_launch) for steps_per_trial steps
 * on the batch and keeps the fastest.  Call it before
 * lbm_batch_init_fields (it advances whatever state the buffers hold).
 * Writes "tile_x,blocks,one_launch,ms_per_step" of the choice into choice
 * (may be NULL). */
int lbm_batch_tune(lbm_batch *b, int32_t steps_per_trial, char *choice, int32_t choice_len);
/* Apply a tiling directly (for example one the autotuner chose earlier). */
int lbm_batch_set_tiling(lbm_batch *b, int32_t tile_x, int32_t blocks);
/* Launch form of a step: one_launch 1 (the default), one kernel over every
 * tile, each wave taking its class's path; 0, two kernels: the tiles whose
 * waves are all interior (compiled without the boundary code), then the
 * others. */
int lbm_batch_set_launch(lbm_batch *b, int32_t one_launch);

/* rho and u of every domain concatenated in domain order (compact,
 * x fastest); u cell-major [cell][dim]. */
int lbm_batch_init_fields(lbm_batch *b, const double *rho, const double *u);

/* Run cycles of `interval` steps (even, >= 2; the last two steps of a cycle
 * are the snapshot and the check) captured in a hipGraph until every domain
 * has converged below tol or max_steps is reached.  results[ndom] may be
 * NULL; ms receives the wall time of the run on the device stream. */
int lbm_batch_run(lbm_batch *b, int32_t interval, double tol, int64_t max_steps, lbm_domain_result *results,
                  double *ms);

/* Macroscopic fields of one domain (compact layout). */
int lbm_batch_get_macros(lbm_batch *b, int32_t dom, double *rho, double *u);

/* Post-collision populations of one domain, SoA [Q][cells], FP64. */
int lbm_batch_get_populations(lbm_batch *b, int32_t dom, double *f);

/* number of cells (sum over domains) and of tiles of the current tiling */
int lbm_batch_info(const lbm_batch *b, int64_t *ncell, int64_t *ntiles, int32_t *tile_x, int32_t *blocks);

const char *lbm_batch_last_error(const lbm_batch *b);
int lbm_batch_destroy(lbm_batch *b);

#ifdef __cplusplus
}
#endif
#endif
