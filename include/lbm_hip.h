/*
 * lbm_hip.h -- C ABI of the MI355X lattice-Boltzmann step library
 * (liblbm_hip.so, HIP kernels for gfx950).
 *
 * This ABI replaces the per-cell passes that SFCMM/LBM's LBMSolver::timeStep
 * runs on the CPU (src/lbm/solver.cpp:307-320) together with the boundary
 * application of LBMBndManager (src/lbm/bnd/bnd.h:48-65).  Plain types only:
 * no C++ or torch types cross it, every entry point returns an int status
 * (LBM_OK = 0, negative on error) and never throws.  A handle owns its device
 * buffers and one HIP stream; it is not thread-safe.
 *
 * Cell order on this ABI is structured: index = x + nx*(y + ny*z).  The host
 * layer (lbm_amd/host) maps the reference's Hilbert-ordered cell ids onto it.
 */
#ifndef LBM_HIP_H
#define LBM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define LBM_OK 0
#define LBM_EINVAL (-1)   /* bad argument */
#define LBM_EHIP (-2)     /* HIP runtime error */
#define LBM_ENOMEM (-3)   /* device allocation failed */
#define LBM_DIVERGED (-4) /* NaN/Inf found by a checked step */

/* lattice: the model string of the reference (src/lbm/constants.h:155-157) */
#define LBM_D2Q9 9
#define LBM_D3Q19 19

/* population storage.  Every form stores f_i - w_i ("shifted"); moments,
 * equilibrium and collision run in FP64 for F32 and F64 storage and in FP32
 * for F16/BF16 storage. */
#define LBM_STORE_F32 0
#define LBM_STORE_F64 1
#define LBM_STORE_F16 2
#define LBM_STORE_BF16 3

/* face kinds for the six faces -x,+x,-y,+y,-z,+z (direction ids of the
 * reference, include/common/constants.h:104-125) */
#define LBM_FACE_PERIODIC 0    /* 'periodic', generateBndry=false (src/cartesiangrid.h:645-706) */
#define LBM_FACE_WALL 1        /* 'wall'/'bounceback' (+tangentialVelocity), 'dirichlet'/'bounceback'
                                  (src/lbm/bnd/bnd_dirichlet.h:79-121) */
#define LBM_FACE_PRESSURE_ABB 2 /* 'pressure' anti-bounce-back (src/lbm/bnd/bnd_pressure.h:55-106) */
#define LBM_FACE_VELOCITY_ZH 3 /* Zou-He velocity inlet (no reference counterpart) */
#define LBM_FACE_PRESSURE_ZH 4 /* Zou-He pressure outlet (no reference counterpart) */
#define LBM_FACE_PRESSURE_NEEM 5 /* non-equilibrium extrapolation pressure face (Guo, Zheng & Shi 2002):
                                    every population of a face cell from its inward neighbour (no
                                    reference counterpart; the faces next to it must be walls or
                                    periodic) */
#define LBM_FACE_VELOCITY_NEEM 6 /* non-equilibrium extrapolation velocity face, density from the inward
                                    neighbour (no reference counterpart; same neighbour rule) */
#define LBM_FACE_PERIODIC_PRESSURE 7 /* 'periodic' with 'pressure' (src/lbm/bnd/bnd_periodic.h:102-119): a
                                    population entering across the face is feq(p', u_c) + f_post(c) -
                                    feq(rho_c, u_c) of the cell c straight across on the opposite face,
                                    p' = rho of the opposite face; both faces of the axis carry it, and
                                    it takes the lowest priority (walls win at corners, as the
                                    reference's wall objects overwrite after its preApply) */
/* Wet-node walls: every population of a face cell is rebuilt after streaming
 * from the density of the populations known there (LBMBnd_wallWetnode,
 * src/lbm/bnd/bnd_wetnode.h:20-55; calcDensity_limited,
 * src/lbm/moments.h:120-153) and the wall velocity `value`; populations whose
 * upstream face has a higher priority keep that face's rule. */
#define LBM_FACE_WALL_EQ 8   /* 'wall'/'equilibrium': feq(rho, u_w) (src/lbm/bnd/bnd_wall.h:100-185) */
#define LBM_FACE_WALL_NEEM 9 /* 'wall'/'neem': feq(rho, u_w) + f(n1) - feq(rho1, u1), n1 the inward
                                neighbour (src/lbm/bnd/bnd_wall.h:187-300); at least 3 cells along the
                                normal, no wet-node, Zou-He or NEEM face next to it */
#define LBM_FACE_WALL_NEBB 10 /* 'wall'/'nebb' with 'velocity', D2Q9 only: normal population
                                 reflected, diagonals from the known ones (apply_constV,
                                 src/lbm/bnd/bnd_wall.h:404-472); the velocity must be tangential */
#define LBM_FACE_WALL_NEBB0 11 /* 'wall'/'nebb' without 'velocity', D2Q9 only: the reference's
                                  apply_0 (src/lbm/bnd/bnd_wall.h:353-400): main populations
                                  reflected wherever a neighbour is missing, then its diagonal
                                  formulas, which read the two diagonals that enter through the
                                  wall; no rule writes those, so they hold their initial value
                                  (taken as the equilibrium at rho 1, u 0) or, on a periodic-
                                  pressure face, the value that face supplies */

/* cell flags (optional byte per cell) */
#define LBM_CELL_FLUID 0
#define LBM_CELL_SOLID 1 /* bounce-back obstacle; its populations are never read or written */

typedef struct lbm_handle lbm_handle;

/* Structured lattice.  Replaces the grid handed over by
 * LBMSolver::transferGrid (src/lbm/solver.cpp:759-801) for single-level
 * grids.  solid: optional nx*ny*nz bytes (LBM_CELL_*), copied. */
typedef struct {
  int32_t dim;         /* 2 or 3 */
  int32_t nx, ny, nz;  /* nz = 1 in 2D */
  const uint8_t *solid;
} lbm_grid_desc;

/* Solver parameters.  omega = 1/tau as set from 'relaxation'
 * (src/lbm/solver.cpp:108-112).  force is a constant Guo body force (new). */
typedef struct {
  int32_t lattice; /* LBM_D2Q9 or LBM_D3Q19 */
  int32_t storage; /* LBM_STORE_* */
  double omega;
  double force[3];
} lbm_params;

/* One boundary face.  Replaces the per-surface LBMBndInterface objects that
 * LBMBndManager::setupBndryCnds builds (src/lbm/bnd/bnd.h:71-142).
 * value: wall/inlet velocity; rho: pressure (density) for pressure faces.
 * priority: a population that leaves the box across two faces (an edge or
 * corner cell's diagonal) takes the rule of the face with the higher
 * priority, ties going to the lower face index.  The reference applies its
 * boundary objects in order and the last write wins (src/lbm/bnd/bnd.h:60-65),
 * so a host passes the application order here. */
typedef struct {
  int32_t face; /* 0..5 = -x,+x,-y,+y,-z,+z */
  int32_t kind; /* LBM_FACE_* */
  double value[3];
  double rho;
  int32_t priority;
} lbm_bc_desc;

/* Reductions in FP64, summed in a fixed order (deterministic).
 * abs_diff follows the reference's variable order u,v,(w),rho and is its
 * convergence measure sum|q^n - q^(n-1)| (src/lbm/solver.cpp:233-263,808-815);
 * res_rel = sqrt(diff_sq / norm_sq) is the relative L2 residual. */
typedef struct {
  double mass;
  double momentum[3];
  double abs_diff[4];
  double diff_sq;
  double norm_sq;
  double res_rel;
  double nonfinite;
  double fluid_cells;
} lbm_sums;

/* Create a solver.  Allocates two population buffers (A-B pull scheme). */
int lbm_create(const lbm_grid_desc *grid, const lbm_params *params, lbm_handle **out);

/* Set boundary faces; every non-periodic face must be described once.
 * Faces not listed are periodic. */
int lbm_set_bc(lbm_handle *h, const lbm_bc_desc *bcs, int32_t n);

/* Initial condition of src/lbm/solver.cpp:267-295: f = feq(rho, u), then the
 * first collision.  rho[n], u[n*dim] (cell-major, structured order). */
int lbm_init_fields(lbm_handle *h, const double *rho, const double *u);

/* Advance nsteps time steps (LBMSolver::timeStep, src/lbm/solver.cpp:307-320). */
int lbm_step(lbm_handle *h, int32_t nsteps);

/* Advance two steps; the second reports the one-step change of the
 * macroscopic fields plus mass and momentum (convergenceCondition,
 * src/lbm/solver.cpp:233-263).  Returns LBM_DIVERGED if NaN/Inf appear. */
int lbm_step_check(lbm_handle *h, lbm_sums *out);

/* One step with an explicit reduction mode:
 *   0 plain, 1 snapshot (store u, rho), 2 check (one-step change against the
 *   last snapshot, sums in *out), 3 check and snapshot.
 * lbm_init_fields stores the snapshot of the initial fields.  This lets a
 * host reproduce the reference's schedule, where the check at step t
 * compares the fields of steps t-1 and t-2 (src/lbm/solver.cpp:189-197). */
#define LBM_STEP_PLAIN 0
#define LBM_STEP_SNAPSHOT 1
#define LBM_STEP_CHECK 2
#define LBM_STEP_CHECK_SNAPSHOT 3
int lbm_step_ex(lbm_handle *h, int32_t mode, lbm_sums *out);

/* lbm_step_ex without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.  The FP64
 * per-block partial sums and the reduced sums of a check step live in one
 * library-wide device array per GPU (LBM_NOALLOC_MAX_BLOCKS blocks of 256
 * cells per launch, about 12.6 MB, held until the process ends) instead of
 * the handle's buffers; a launch with more blocks is rejected with
 * LBM_EINVAL.  The snapshot buffers must exist (lbm_init_fields creates
 * them); the call never creates them.  Results equal lbm_step_ex's bit for
 * bit.  All calls on one GPU share the array: calls that overlap in time on
 * one GPU (two host threads, or two handles' streams) race. */
#define LBM_NOALLOC_MAX_BLOCKS 131072
int lbm_step_ex_noalloc(lbm_handle *h, int32_t mode, lbm_sums *out);

/* Macroscopic fields after the last step (updateMacroscopicValues,
 * src/lbm/solver.cpp:513-553): rho[n], u[n*dim], cell-major. */
int lbm_get_macros(lbm_handle *h, double *rho, double *u);

/* Mass and momentum of the current state (diff fields are zero). */
int lbm_reduce(lbm_handle *h, lbm_sums *out);

/* Post-collision populations f_i (unshifted), SoA [Q][n], FP64. */
int lbm_get_populations(lbm_handle *h, double *f);
int lbm_set_populations(lbm_handle *h, const double *f);

/* Change the population storage (LBM_STORE_*) keeping the state: the
 * populations are converted on the device (exact when widening, round to
 * nearest even when narrowing).  A residual check needs a snapshot step
 * after the change (lbm_step_check and LBM_STEP_SNAPSHOT take one).  No
 * reference counterpart: used to run the bulk of a solve in 16-bit storage
 * and finish in FP32. */
int lbm_set_storage(lbm_handle *h, int32_t storage);

/* Time nsteps steps on the handle's stream with HIP events (ms total). */
int lbm_time_steps(lbm_handle *h, int32_t nsteps, double *ms);

/* Time ngroups consecutive groups of steps (steps[g] steps each) as one
 * uninterrupted run on the handle's stream: an event is recorded between
 * groups, the host waits once at the end; ms[g] = time of group g.  Used to
 * report the median of equal repetitions inside one timed run. */
int lbm_time_step_groups(lbm_handle *h, const int32_t *steps, int32_t ngroups, double *ms);

/* Launch shape of the step kernel: block_x threads along x (multiple of 64),
 * block_y rows (0 keeps the current value); nontemporal loads/stores on (1)
 * or off (0), -1 keeps the current setting. */
int lbm_set_launch(lbm_handle *h, int32_t block_x, int32_t block_y, int32_t nontemporal);

/* How the step kernel gathers the populations with c_x != 0: 0 (default)
 * straight from global memory, neighbours' cache lines reused through L2;
 * 1 through an LDS stage per 256-cell x row (aligned row loads plus two
 * halo cells per row and population, then shifted reads from LDS).  Same
 * results bit for bit.  1 needs FP32 storage, 32-bit offsets and 256x1
 * blocks (x extent >= 256); it is dropped again when lbm_set_storage
 * leaves FP32, and lbm_set_launch / lbm_set_index_width shapes other than
 * those fall back to 0.  Replaces the push of propagationStep,
 * src/lbm/solver.cpp:715-740. */
int lbm_set_halo(lbm_handle *h, int32_t lds);

/* LBMSolver::forcing (src/lbm/solver.cpp:626-696, configuration key
 * "forcing"), D2Q9: after every collision (and after the first one in
 * lbm_init_fields) the cells of the x = 0 face take
 * f_i = feq_i(p_inlet, u_v) + f_i(v) - feq_i(rho_v, u_v) from the cell v next
 * to the x = nx-1 face in their row, and the cells of the x = nx-1 face the
 * same with p_outlet from the cell next to the x = 0 face; feq_i(p, u_v)
 * takes only u_v's x component into c_i.u (as the reference does).  The
 * pushed values then stream as the step's own.  on = 0 turns it off. */
int lbm_set_pressure_forcing(lbm_handle *h, int32_t on, double p_inlet, double p_outlet);

/* Width of the population offsets of the step kernel: 32 (default whenever a
 * population plane fits in 4 GiB; one 32-bit offset per access) or 64.
 * Both give identical results; 32 is rejected for planes above 4 GiB. */
int lbm_set_index_width(lbm_handle *h, int32_t bits);

/* Size facts: number of cells, populations per cell, bytes per stored
 * population, and the name of the step kernel symbol family. */
int lbm_info(const lbm_handle *h, int64_t *ncell, int32_t *q, int32_t *bytes_per_pop);

int lbm_synchronize(lbm_handle *h);
int lbm_destroy(lbm_handle *h);

/* Last error message of a handle, or of the calling thread when h is NULL. */
const char *lbm_last_error(const lbm_handle *h);

/* Device copy kernel over n bytes (the bandwidth ceiling used to calibrate
 * the roofline counters); ms = average time per launch over reps. */
int lbm_copy_bench(int64_t nbytes, int32_t reps, double *ms);

#ifdef __cplusplus
}
#endif
#endif /* LBM_HIP_H */
