/*
 * lbm_host.h -- C ABI of the drop-in host layer (liblbm_host.so).
 *
 * Exposes the grid the host builds from a reference configuration file --
 * cell ids in the reference's order, centres, neighbour tables, boundary
 * cells and surfaces -- and a whole solver run, so that callers and tests
 * can check them without C++ types.  Replaces what the reference's
 * GridGenerator + CartesianGrid::loadGridInplace produce
 * (src/gridgenerator/gridGenerator.cpp:72-229, src/cartesiangrid.h:264-320)
 * and LBMSolver::run (src/lbm/solver.cpp:176-214).
 */
#ifndef LBM_HOST_H
#define LBM_HOST_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lbm_host_grid lbm_host_grid;

/* Generate the grid of a configuration (JSON text) and prepare it for the
 * solver section (boundary cells, surfaces, periodic links, diagonals). */
int lbm_host_grid_create(const char *json_text, lbm_host_grid **out);
int64_t lbm_host_grid_ncells(const lbm_host_grid *g);
int32_t lbm_host_grid_dim(const lbm_host_grid *g);
/* neighbour slots per cell: 8 (2D) or 26 (3D) */
int32_t lbm_host_grid_stride(const lbm_host_grid *g);
/* centres [n][dim] */
int lbm_host_grid_centers(const lbm_host_grid *g, double *out);
/* neighbour table [n][stride], -1 = none */
int lbm_host_grid_neighbors(const lbm_host_grid *g, int64_t *out);
/* boundary-cell flag [n] */
int lbm_host_grid_bndry(const lbm_host_grid *g, uint8_t *out);
/* structured window extents nxyz[3] and structured index per cell [n]
 * (single-level grids; -1 for a multi-level grid) */
int lbm_host_grid_structured(const lbm_host_grid *g, int64_t *nxyz, int64_t *index);
/* tree level and parent cell (-1 on the partition level) of every cell [n]
 * (CartesianGridBase::level / parent, src/cartesiangrid_base.h; levels
 * above the partition level come from uniformLevel > partitionLevel or
 * maxRfnmtLvl > uniformLevel) */
int lbm_host_grid_levels(const lbm_host_grid *g, int32_t *level, int64_t *parent);
/* surfaces in name order: count, then name / cell list / normal dirs of surface i */
int32_t lbm_host_grid_surface_count(const lbm_host_grid *g);
int lbm_host_grid_surface(const lbm_host_grid *g, int32_t i, char *name, int32_t name_len, int64_t *ncell,
                          int64_t *cells, int64_t *dirs);
/* Cells of the postprocessing function "line" through a[dim], b[dim]
 * (PostprocessCartesianFunctionLine::init,
 * src/postprocess/postprocessing_cartesian.h:17-42): centre within half a cell
 * length of the hyperplane through a and b, sorted by centre.  Writes at
 * most cap ids to out; returns the number of cells, or -1. */
int64_t lbm_host_grid_line_cells(const lbm_host_grid *g, const double *a, const double *b, int64_t *out, int64_t cap);
void lbm_host_grid_destroy(lbm_host_grid *g);

/* Run the solver on a configuration file like `lbm <file>`; storage may be
 * NULL (use the file's key).  The final fields are written in reference
 * cell order into rho[n] and u[n*dim] when those are non-NULL and n matches.
 * Returns 0, or the process exits with status 255 on a fatal error exactly
 * like the CLI. */
int lbm_host_run_file(const char *config_file, const char *storage, int64_t n, double *rho, double *u,
                      int64_t *steps_done);

/* Value of a boundary-value expression ("cos(pi*x)", numbers, x y z, pi,
 * + - * / ^, parentheses, sin cos tan sinh cosh tanh exp log sqrt abs asin
 * acos atan pow atan2 min max) at (x, y, z): the subset of exprtk the
 * reference's configurations use (src/common/mathexpr.h:8-40).  Returns 0,
 * or -1 with lbm_host_last_error() on a syntax error. */
int lbm_host_eval_expression(const char *text, double x, double y, double z, double *out);

/* The same for a configuration whose solver "equation" is "poisson" (the
 * reference's Poisson LBMSolver, src/lbm/solverExe.h:37-80): the electric
 * potential after the run in reference cell order into phi[n]. */
int lbm_host_run_poisson_file(const char *config_file, int64_t n, double *phi, int64_t *steps_done);

/* Write <file_base>.vtp, the reference's base64 binary point file
 * (VTK::BINARY::writePoints, src/common/IO.h:414-507, as LBMSolver::output
 * calls it, src/lbm/solver.cpp:335-381): n points with dim coordinates
 * coords[n][dim] and nfields Float64 fields names[k] / values[k][n].  Values
 * are written as given (the solver rounds them to 15 decimals first, as the
 * reference's text round trip does).  Returns 0, or -1 if n < 1 or the file
 * cannot be written. */
int lbm_host_write_points_vtk(const char *file_base, int32_t dim, int64_t n, const double *coords, int32_t nfields,
                              const char *const *names, const double *const *values);

/* value -> "%.15f" text -> double, the reference's field round trip
 * (toStringVector, include/common/util/string_helper.h:95-105) */
double lbm_host_text_round_trip(double v);

const char *lbm_host_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
