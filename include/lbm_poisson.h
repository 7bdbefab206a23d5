/* lbm_poisson.h -- C ABI of the MI355X Poisson (screened Poisson /
 * diffusion-reaction) lattice-Boltzmann step, FP64, on a structured 1D or 2D
 * window.  Replaces the reference's Poisson instantiations of LBMSolver
 * (src/lbm/solverExe.h:37-80): updateMacroscopicValues' electric potential
 * (src/lbm/solver.cpp:542-545), eq::poisson (src/lbm/equilibrium_func.h:
 * 149-163), the Poisson collision with its reaction source
 * (src/lbm/solver.cpp:589-611), the push propagation (:715-740) and the
 * boundary objects LBMBnd_DirichletNEEM (src/lbm/bnd/bnd_dirichlet.h:250-394)
 * and LBMBnd_NeumannNEEM (src/lbm/bnd/bnd_neumann.h:14-66).
 *
 * Scheme: one fused pull step per call of the step kernel.  The stored state
 * is the post-collision population set f^n (structure of arrays, x fastest);
 * a step gathers f^n(x - c_i, i), which is the reference's post-propagation
 * fold(x, i); a cell on a face with a boundary object takes that object's
 * rule instead for all of its populations (the rule of the face of highest
 * priority, i.e. applied last, at a corner), then phi and the collision run
 * in registers and f^(n+1) is written once.  phi of every step is written
 * too (the Neumann rule reads the previous step's phi, as the reference's
 * vars hold it). */
#ifndef LBM_POISSON_H
#define LBM_POISSON_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lbm_poisson lbm_poisson;

/* lattices of the Poisson equation (src/lbm/constants.h:248-319) */
#define LBM_POISSON_D1Q3 3
#define LBM_POISSON_D2Q5 5
#define LBM_POISSON_D2Q9 9

typedef struct {
  int32_t lattice; /* 3, 5 or 9 */
  int32_t nx, ny;  /* structured window; ny = 1 for D1Q3; each extent >= 5 */
  double omega;    /* 1 / relaxation */
  double source;   /* dt * D * kappa^2 (collisionStep, solver.cpp:605-608): added times w'_i * phi */
} lbm_poisson_params;

/* boundary objects of the faces -x, +x, -y, +y */
#define LBM_POISSON_DIRICHLET 1 /* 'dirichlet' / 'neem': fold = feq(value) + fold(e) - feq(phi_e) */
#define LBM_POISSON_NEUMANN 2   /* 'neumann' / 'neem' (zero gradient): value = (4 phi(e) - phi(e2)) / 3 */

typedef struct {
  int32_t face;        /* 0..3: -x, +x, -y, +y (2D), 0..1 (1D) */
  int32_t kind;        /* LBM_POISSON_DIRICHLET or LBM_POISSON_NEUMANN */
  int32_t priority;    /* position in the reference's application order; a corner takes the highest */
  const double *value; /* Dirichlet: one value per cell of the face, along the face (ny values on
                          x faces, nx on y faces; 1 in 1D); Neumann: ignored */
} lbm_poisson_face;

/* Every face needs a boundary object (the reference's test cases give every
 * face one; a face without would leave its cells' missing populations as
 * the push left them, which a pull step does not keep). */
int lbm_poisson_create(const lbm_poisson_params *p, lbm_poisson **out);
int lbm_poisson_set_faces(lbm_poisson *h, const lbm_poisson_face *faces, int32_t n);

/* initialCondition (src/lbm/solver.cpp:267-295): f = fold = feq(phi0), then
 * the first step's phi and collision.  phi0[nx*ny], x fastest. */
int lbm_poisson_init(lbm_poisson *h, const double *phi0);

/* nsteps steps. */
int lbm_poisson_step(lbm_poisson *h, int32_t nsteps);

/* One step that also returns sum over cells of |phi(step) - phi(step - 1)|
 * (sumAbsDiff of the electric potential, solver.cpp:808-815), FP64 with a
 * fixed-order reduction; LBM_DIVERGED when it is not finite. */
int lbm_poisson_step_check(lbm_poisson *h, double *abs_diff);

/* phi of the last step (the reference's vars after updateMacroscopicValues). */
int lbm_poisson_get(lbm_poisson *h, double *phi);

int lbm_poisson_destroy(lbm_poisson *h);
const char *lbm_poisson_last_error(const lbm_poisson *h);

#ifdef __cplusplus
}
#endif
#endif /* LBM_POISSON_H */
