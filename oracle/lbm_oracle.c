/*
 * FP64 CPU oracle for the lattice-Boltzmann collide/stream/boundary step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates, in plain C, the algorithm of
 * SFCMM/LBM's src/lbm hot path so that the HIP product path can be checked
 * against it.  Only the test suite, the smoke check and the CPU-baseline leg
 * of the benchmark load it; the product library never links or calls it.
 *
 * What is restated (reference file:line):
 *   - lattice tables D2Q9 / D3Q19 ............ src/lbm/constants.h:295-365
 *   - equilibrium, "defaultEq" grouping ........ src/lbm/equilibrium_func.h:52-84
 *   - symmetric equilibrium "symmEq" ........... src/lbm/equilibrium_func.h:109-147
 *   - macroscopic values ....................... src/lbm/solver.cpp:513-553
 *   - collision f=(1-w)fold+w feq .............. src/lbm/solver.cpp:574-622
 *   - push propagation through neighbour table . src/lbm/solver.cpp:715-740
 *   - halfway bounce-back (+Ladd wall velocity)  src/lbm/bnd/bnd_dirichlet.h:79-121
 *   - anti-bounce-back pressure ................ src/lbm/bnd/bnd_pressure.h:55-106
 *   - periodic pair with a density jump ........ src/lbm/bnd/bnd_periodic.h:102-108
 *   - wet-node walls (equilibrium, NEEM, NEBB) . src/lbm/bnd/bnd_wall.h:100-479,
 *     limited density ........................... bnd_wetnode.h:20-55, moments.h:120-153
 *   - residual sum |q^n - q^(n-1)| ............. src/lbm/solver.cpp:233-263,808-815
 *   - Hilbert index ............................ include/common/math/hilbert.h:15-48
 * New components with no reference counterpart (pinned by closed forms and
 * invariants in the tests): Guo body force, Zou-He velocity/pressure faces
 * (Hecht & Harting 2010 form with transverse-momentum correction), the
 * relative L2 residual ||u^n-u^(n-1)||/||u^n||, and shifted storage
 * quantisation helpers (fp16/bf16 round-to-nearest-even).
 *
 * Gate rule of the reference's bounce-back: a population is bounced when the
 * neighbour is missing AND normal.c_i >= eps (bnd_dirichlet.h:85-86).  The
 * caller encodes that gate in each boundary record's direction mask, so this
 * file applies exactly the directions it is told to.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------- tables */
/* src/lbm/constants.h:295-319: rest population last, opposite map. */
static const double D2Q9_C[9][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {1, 1, 0},
                                    {1, -1, 0}, {-1, -1, 0}, {-1, 1, 0}, {0, 0, 0}};
static const int D2Q9_OPP[9] = {1, 0, 3, 2, 6, 7, 4, 5, 8};
static const double D2Q9_W[9] = {1.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0, 1.0 / 36.0,
                                 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 4.0 / 9.0};
/* src/lbm/constants.h:321-365 */
static const double D3Q19_C[19][3] = {{-1, 0, 0},  {1, 0, 0},   {0, -1, 0},  {0, 1, 0},  {0, 0, -1},
                                      {0, 0, 1},   {-1, -1, 0}, {-1, 1, 0},  {1, -1, 0}, {1, 1, 0},
                                      {-1, 0, -1}, {-1, 0, 1},  {1, 0, -1},  {1, 0, 1},  {0, -1, -1},
                                      {0, -1, 1},  {0, 1, -1},  {0, 1, 1},   {0, 0, 0}};
static const int D3Q19_OPP[19] = {1, 0, 3, 2, 5, 4, 9, 8, 7, 6, 13, 12, 11, 10, 17, 16, 15, 14, 18};
static const double D3Q19_W[19] = {1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0,
                                   1.0 / 18.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                                   1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                                   1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 3.0};

/* lbm_cssq, src/lbm/constants.h:27 */
#define CSSQ (1.0 / 3.0)

typedef struct {
  int Q, dim;
  const double (*c)[3];
  const double *w;
  const int *opp;
} lattice_t;

static int get_lattice(int Q, lattice_t *L) {
  if (Q == 9) {
    L->Q = 9; L->dim = 2; L->c = D2Q9_C; L->w = D2Q9_W; L->opp = D2Q9_OPP;
    return 0;
  }
  if (Q == 19) {
    L->Q = 19; L->dim = 3; L->c = D3Q19_C; L->w = D3Q19_W; L->opp = D3Q19_OPP;
    return 0;
  }
  return -1;
}

/* Export the tables so the tests can pin them against closed-form invariants. */
int orc_tables(int Q, double *c_out /*[Q][3]*/, double *w_out, int *opp_out) {
  lattice_t L;
  if (get_lattice(Q, &L)) return -1;
  for (int i = 0; i < Q; ++i) {
    for (int d = 0; d < 3; ++d) c_out[3 * i + d] = L.c[i][d];
    w_out[i] = L.w[i];
    opp_out[i] = L.opp[i];
  }
  return 0;
}

/* eq::defaultEq, src/lbm/equilibrium_func.h:52-54 */
static inline double default_eq(double w, double rho, double cu, double vsq) {
  return w * rho * (1.0 + cu / CSSQ + cu * cu / (2.0 * CSSQ * CSSQ) - vsq / (2.0 * CSSQ));
}
/* eq::symmEq, src/lbm/equilibrium_func.h:109-111 */
static inline double symm_eq(double w, double rho, double cu, double vsq) {
  return w * rho * (1.0 + cu * cu / (2.0 * CSSQ * CSSQ) - vsq / (2.0 * CSSQ));
}

/* Array form, src/lbm/equilibrium_func.h:67-84: vsq summed over dims first,
 * then cu per population in dimension order. */
static inline void feq_all(const lattice_t *L, double rho, const double *u, double *feq) {
  double vsq = 0;
  for (int d = 0; d < L->dim; ++d) vsq += u[d] * u[d];
  for (int i = 0; i < L->Q; ++i) {
    double cu = 0;
    for (int d = 0; d < L->dim; ++d) cu += u[d] * L->c[i][d];
    feq[i] = default_eq(L->w[i], rho, cu, vsq);
  }
}

int orc_feq(int Q, double rho, const double *u, double *feq) {
  lattice_t L;
  if (get_lattice(Q, &L)) return -1;
  feq_all(&L, rho, u, feq);
  return 0;
}

/* ------------------------------------------------------------- macros */
/* updateMacroscopicValues, src/lbm/solver.cpp:513-553.  rho is the sum over
 * all Q populations in distribution order (std::accumulate); u_d sums the
 * first Q-1 populations (the rest population has c=0) and divides by rho.
 * With a body force (new) the reported velocity is Guo's u* = u + F/(2 rho).
 * vars layout per cell: u_0..u_{dim-1}, rho (src/lbm/variables.h:74-79). */
static inline void macros_cell(const lattice_t *L, const double *f, const double *force, double *v) {
  double rho = 0;
  for (int i = 0; i < L->Q; ++i) rho += f[i];
  for (int d = 0; d < L->dim; ++d) {
    double m = 0;
    for (int i = 0; i < L->Q - 1; ++i) m += L->c[i][d] * f[i];
    m /= rho;
    if (force) m += 0.5 * force[d] / rho;
    v[d] = m;
  }
  v[L->dim] = rho;
}

int orc_macros(int Q, int64_t ncell, const double *fold, const double *force, double *vars) {
  lattice_t L;
  if (get_lattice(Q, &L)) return -1;
  const int nv = L.dim + 1;
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < ncell; ++c) macros_cell(&L, fold + c * Q, force, vars + c * nv);
  return 0;
}

/* ------------------------------------------------------ boundary records */
/* One record per boundary cell.  kind:
 *   1 BB   halfway bounce-back, optional Ladd wall velocity u (rho0 = 1),
 *          bnd_dirichlet.h:85-116
 *   2 ABB  anti-bounce-back constant pressure rho, velocity extrapolated
 *          1.5*u(n1) - 0.5*u(n2), bnd_pressure.h:55-106
 *   3 ZHV  Zou-He velocity face (new): outward normal axis/sign, velocity u
 *   4 ZHP  Zou-He pressure face (new): outward normal axis/sign, density rho
 *   5 NEEM non-equilibrium extrapolation pressure face (new): all
 *          populations from the inward neighbour n1, density rho
 *   6 NEEM non-equilibrium extrapolation velocity face (new), velocity u
 *   7 PP   periodic face with a density jump, LBMBndCell_periodic::preApply
 *          (bnd_periodic.h:102-108): fold(c, o) = feq_o(rho, u_n1) +
 *          fpost(n1, o) - feq_o(vars(n1)) for the missing directions, n1 the
 *          linked cell on the other face, rho the pressure of n1's surface
 *   8 EQW  wet-node equilibrium wall, LBMBnd_wallEq (bnd_wall.h:100-185)
 *          through LBMBnd_DirichletEQ::apply (bnd_dirichlet.h:203-236)
 *   9 NEEMW wet-node non-equilibrium extrapolation wall, LBMBnd_wallNEEM
 *          (bnd_wall.h:187-300): EQW, then + fold(n1) - feq(moments of n1)
 *  10 NEBBW wet-node non-equilibrium bounce-back wall, D2Q9, LBMBnd_wallNEBB
 *          apply_constV (bnd_wall.h:404-472)
 * dirmask bit i: apply to lattice direction i (the direction of the missing
 * neighbour), only consulted by BB, ABB and PP.
 * Wet-node walls: lim1 / lim2 bit i = population i enters the limited
 * density with constant 1 / 2 (LBMBnd_wallWetnode, bnd_wetnode.h:20-55; the
 * rest population is a lim1 bit); flags bit 0: the wall has a velocity
 * (apply_constV), bit 1: empty limited set (a corner, density unchanged),
 * bit 2: leave vars untouched (the reference writes the wall's velocity and
 * density there, which later records of the same step may read);
 * axis/sign: the outward normal. */
typedef struct {
  int64_t cell;
  int32_t kind;
  int32_t dirmask;
  int32_t axis;
  int32_t sign;
  int64_t n1, n2;
  double u[3];
  double rho;
  int32_t lim1, lim2;
  int32_t flags, pad;
} orc_bc;

/* calcDensity_limited (src/lbm/moments.h:120-153): rho = sum over the
 * limited set (ascending distribution order) of const_i fold(c, i); with a
 * wall velocity rho *= 1 / (1 + n.u) along the normal, u the velocity the
 * cell holds in vars at that moment.  An empty set leaves vars untouched. */
static void wet_density(const lattice_t *L, const orc_bc *r, const double *f, double *v) {
  if (r->flags & 2) return;
  double rho = 0;
  for (int i = 0; i < L->Q; ++i) {
    if ((r->lim2 >> i) & 1) rho += 2.0 * f[i];
    else if ((r->lim1 >> i) & 1) rho += 1.0 * f[i];
  }
  if (r->flags & 1) rho *= 1.0 / (1.0 + (double)r->sign * v[r->axis]);
  v[L->dim] = rho;
}

/* Zou-He with transverse-momentum correction (Hecht & Harting, J. Stat.
 * Mech. 2010 P01018), written for any face normal of D2Q9/D3Q19:
 *   unknown i (c_i . n_out < 0):  f_i = f_opp(i) + 6 w_i rho (c_i . u)
 *                                         - sum_t c_it N_t
 *   N_t = 1/2 sum_{c_in = 0} c_jt f_j - rho u_t / 3
 * rho (velocity face) or u_n (pressure face) from the mass balance
 *   rho (1 + s u_n) = sum_{c_n=0} f + 2 sum_{c.n_out>0} f, s = outward sign. */
static void zou_he_cell(const lattice_t *L, double *f, int axis, int sign, int is_velocity, const double *uin,
                        double rho_in) {
  const int Q = L->Q;
  double s0 = 0, sout = 0;
  for (int i = 0; i < Q; ++i) {
    const double cn = L->c[i][axis] * sign;
    if (cn == 0) s0 += f[i];
    else if (cn > 0) sout += f[i];
  }
  double u[3] = {0, 0, 0};
  double rho;
  if (is_velocity) {
    for (int d = 0; d < 3; ++d) u[d] = uin[d];
    const double un = u[axis] * sign; /* outward component */
    rho = (s0 + 2.0 * sout) / (1.0 + un);
  } else {
    rho = rho_in;
    const double un = (s0 + 2.0 * sout) / rho - 1.0;
    u[axis] = un * sign;
  }
  double N[3] = {0, 0, 0};
  for (int t = 0; t < L->dim; ++t) {
    if (t == axis) continue;
    double m = 0;
    for (int i = 0; i < Q; ++i)
      if (L->c[i][axis] == 0) m += L->c[i][t] * f[i];
    N[t] = 0.5 * m - rho * u[t] / 3.0;
  }
  for (int i = 0; i < Q; ++i) {
    const double cn = L->c[i][axis] * sign;
    if (!(cn < 0)) continue;
    double cu = 0, corr = 0;
    for (int d = 0; d < L->dim; ++d) cu += L->c[i][d] * u[d];
    for (int t = 0; t < L->dim; ++t)
      if (t != axis) corr += L->c[i][t] * N[t];
    f[i] = f[L->opp[i]] + 6.0 * L->w[i] * rho * cu - corr;
  }
}

/* ----------------------------------------------------------------- step */
/* One time step in the reference's order (src/lbm/solver.cpp:307-320):
 *   vars <- macros(fold); feq(vars); f = (1-w) fold + w feq (+ Guo force);
 *   fold(nbr(c,i), i) = f(c,i) for i < Q-1, fold(c,Q-1) = f(c,Q-1);
 *   boundary records in list order.
 * nbr[c*(Q-1)+i] is the downstream neighbour of cell c in lattice direction
 * i, or -1.  fold is updated in place; fpost is caller scratch [ncell*Q];
 * vars receives macros(fold) at the start of the step (what the reference
 * holds in m_vars after this step).
 */
static int step_impl(int Q, int64_t ncell, const int64_t *nbr, double omega, const double *force, double *fold,
                     double *fpost, double *vars, int64_t nbc, const orc_bc *bcs, int64_t nforce, const int64_t *fcell,
                     const int64_t *vcell, const double *fpres);

int orc_step(int Q, int64_t ncell, const int64_t *nbr, double omega, const double *force, double *fold,
             double *fpost, double *vars, int64_t nbc, const orc_bc *bcs) {
  return step_impl(Q, ncell, nbr, omega, force, fold, fpost, vars, nbc, bcs, 0, NULL, NULL, NULL);
}

/* The same step with LBMSolver::forcing (src/lbm/solver.cpp:626-696)
 * between the collision and the propagation: cell fcell[k] takes
 * f_i = feq_i(fpres[k], u_v) + f_i(v) - feq_i(rho_v, u_v) from its value
 * cell v = vcell[k] (post-collision f, the step's macros and equilibrium of
 * v), where feq_i(p, u_v) puts only u_v's x component into c_i.u and all
 * of u_v into u.u (the reference's cu = velocity(v, 0) * dirs[i][0]). */
int orc_step_forced(int Q, int64_t ncell, const int64_t *nbr, double omega, const double *force, double *fold,
                    double *fpost, double *vars, int64_t nbc, const orc_bc *bcs, int64_t nforce, const int64_t *fcell,
                    const int64_t *vcell, const double *fpres) {
  return step_impl(Q, ncell, nbr, omega, force, fold, fpost, vars, nbc, bcs, nforce, fcell, vcell, fpres);
}

static int step_impl(int Q, int64_t ncell, const int64_t *nbr, double omega, const double *force, double *fold,
                     double *fpost, double *vars, int64_t nbc, const orc_bc *bcs, int64_t nforce, const int64_t *fcell,
                     const int64_t *vcell, const double *fpres) {
  lattice_t L;
  if (get_lattice(Q, &L)) return -1;
  const int nv = L.dim + 1;
  const int has_force = force && (force[0] != 0 || force[1] != 0 || force[2] != 0);

#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < ncell; ++c) {
    const double *fo = fold + c * Q;
    double *v = vars + c * nv;
    macros_cell(&L, fo, has_force ? force : NULL, v);
    double feq[19];
    feq_all(&L, v[L.dim], v, feq);
    double *fp = fpost + c * Q;
    for (int i = 0; i < Q; ++i) {
      fp[i] = (1 - omega) * fo[i] + omega * feq[i];
      if (has_force) {
        /* Guo, Zheng & Shi, PRE 65 046308 (2002):
         * (1 - w/2) w_i [ (c_i - u)/cs^2 + (c_i.u) c_i / cs^4 ] . F */
        double cu = 0;
        for (int d = 0; d < L.dim; ++d) cu += L.c[i][d] * v[d];
        double s = 0;
        for (int d = 0; d < L.dim; ++d) s += ((L.c[i][d] - v[d]) / CSSQ + cu * L.c[i][d] / (CSSQ * CSSQ)) * force[d];
        fp[i] += (1 - 0.5 * omega) * L.w[i] * s;
      }
    }
  }

  /* forcing: the face cells from their value cells, in list order */
  for (int64_t k = 0; k < nforce; ++k) {
    const int64_t c = fcell[k], v = vcell[k];
    const double *vv = vars + v * nv;
    double vsq = 0;
    for (int d = 0; d < L.dim; ++d) vsq += vv[d] * vv[d];
    double feqv[19];
    feq_all(&L, vv[L.dim], vv, feqv);
    for (int i = 0; i < Q; ++i) {
      const double cu = vv[0] * L.c[i][0];
      fpost[c * Q + i] = default_eq(L.w[i], fpres[k], cu, vsq) + fpost[v * Q + i] - feqv[i];
    }
  }

  /* propagation, src/lbm/solver.cpp:727-739 */
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < ncell; ++c) {
    for (int i = 0; i < Q - 1; ++i) {
      const int64_t n = nbr[c * (Q - 1) + i];
      if (n >= 0) fold[n * Q + i] = fpost[c * Q + i];
    }
    fold[c * Q + Q - 1] = fpost[c * Q + Q - 1];
  }

  /* boundaries, applied in list order like LBMBndManager::apply (bnd.h:60-65) */
  for (int64_t b = 0; b < nbc; ++b) {
    const orc_bc *r = bcs + b;
    const int64_t c = r->cell;
    if (r->kind == 1 || r->kind == 2) {
      double uext[3] = {0, 0, 0};
      double vsq = 0;
      if (r->kind == 2) {
        for (int d = 0; d < L.dim; ++d) {
          uext[d] = 1.5 * vars[r->n1 * nv + d] - 0.5 * vars[r->n2 * nv + d];
          vsq += uext[d] * uext[d];
        }
      }
      for (int i = 0; i < Q - 1; ++i) {
        if (!((r->dirmask >> i) & 1)) continue;
        if (nbr[c * (Q - 1) + i] >= 0) continue;
        const int o = L.opp[i];
        if (r->kind == 1) {
          double val = fpost[c * Q + i];
          /* += rho0 * 2/cs^2 * w_opp * c_opp . u_wall, rho0 = 1 */
          for (int d = 0; d < L.dim; ++d) val += 1.0 * 2.0 / CSSQ * L.w[o] * L.c[o][d] * r->u[d];
          fold[c * Q + o] = val;
        } else {
          double cu = 0;
          for (int d = 0; d < L.dim; ++d) cu += uext[d] * L.c[i][d];
          fold[c * Q + o] = -fpost[c * Q + i] + 2 * symm_eq(L.w[i], r->rho, cu, vsq);
        }
      }
    } else if (r->kind == 3 || r->kind == 4) {
      zou_he_cell(&L, fold + c * Q, r->axis, r->sign, r->kind == 3, r->u, r->rho);
    } else if (r->kind == 7) {
      const double *v1 = vars + r->n1 * nv;
      double vsq = 0;
      for (int d = 0; d < L.dim; ++d) vsq += v1[d] * v1[d];
      for (int i = 0; i < Q - 1; ++i) {
        if (!((r->dirmask >> i) & 1)) continue;
        if (nbr[c * (Q - 1) + i] >= 0) continue;
        const int o = L.opp[i];
        double cu = 0;
        for (int d = 0; d < L.dim; ++d) cu += L.c[o][d] * v1[d];
        fold[c * Q + o] = default_eq(L.w[o], r->rho, cu, vsq) + fpost[r->n1 * Q + o] -
                          default_eq(L.w[o], v1[L.dim], cu, vsq);
      }
    } else if (r->kind == 8 || r->kind == 9) {
      /* DirichletEQ::apply: the wall velocity into vars, the limited
       * density, then every population = feq(rho, u_w) (w_i rho without a
       * velocity, VALZERO) */
      double vloc[4];
      double *v = vars + c * nv;
      if (r->flags & 4) {
        for (int k = 0; k < nv; ++k) vloc[k] = v[k];
        v = vloc;
      }
      for (int d = 0; d < L.dim; ++d) v[d] = r->u[d];
      wet_density(&L, r, fold + c * Q, v);
      const double rho = v[L.dim];
      if (r->flags & 1) {
        feq_all(&L, rho, r->u, fold + c * Q);
      } else {
        for (int i = 0; i < Q; ++i) fold[c * Q + i] = default_eq(L.w[i], rho, 0.0, 0.0);
      }
      if (r->kind == 9) {
        /* apply_general: moments of the extrapolation cell n1 (calcDensity /
         * calcVelocity, moments.h:17-101: no force term), then
         * fold(c, i) += fold(n1, i) - feq_i(rho1, u1).  The reference
         * also stores rho1, u1 in vars(n1); later objects of the same
         * step would read them there.  This restatement keeps them local. */
        const double *f1 = fold + r->n1 * Q;
        double rho1 = f1[0];
        for (int i = 1; i < Q; ++i) rho1 += f1[i];
        double u1[3] = {0, 0, 0};
        for (int d = 0; d < L.dim; ++d) {
          for (int i = 0; i < Q - 1; ++i) u1[d] += L.c[i][d] * f1[i];
          u1[d] /= rho1;
        }
        double vsq = 0;
        for (int d = 0; d < L.dim; ++d) vsq += u1[d] * u1[d];
        for (int i = 0; i < Q; ++i) {
          double cu = 0;
          for (int d = 0; d < L.dim; ++d) cu += u1[d] * L.c[i][d];
          fold[c * Q + i] += f1[i] - default_eq(L.w[i], rho1, cu, vsq);
        }
      }
    } else if (r->kind == 10) {
      /* NEBB (D2Q9): limited density with the velocity the cell holds, then
       * the wall velocity into vars; the normal population reflected with
       * -/+ 2/3 rho u_n; the two unknown diagonals from the known ones,
       * the tangential difference and the wall velocity (Zou & He 1997). */
      if (Q != 9) return -3;
      double vloc[4];
      double *v = vars + c * nv;
      if (r->flags & 4) {
        for (int k = 0; k < nv; ++k) vloc[k] = v[k];
        v = vloc;
      }
      double *f = fold + c * Q;
      wet_density(&L, r, f, v);
      for (int d = 0; d < L.dim; ++d) v[d] = r->u[d];
      const double rho = v[L.dim];
      const double *uw = r->u;
      const int out = 2 * r->axis + (r->sign > 0 ? 1 : 0); /* outward main direction */
      if (nbr[c * (Q - 1) + out] < 0) {
        f[L.opp[out]] = f[out];
        if (r->sign < 0) f[L.opp[out]] -= 2.0 / 3.0 * rho * uw[r->axis];
        else f[L.opp[out]] += 2.0 / 3.0 * rho * uw[r->axis];
      }
      if (r->axis == 0 && r->sign > 0) {
        f[6] = f[4] + 0.5 * (f[3] - f[2]) - 0.5 * rho * uw[1] - 1.0 / 6.0 * rho * uw[0];
        f[7] = f[5] - 0.5 * (f[3] - f[2]) + 0.5 * rho * uw[1] - 1.0 / 6.0 * rho * uw[0];
      } else if (r->axis == 0) {
        f[4] = f[6] - 0.5 * (f[3] - f[2]) + 0.5 * rho * uw[1] + 1.0 / 6.0 * rho * uw[0];
        f[5] = f[7] + 0.5 * (f[3] - f[2]) - 0.5 * rho * uw[1] + 1.0 / 6.0 * rho * uw[0];
      } else if (r->sign > 0) {
        f[5] = f[7] - 0.5 * (f[1] - f[0]) + 0.5 * rho * uw[0] - 1.0 / 6.0 * rho * uw[1];
        f[6] = f[4] + 0.5 * (f[1] - f[0]) - 0.5 * rho * uw[0] - 1.0 / 6.0 * rho * uw[1];
      } else {
        f[7] = f[5] + 0.5 * (f[1] - f[0]) - 0.5 * rho * uw[0] + 1.0 / 6.0 * rho * uw[1];
        f[4] = f[6] - 0.5 * (f[1] - f[0]) + 0.5 * rho * uw[0] + 1.0 / 6.0 * rho * uw[1];
      }
    } else if (r->kind == 11) {
      /* NEBB without a wall velocity, LBMBnd_wallNEBB::apply_0
       * (bnd_wall.h:353-400): zero velocity, limited density, every main
       * population whose neighbour is missing reflected, then the
       * diagonal formulas of the branch its outward normal selects; the
       * populations they read hold whatever fold holds (values no rule has
       * written this step keep older ones). */
      if (Q != 9) return -3;
      double vloc[4];
      double *v = vars + c * nv;
      if (r->flags & 4) {
        for (int k = 0; k < nv; ++k) vloc[k] = v[k];
        v = vloc;
      }
      double *f = fold + c * Q;
      for (int d = 0; d < L.dim; ++d) v[d] = 0.0;
      wet_density(&L, r, f, v);
      for (int d = 0; d < 4; ++d)
        if (nbr[c * (Q - 1) + d] < 0) f[L.opp[d]] = f[d];
      if (r->axis == 0 && r->sign < 0) {
        f[6] = f[4] + 0.5 * (f[3] - f[2]);
        f[7] = f[5] - 0.5 * (f[3] - f[2]);
      } else if (r->axis == 0) {
        f[4] = f[6] - 0.5 * (f[3] - f[2]);
        f[5] = f[7] + 0.5 * (f[3] - f[2]);
      } else if (r->sign < 0) {
        f[5] = f[7] - 0.5 * (f[1] - f[0]);
        f[6] = f[4] + 0.5 * (f[1] - f[0]);
      } else {
        f[7] = f[5] + 0.5 * (f[1] - f[0]);
        f[4] = f[6] - 0.5 * (f[1] - f[0]);
      }
    } else if (r->kind == 5 || r->kind == 6) {
      /* NEEM faces (Guo, Zheng & Shi, Chin. Phys. 11 366 (2002)): every
       * population of c from the inward neighbour n1 (complete after the
       * bounce-back records),
       *   f_i(c) = f_i(n1) + feq_i(rho_b, u_b) - feq_i(rho1, u1),
       * rho1, u1 the moments of f(n1) (u1 with the Guo half-force shift);
       * pressure face (5): rho_b = rho, u_b = u1; velocity face (6):
       * rho_b = rho1, u_b = u. */
      const double *f1 = fold + r->n1 * Q;
      double v1[4];
      macros_cell(&L, f1, has_force ? force : NULL, v1);
      const double rho1 = v1[L.dim];
      const double rhob = r->kind == 5 ? r->rho : rho1;
      const double *ub = r->kind == 5 ? v1 : r->u;
      double vsq1 = 0, vsqb = 0;
      for (int d = 0; d < L.dim; ++d) {
        vsq1 += v1[d] * v1[d];
        vsqb += ub[d] * ub[d];
      }
      double fn[19];
      for (int i = 0; i < Q; ++i) {
        double cu1 = 0, cub = 0;
        for (int d = 0; d < L.dim; ++d) {
          cu1 += L.c[i][d] * v1[d];
          cub += L.c[i][d] * ub[d];
        }
        fn[i] = f1[i] + default_eq(L.w[i], rhob, cub, vsqb) - default_eq(L.w[i], rho1, cu1, vsq1);
      }
      for (int i = 0; i < Q; ++i) fold[c * Q + i] = fn[i];
    } else {
      return -2;
    }
  }
  return 0;
}

/* Initial condition, src/lbm/solver.cpp:267-295: fold = feq(rho, u). */
int orc_init(int Q, int64_t ncell, const double *rho, const double *u /*[ncell][dim]*/, double *fold) {
  lattice_t L;
  if (get_lattice(Q, &L)) return -1;
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < ncell; ++c) feq_all(&L, rho[c], u + c * L.dim, fold + c * Q);
  return 0;
}

/* Residual sums between two macro snapshots (vars layout u..,rho):
 * out[0..nv-1] = sum |q - q_old| per variable (sumAbsDiff, solver.cpp:808-815)
 * out[nv]      = sum |u - u_old|^2,  out[nv+1] = sum |u|^2   (new: R_rel)
 * Summed serially in cell order, as the reference does. */
int orc_residual(int dim, int64_t ncell, const double *vars, const double *vars_old, double *out) {
  const int nv = dim + 1;
  for (int k = 0; k < nv + 2; ++k) out[k] = 0;
  for (int64_t c = 0; c < ncell; ++c) {
    double d2 = 0, n2 = 0;
    for (int k = 0; k < nv; ++k) out[k] += fabs(vars[c * nv + k] - vars_old[c * nv + k]);
    for (int d = 0; d < dim; ++d) {
      const double du = vars[c * nv + d] - vars_old[c * nv + d];
      d2 += du * du;
      n2 += vars[c * nv + d] * vars[c * nv + d];
    }
    out[nv] += d2;
    out[nv + 1] += n2;
  }
  return 0;
}

/* hilbert::index, include/common/math/hilbert.h:15-48 (no rotation; fixed
 * quadrant lookup table). x in unit-cube coordinates. */
int64_t orc_hilbert_index(int dim, const double *x, int64_t level) {
  static const int64_t LUT[16] = {0, 3, 1, 2, 5, 4, 6, 7, 10, 9, 11, 8, 15, 14, 12, 13};
  double p[4];
  for (int d = 0; d < dim; ++d) p[d] = x[d];
  int64_t index = 0;
  for (int64_t l = 0; l < level; ++l) {
    int q = 0;
    int qb[4];
    for (int d = 0; d < dim; ++d) {
      qb[d] = p[d] >= 0.5;
      q |= qb[d] << d;
    }
    int64_t mult = 1;
    for (int64_t k = 0; k < dim * (level - 1 - l); ++k) mult *= 2;
    index += mult * LUT[q];
    for (int d = 0; d < dim; ++d) p[d] = 2 * p[d] - qb[d];
  }
  return index;
}

/* Round-to-nearest-even quantisers used to model shifted reduced storage. */
static inline float bits_to_f32(uint32_t b) {
  float f;
  memcpy(&f, &b, 4);
  return f;
}
static inline uint32_t f32_to_bits(float f) {
  uint32_t b;
  memcpy(&b, &f, 4);
  return b;
}
double orc_round_bf16(double x) {
  uint32_t b = f32_to_bits((float)x);
  if ((b & 0x7f800000u) == 0x7f800000u) return (double)bits_to_f32(b);
  const uint32_t lsb = (b >> 16) & 1u;
  b += 0x7fffu + lsb;
  b &= 0xffff0000u;
  return (double)bits_to_f32(b);
}
/* IEEE binary16 round-to-nearest-even of the float value of x, returned as
 * double (exact).  Written from the format definition: 1 sign, 5 exponent
 * (bias 15), 10 mantissa bits; subnormals below 2^-14; overflow to inf. */
double orc_round_f16(double x) {
  const float f = (float)x;
  if (f != f) return (double)f;
  const double a = fabs((double)f);
  if (a == 0) return (double)f;
  int e;
  frexp(a, &e); /* a = m * 2^e, m in [0.5,1) */
  int ulp_exp = e - 11; /* 11 significant bits */
  if (ulp_exp < -24) ulp_exp = -24; /* subnormal spacing 2^-24 */
  const double ulp = ldexp(1.0, ulp_exp);
  double q = a / ulp; /* exact: power-of-two scaling */
  double r = floor(q);
  const double frac = q - r;
  if (frac > 0.5 || (frac == 0.5 && fmod(r, 2.0) != 0)) r += 1;
  double y = r * ulp;
  if (y > 65504.0) y = INFINITY;
  return f < 0 ? -y : y;
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
