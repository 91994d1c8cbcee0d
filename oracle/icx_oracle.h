/*
 * icx_oracle.h -- CPU restatement of ICARAT's SIMP linear-elastic topology
 * optimisation hot path (FP64).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * path and the timed CPU baseline.  Only tests/, the smoke check and the CPU
 * baseline leg of bench.py may load it.  The product library (libicx) never
 * links or calls it.
 *
 * Every function names the reference file:line it restates.  Paths are
 * relative to the reference repository root.
 *
 * Pins (tests/test_oracle_*.py, tests/test_golden.py): mesh, DOF map and loads by
 * closed forms and hand-checked grids; element matrices by closed forms, rigid modes
 * and the patch test; solves by direct solutions and scipy; sensitivities, projection
 * and the stress constraint's adjoint derivative by finite differences; filters by
 * brute force, invariants and adjoint identities; OC by the closed-form limit lambda*
 * of a separable problem (exponent, move limits, xmin, bracket) and a replayed
 * bisection; MMA by the asymptote history worked by hand over designed x sequences
 * (init rule, gamma 0.7 / 1.2 / 1, both clamps, the out-of-box rule), literal GenSub
 * coefficients (0.001 and raa0 terms, alpha/beta with albefa and the move limit),
 * first-order exactness at xval and the KKT conditions of the m = 1 subproblem;
 * GCMMA by closed-form optima, SLSQP on the subproblem, first-order exactness and the
 * RaaUpdate identity.  Parity unpinned: the exact Eigen CG iterates (the algorithm is
 * restated, iteration counts agree within about 1%) and SimplicialLDLT's rounding.
 */
#ifndef ICX_ORACLE_H
#define ICX_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* material state, elasticity.cpp:19-32 ("plane_stress", "plane_strain", "3D"), and
 * heat conduction (src/problem/heat.hpp, one DOF per node) */
enum { ICXO_PLANE_STRESS = 0, ICXO_PLANE_STRAIN = 1, ICXO_SOLID3D = 2, ICXO_HEAT = 3 };

/* ---- mesh / boundary conditions (integer, bit-exact) ---------------------- */
int icxo_mesh_box(int ne, const int *div, const double *len, double *x, int *conn);
void icxo_dirichlet(int ndim, int dofnp, int numnp, const double *x,
                    const int *flag6, const double *val6, const double *range6,
                    int *has, int *node_flag, double *node_val);
int icxo_count_dof(int numnp, int dofnp, const int *has, const int *node_flag, int *dof);
void icxo_neumann(int ndim, int dofnp, int ne, int nelx, const double *x,
                  const int *conn, const int *dof, int numeq,
                  const double *val3, const double *range6, double *fext);

/* ---- element level --------------------------------------------------------- */
double icxo_simp(double young1, double young2, double design, double p);
void icxo_element_stiffness(int ne, int ndim, int state, double E, double nu,
                            const double *X, double *Ke, double *volume);
void icxo_helm_element(int ne, int ndim, double radius, const double *X,
                       double *Ke, double *Nint);
/* heat: conductivity tensor (ndim x ndim) of a 6-entry row; element conductivity matrix */
void icxo_heat_tensor(int ndim, const double *c6, double *D);
void icxo_heat_element(int ne, int ndim, const double *D, const double *X, double *Ke, double *volume);

/* ---- global structural system -------------------------------------------- */
typedef struct icxo_sys icxo_sys;
icxo_sys *icxo_sys_create(int ne, int ndim, int dofnp, int state, int numnp, int nelx,
                          const double *x, const int *conn, const int *dof, int numeq,
                          double young1, double young2, double nu, double p);
void icxo_sys_destroy(icxo_sys *s);
long long icxo_sys_nnz(const icxo_sys *s);
void icxo_sys_assemble(icxo_sys *s, const double *design, const int *has,
                       const int *node_flag, const double *node_val, double *fint);
void icxo_sys_csr(const icxo_sys *s, int *rowptr, int *col, double *val);
void icxo_sys_spmv(const icxo_sys *s, const double *x, double *y);
int icxo_sys_cg(const icxo_sys *s, const double *b, double *x, double tol, int maxit,
                double *err);
int icxo_sys_direct(const icxo_sys *s, const double *b, double *x);
void icxo_sys_element_energy(const icxo_sys *s, const double *nodeval, double *ue_K_ue);
/* element-mean von Mises stress (Gauss-point average, elasticity.cpp:165-208) */
void icxo_sys_mises(const icxo_sys *s, const double *nodeval, double *mises);
/* reactions Ke u_e summed over all DOFs, flushed, 3 per node (linear.hpp:90-91) */
void icxo_sys_reactions(const icxo_sys *s, const double *nodeval, double *out);
/* heat (state ICXO_HEAT): conductivities of materials 1 and 2 (6 entries each), and
 * the sensitivity -T_e^T dKe/ds T_e (opt-heat/opt_heat.cpp:158-194) */
void icxo_sys_set_conductivity(icxo_sys *s, const double *k1, const double *k2);
void icxo_sys_heat_sensitivity(const icxo_sys *s, const double *nodeval, double *dfds);

/* ---- Eigen CG restatement on a CSR matrix -------------------------------- */
int icxo_csr_cg(int n, const int *rowptr, const int *col, const double *val,
                const double *b, double *x, double tol, int maxit, double *err);

/* ---- solve-level helpers ---------------------------------------------------- */
void icxo_divide_val_into_node(int numnp, int dofnp, int numeq, const int *dof,
                               const double *u, const double *node_val_dirich,
                               double *nodeval);
void icxo_sensitivity(int nelx, int driver_b, double young1, double young2, double p,
                      const double *design, const double *ue_K0_ue_over_E,
                      const double *volume, double V0, double *dfds, double *dgds);

/* ---- cone filter, projection ------------------------------------------------ */
typedef struct icxo_filter icxo_filter;
icxo_filter *icxo_filter_create(int ne, int ndim, int nelx, const double *x,
                                const int *conn, double radius, int brute_force);
void icxo_filter_destroy(icxo_filter *f);
long long icxo_filter_nnz(const icxo_filter *f);
void icxo_filter_neighbours(const icxo_filter *f, int *ptr, int *id, double *dist);
void icxo_filter_apply(icxo_filter *f, const double *in, double *out, int is_sens);
/* Filter::sensitivityFilter (filter.hpp:72-100): volume per element */
void icxo_filter_sensitivity(const icxo_filter *f, const double *design, const double *dfds,
                             const double *volume, double *out);
void icxo_threshold(int n, const double *in, double beta, double T, double *out);
void icxo_threshold_derivative(int n, const double *in, double beta, double T, double *out);

/* ---- Helmholtz PDE filter --------------------------------------------------- */
typedef struct icxo_helm icxo_helm;
icxo_helm *icxo_helm_create(int ne, int ndim, int numnp, int nelx, const double *x,
                            const int *conn, double radius);
void icxo_helm_destroy(icxo_helm *h);
int icxo_helm_apply(icxo_helm *h, const double *in, double *out, int sens,
                    double tol, int maxit, int direct);

/* ---- design updaters ------------------------------------------------------- */
typedef double (*icxo_volume_fn)(const double *x, void *user);
int icxo_oc_update(int n, double *x, const double *dfdx, double g, const double *dgdx,
                   const double *xmin, const double *xmax, double movelimit,
                   icxo_volume_fn gk, void *user);

typedef struct icxo_mma icxo_mma;
icxo_mma *icxo_mma_create(int n, int m, double a, double c, double d);
void icxo_mma_destroy(icxo_mma *m);
int icxo_mma_update(icxo_mma *m, double *x, const double *dfdx, const double *gx,
                    const double *dgdx, const double *xmin, const double *xmax);
/* the subproblem of the last update (any pointer may be NULL): asymptotes low/upp, move
 * bounds alpha/beta (n each), p0/q0 (n), pij/qij (n*m, row i*m+j), b (m), and the dual
 * solution lam, y (m) and z */
void icxo_mma_subproblem(const icxo_mma *m, double *low, double *upp, double *alpha,
                         double *beta, double *p0, double *q0, double *pij, double *qij,
                         double *b, double *lam, double *y, double *z);
/* SetAsymptotes(init, decrease, increase) (MMASolver.hpp:91-97) */
void icxo_mma_set_asymptotes(icxo_mma *m, double init, double decrease, double increase);
/* GCMMASolver(n, m, a=0, c=1000, d=0): OuterUpdate, InnerUpdate, ConCheck
 * (src/optimization/opt-solver/GCMMASolver.cpp:41-129); destroy with icxo_mma_destroy */
icxo_mma *icxo_gcmma_create(int n, int m, double a, double c, double d);
int icxo_gcmma_outer(icxo_mma *m, double *xmma, const double *xval, double f0x,
                     const double *df0dx, const double *fx, const double *dfdx,
                     const double *xmin, const double *xmax);
int icxo_gcmma_inner(icxo_mma *m, double *xmma, double f0xnew, const double *fxnew,
                     const double *xval, double f0x, const double *df0dx, const double *fx,
                     const double *dfdx, const double *xmin, const double *xmax);
int icxo_gcmma_concheck(const icxo_mma *m, double f0xnew, const double *fxnew);
int icxo_stress_constraint(const icxo_sys *s, const double *nodeval, const double *node_val_dirich,
                           double qp, double pnorm, double slim, int direct, double cg_tol,
                           double *g, double *dgds, double *sqp_out);
void icxo_gcmma_approx_at(const icxo_mma *m, const double *x, double *f0, double *fj);
void icxo_gcmma_state(const icxo_mma *m, double *f0app, double *fapp, double *raa0, double *raa);

/* ---- drivers ----------------------------------------------------------------
 * A: opt-macro-oc/opt_macro_oc.cpp:84-172  (cone filter, projection, CG, OC)
 * B: opt-mma-HelmholtzFilter/opt_mma_Helmholtz.cpp:81-163 (Helmholtz, MMA)
 * 2: opt_macro_gcmma.cpp:83-216 (cone filter r, beta = min(int(max((k-100)/40, 1)), 8),
 *    GCMMA with up to 15 inner loops, each re-solving a trial design)
 * 3: constraint/opt-stress-constraint/opt_stress_constraint.cpp:85-158 (2D: cone filter r,
 *    beta = min(2^(k/40), 8), f = volume, g = stress p-norm, MMA with asyminit 0.1);
 *    obj_hist = volume, vol_hist = g, last_dfds = the filtered dg */
typedef struct {
  int driver;          /* 0 = A (OC + cone), 1 = B (MMA + Helmholtz), 2 = GCMMA + cone */
  int ne, ndim, dofnp, state;
  int numnp, nelx, numeq;
  const double *x; const int *conn; const int *dof;
  const int *has; const int *node_flag; const double *node_val;
  const double *fext;
  double young1, young2, nu, p;
  double filter_radius;
  double xmin, xmax, move;
  int iterations;
  int projection;      /* 1: beta schedule + tanh projection (reference), 0: off */
  int direct;          /* 1: direct solve for the structure, 0: Eigen CG restatement */
  double cg_tol; int cg_maxit;
  double helm_tol;     /* Helmholtz CG tolerance (reference 1e-10) */
  int helm_direct;
  int *inner_hist;     /* driver 2: GCMMA inner loops per step (may be NULL) */
  double stress_limit, p_norm, qp;  /* driver 3 ([this_problem] of the stress driver) */
} icxo_driver_cfg;

int icxo_run_driver(const icxo_driver_cfg *cfg, double *design /* in: x0, out: final */,
                    double *obj_hist, double *vol_hist, double *last_dfds,
                    double *last_u /* numeq */, int *cg_iters_hist);

#ifdef __cplusplus
}
#endif
#endif
