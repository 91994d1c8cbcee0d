/*
 * icx.h -- C ABI of the MI355X (gfx950) implementation of ICARAT's
 * SIMP linear-elastic topology-optimisation hot path.
 *
 * Conventions
 *  - Every entry point returns an icx_status (0 = ok).  Nothing calls exit().
 *    icx_last_error() returns a message for the last failure of a context.
 *  - Host buffers are caller-owned and copied; device state is owned by the
 *    job.  Vectors on the DOF axis use the reference's equation numbering
 *    (free DOFs 0..numeq-1 in node order, src/io/boundary_condition_base.hpp:35-88);
 *    element vectors use the reference's element ids (x fastest,
 *    src/io/mesh_icarat.hpp:325-355, 440-494).
 *  - One context = one GPU + one HIP stream; use a context from one host thread
 *    at a time.  Jobs of the same context run in submission order.
 *
 * Each declaration names the reference interface it replaces (paths relative
 * to the reference repository root).
 */
#ifndef ICX_H
#define ICX_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int icx_status;
enum {
  ICX_OK = 0,
  ICX_ERR_ARG = 1,      /* invalid argument */
  ICX_ERR_STATE = 2,    /* call out of order (e.g. solve before icx_make_dof) */
  ICX_ERR_HIP = 3,      /* HIP runtime error */
  ICX_ERR_NOMEM = 4,    /* device allocation failed / over the memory budget */
  ICX_ERR_NOCONV = 5,   /* iterative solve hit maxit (results are still written) */
  ICX_ERR_NUMERIC = 6   /* non-finite value / zero pivot */
};

/* element types (src/io/mesh_base.hpp:50-115) */
enum { ICX_QUAD4 = 4, ICX_HEXA8 = 8 };
/* material state (src/problem/elasticity.cpp:19-32) */
enum { ICX_PLANE_STRESS = 0, ICX_PLANE_STRAIN = 1, ICX_SOLID3D = 2, ICX_HEAT = 3 };
/* solver forms */
enum {
  ICX_SOLVE_PLAIN = 0, /* FP64 Jacobi-PCG, Eigen ConjugateGradient loop (parity form) */
  ICX_SOLVE_FUSED = 1, /* FP64 Jacobi-PCG, fused kernels (same iterates up to rounding) */
  ICX_SOLVE_MIXED = 2  /* FP32 inner PCG + FP64 residual replacement (reliable updates) */
};
/* filter kinds */
enum { ICX_FILTER_NONE = 0, ICX_FILTER_CONE = 1, ICX_FILTER_HELMHOLTZ = 2 };

typedef struct icx_ctx icx_ctx;
typedef struct icx_job icx_job;

/* Structured box problem (src/io/mesh_icarat.hpp:69-128 [mesh] keys and
 * src/problem/elasticity.cpp:19-32 [material] keys). */
typedef struct {
  int eltype;        /* ICX_QUAD4 or ICX_HEXA8 */
  int div[3];        /* [mesh] division (div[2] ignored for quad4) */
  double length[3];  /* [mesh] length */
  int state;         /* ICX_PLANE_STRESS / ICX_PLANE_STRAIN / ICX_SOLID3D, or ICX_HEAT:
                        heat conduction (LinearHeat, src/problem/heat.hpp), one DOF per
                        node, quad4 or hexa8; young and poisson unused, penalty = p */
  double young1;     /* [material] young[0] (E1; Emin when young2 > young1) */
  double young2;     /* [material] young[1] (E2; E0) */
  double poisson;    /* [material] poisson */
  double penalty;    /* [material] penalty (SIMP p) */
} icx_problem;

typedef struct {
  int iterations;        /* Eigen semantics: passes that did not converge */
  int matvecs;           /* operator applications (inner + FP64 residual) */
  int fp64_matvecs;      /* of which FP64 */
  int refinements;       /* FP64 residual replacements (mixed form) */
  double relres;         /* |b - K u| / |b| (recursive for plain, true for mixed) */
  double seconds;        /* device time of the solve (HIP events) */
} icx_solve_info;

/* ---- context ---------------------------------------------------------------- */
icx_status icx_ctx_create(int device, icx_ctx **out);
icx_status icx_ctx_destroy(icx_ctx *ctx);
const char *icx_last_error(const icx_ctx *ctx);
icx_status icx_ctx_synchronize(icx_ctx *ctx);

/* ---- job set-up --------------------------------------------------------------
 * Replaces MeshIcarat::setParameter/generate (src/io/mesh_icarat.hpp:69-136),
 * LinearElastic::setParameter (src/problem/elasticity.cpp:19-32). */
icx_status icx_job_create(icx_ctx *ctx, const icx_problem *prob, icx_job **out);
icx_status icx_job_destroy(icx_job *job);
/* A box job from the text of a reference job file (TOML), as the drivers set one up:
 * [mesh] dimension / type / division / length (MeshIcarat::setParameter,
 * src/io/mesh_icarat.hpp:69-128), [material] (Elastic::setParameter,
 * src/problem/elasticity.cpp:19-32; Heat::setParameter, heat.cpp:19-60), then the
 * [boundary_condition] dirichlet ranges, countDOF and the neumann tractions
 * (BCIcarat::input_dirich / makeDOF / input_neumann,
 * src/io/boundary_condition_icarat.hpp:68-396; boundary_condition_base.cpp:143-180).  The
 * job comes back with its DOF map and load; numeq and design0 (either may be NULL)
 * receive the free DOF count and designs_0.  A malformed or unsupported spec returns
 * ICX_ERR_ARG with the reason in icx_last_error.  The C++ drivers' StructuredModel builds
 * its jobs with the same code. */
icx_status icx_job_create_toml(icx_ctx *ctx, const char *toml_text, icx_job **out, int *numeq,
                               double *design0);

/* Unstructured mesh of one element type (MeshMDPA, src/io/mesh_mdpa.hpp:84-238, with
 * the elements of LinearElastic, src/problem/elasticity.cpp:210-254): node coordinates
 * and 0-based connectivity in the reference's local node order (quad4: counter-clockwise;
 * hexa8: the z-low face counter-clockwise, then the z-high face).  Element matrices are
 * integrated per element (2 x 2 (x 2) Gauss points); an element with det J <= 0 at a
 * Gauss point is rejected (ICX_ERR_NUMERIC).  Mesh jobs take every solve form (plain,
 * fused, mixed), the cone and Helmholtz filters, von Mises stress and the design loop;
 * the box-only entry points (icx_matvec at FP32, submitted solves) return ICX_ERR_ARG. */
typedef struct {
  int eltype;          /* ICX_QUAD4 or ICX_HEXA8 */
  int numnp, nelx;
  const double *x;     /* numnp x 3 (z ignored for quad4) */
  const int *conn;     /* nelx x ne, 0-based */
  int state;           /* ICX_PLANE_STRESS / ICX_PLANE_STRAIN (quad4), ICX_SOLID3D (hexa8) */
  double young1, young2, poisson, penalty;
} icx_mesh;
icx_status icx_job_create_mesh(icx_ctx *ctx, const icx_mesh *mesh, icx_job **out);
/* MDPA file (MeshMDPA::setParameter / generate, src/io/mesh_mdpa.hpp:84-216): the nodes
 * of the "Begin Nodes" block in file order and the elements of every "Begin Elements"
 * block whose name carries the element's tag ("2D4" quad4, "3D8" hexa8), in file order;
 * connectivity as node id - 1.  x and conn may be NULL to query the sizes first. */
icx_status icx_mdpa_read(const char *path, int eltype, int *numnp, int *nelx, double *x, int *conn);
/* numnp, nelx, neq (= numnp * dofnp) */
icx_status icx_job_sizes(const icx_job *job, int *numnp, int *nelx, int *neq);
/* the job's problem descriptor (mesh jobs: eltype, state and material; div and length
 * zero) -- what FEM / the [mesh] and [material] keys hold for the reference's drivers */
icx_status icx_job_get_problem(const icx_job *job, icx_problem *prob);
/* node coordinates (numnp x 3) and connectivity (nelx x ne), as MeshIcarat builds them */
icx_status icx_get_mesh(const icx_job *job, double *x, int *conn);

/* BCIcarat::input_dirich for one [boundary_condition].dirichlet entry
 * (src/io/boundary_condition_icarat.hpp:68-121): flag[6], val[6],
 * range[6] = {xmin,xmax,ymin,ymax,zmin,zmax}; tolerance 1e-7. */
icx_status icx_add_dirichlet(icx_job *job, const int *flag6, const double *val6,
                             const double *range6);
/* BCIcarat::makeDOF -> countDOF (src/io/boundary_condition_icarat.hpp:191-200). */
icx_status icx_make_dof(icx_job *job, int *numeq);
/* node.dof[] as numnp x dofnp (free 0..numeq-1, constrained numeq..neq-1) */
icx_status icx_get_dof_map(const icx_job *job, int *dof);
/* BCIcarat::input_neumann for one [boundary_condition].neumann entry
 * (src/io/boundary_condition_icarat.hpp:253-396), accumulated into fext_org. */
icx_status icx_add_neumann(icx_job *job, const double *val3, const double *range6);
/* Force::fext_org (src/problem/base.hpp:108-123) read/write, numeq doubles. */
icx_status icx_set_fext(icx_job *job, const double *fext);
icx_status icx_get_fext(const icx_job *job, double *fext);
/* adds `value` to fext_org of (node, comp) if that DOF is free */
icx_status icx_add_nodal_load(icx_job *job, int node, int comp, double value);

/* ---- analysis ------------------------------------------------------------------
 * LinearElastic::setDesignS for every element (src/problem/elasticity.hpp:60):
 * the SIMP modulus of element e is computed from design[e] on the device.  Heat jobs:
 * LinearHeat::setDesign (src/problem/heat.hpp:34), conductivity
 * design^p De1 + (1 - design^p) De2 (Heat::makeSIMPDe, heat.cpp:62-77).
 * A design outside [0, 1] (negative, above 1, NaN or infinite) is refused with
 * ICX_ERR_NUMERIC and leaves the job's design unchanged (the reference does not check).
 * While a submitted solve of the job runs, this call and the others that change what
 * it reads (loads, Dirichlet ranges, solves, design-loop steps) return ICX_ERR_STATE. */
icx_status icx_set_design(icx_job *job, const double *design);
/* Heat jobs: the [material] conductivity rows of materials 1 and 2, six entries each
 * (xx, yy, zz, xy, yz, xz), as read by Heat::setParameter (src/problem/heat.cpp:19-60;
 * 2D uses xx, yy, xy).  Material 1 must conduct at least as well as material 2 in xx
 * (heat.cpp:66-70).  Call before icx_set_design.  (The reference's 3D branch assigns
 * material 2 to De1_ and leaves De2_ unset; both tensors are taken as given here.)
 * Heat jobs solve with ICX_SOLVE_PLAIN only; icx_sensitivity gives
 * -p s^(p-1) T_e^T (K1 - K2) T_e (opt-heat/opt_heat.cpp:158-194 before the projection
 * factor); icx_write_vtk writes temperature, threshold_s and design_s
 * (opt_heat.cpp:196-236). */
icx_status icx_set_conductivity(icx_job *job, const double *k1, const double *k2);
/* y = K x on the free DOFs (matrix-free; x, y numeq doubles).
 * prec: 0 = FP64, 1 = FP32 operator (inputs rounded to FP32). */
icx_status icx_matvec(icx_job *job, int prec, const double *x, double *y);
/* No counterpart in the reference.  icx_matvec for callers that cannot allocate device
 * memory (serving loops and harnesses that forbid hipMalloc): the same kernels, with the
 * device staging (x and y on the DOF axis; for prec 1 also the FP32 operand, result and
 * moduli) taken from a 64 MiB array that the library holds on each GPU until the process
 * ends, instead of buffers allocated per call.  Calls no allocating routine.
 * Limit: 8 numeq bytes (prec 0), plus 8 N_grid + 4 nelx bytes for prec 1, each rounded
 * up to 256, at most 64 MiB (N_grid = nodes x DOFs per node); larger jobs get
 * ICX_ERR_ARG.  Every call on one GPU uses the same array: calls that overlap on one GPU
 * (two contexts or threads at once) race; each call waits for its work before it
 * returns, so calls made one after another are safe. */
icx_status icx_matvec_noalloc(icx_job *job, int prec, const double *x, double *y);
/* Jacobi diagonal of K (numeq doubles) */
icx_status icx_diagonal(icx_job *job, double *diag);
/* LinearAnalysis::solve (src/analysis/linear.hpp:54-104) with the solver
 * facade Solvers::solve "Eigen_CG" (src/analysis/solver/solvers.cpp:64-92):
 * rhs = fext_org - K u_dirichlet; x0 = 0; stop on |r| < tol |b|.
 * u (numeq doubles) may be NULL: the solution stays on the device. */
icx_status icx_solve(icx_job *job, int form, double tol, int maxit, double *u,
                     icx_solve_info *info);
/* Value::val of the last solve (numeq doubles) */
icx_status icx_get_displacement(const icx_job *job, double *u);
/* node.val after divideValIntoNode (numnp x 3, |v| < 1e-15 flushed,
 * src/analysis/analysis_utility.hpp:8-31) */
icx_status icx_get_node_values(const icx_job *job, double *nodeval);
/* compliance u . fext (opt-macro-oc/opt_macro_oc.cpp:110) */
icx_status icx_compliance(icx_job *job, double *c);

/* ---- output step ---------------------------------------------------------------
 * Element-mean von Mises stress of the last solve and the current design
 * (Elastic::makeMisesStress, src/problem/elasticity.cpp:165-208: strain and stress
 * averaged over the Gauss points, sqrt(s^T V s); in 2D the reference's V without s_zz,
 * also for plane strain).  out: nelx doubles in element-id order. */
icx_status icx_mises(icx_job *job, double *out);
/* Reactions: the internal force K u over all DOFs after a solve (the re-assembly in
 * LinearAnalysis::solve, src/analysis/linear.hpp:90-91, 117-166, assemblingReact
 * src/misc/assembling.hpp:77-84) scattered to nodes like divideReactionIntoNode
 * (src/analysis/analysis_utility.hpp:33-51): out numnp x 3, |r| < 1e-15 flushed;
 * equals fext on free DOFs at convergence. */
icx_status icx_reactions(icx_job *job, double *out);
/* The drivers' output() (opt-macro-oc/opt_macro_oc.cpp:214-238,
 * opt-mma-HelmholtzFilter/opt_mma_Helmholtz.cpp:207-235, src/io/export_vtk.hpp:47-156)
 * as a binary legacy VTK file, in place of the ASCII file and the meshio-binary
 * conversion: points, cells (reference node order), cell types, displacement (node.val),
 * density (the current design) and mises.  icx_output_begin queues the device work and
 * the copies to pinned host memory on the job's stream and returns at once (later
 * device work on the stream, e.g. the next design step, does not disturb them);
 * icx_output_write waits for them and writes `path`.  icx_write_vtk does both. */
icx_status icx_output_begin(icx_job *job);
icx_status icx_output_write(icx_job *job, const char *path);
icx_status icx_write_vtk(icx_job *job, const char *path);
/* the drivers' sensitivity() (opt-macro-oc/opt_macro_oc.cpp:182-212,
 * opt-mma-HelmholtzFilter/opt_mma_Helmholtz.cpp:173-205):
 * dfds = -(dE/E) u_e^T Ke u_e, dgds = v_e / vscale (vscale = 1 or V0).
 * Either output may be NULL. */
icx_status icx_sensitivity(icx_job *job, double vscale, double *dfds, double *dgds);
/* element volumes (Element::volume, src/problem/elasticity.cpp:249) */
icx_status icx_element_volume(const icx_job *job, double *vol);

/* ---- filters and projection -------------------------------------------------
 * Filter(fem, element, node, r) + densityFilter(v, isSens)
 * (src/optimization/filter.hpp:17-151, 191-234); adjoint = isSens. */
icx_status icx_cone_filter_setup(icx_job *job, double radius);
icx_status icx_cone_filter(icx_job *job, int adjoint, const double *in, double *out);
/* Filter::sensitivityFilter(design_s, dfds) (src/optimization/filter.hpp:72-100): the
 * classic sensitivity filter over the same neighbours (icx_cone_filter_setup first),
 * out_i = ((x_i v_i) / (x_i a_i)) sum_j (w_ij x_j dfds_j) / (x_j v_j), a_i = sum_j w_ij,
 * v the element volumes; as in the reference an element with x = 0 gives a non-finite
 * value.  Neither reference driver calls it. */
icx_status icx_sensitivity_filter(icx_job *job, const double *design, const double *dfds, double *out);
/* FilterHelmholtz(fem, element, node, R) + densityFilter(v, "Eigen_CG", log, sensFlag)
 * (src/optimization/filter_helm.hpp:23-137, 178-256); solves to tol (1e-10 in
 * the reference).  iters may be NULL. */
icx_status icx_helmholtz_setup(icx_job *job, double radius);
icx_status icx_helmholtz_filter(icx_job *job, int sens, double tol, const double *in,
                                double *out, int *iters);
/* threshold / thresholdDerivative (src/optimization/filter.hpp:153-185);
 * out or dout may be NULL. */
icx_status icx_project(icx_job *job, double beta, double T, const double *in, double *out,
                       double *dout);

/* ---- design updates -----------------------------------------------------------
 * OC::UpdateVariables (src/optimization/opt-solver/OCSolver.hpp:20-68) with the
 * volume lambda of opt-macro-oc/opt_macro_oc.cpp:155-163 evaluated on the
 * device: g(x) = sum_e threshold(filter(x))_e v_e (filter = the job's
 * filter of kind `filter`, projection skipped when beta <= 0).
 * x is updated in place (nelx doubles); xmin/xmax scalars. */
icx_status icx_oc_update(icx_job *job, double *x, const double *dfdx, double g,
                         const double *dgdx, double xmin, double xmax, double move,
                         int filter, double beta, double T, int *bisections);
/* MMASolver(n, m=1) + Update (src/optimization/opt-solver/MMASolver.hpp:62-116,
 * MMASolver.cpp:41-553); the optimiser history lives on the device. */
icx_status icx_mma_reset(icx_job *job);
/* MMASolver::SetAsymptotes(init, decrease, increase) (MMASolver.hpp:71-76; defaults
 * 0.2, 0.7, 1.2; the stress driver sets 0.1, 0.7, 1.2, opt_stress_constraint.cpp:72) */
icx_status icx_mma_set_asymptotes(icx_job *job, double init, double decrease, double increase);
icx_status icx_mma_update(icx_job *job, double *x, const double *dfdx, double gx,
                          const double *dgdx, double xmin, double xmax, int *newton_steps);
/* GCMMASolver(n, m = 1, a = 0, c = 1000, d = 0) (src/optimization/opt-solver/
 * GCMMASolver.hpp:61-145, GCMMASolver.cpp:41-661) with its state on the device; xmin /
 * xmax are scalars.  A job's first use starts from raa0 = raa = 1; icx_gcmma_reset is
 * Reset() (the outer counter only, :88).
 * outer: OuterUpdate(xmma, xval, f0x, df0dx, fx, dfdx, xmin, xmax) (:66-90).
 * inner: InnerUpdate(xmma, f0xnew, fxnew, xval, ...) (:92-113); xmma holds the point
 *   the last update returned (where f0xnew / fxnew were evaluated) and is overwritten.
 * concheck: ConCheck(f0xnew, fxnew) (:115-129), 1 = conservative. */
icx_status icx_gcmma_reset(icx_job *job);
icx_status icx_gcmma_outer_update(icx_job *job, double *xmma, const double *xval, double f0x,
                                  const double *df0dx, double fx, const double *dfdx,
                                  double xmin, double xmax, int *newton_steps);
icx_status icx_gcmma_inner_update(icx_job *job, double *xmma, double f0xnew, double fxnew,
                                  const double *xval, double f0x, const double *df0dx,
                                  double fx, const double *dfdx, double xmin, double xmax,
                                  int *newton_steps);
icx_status icx_gcmma_concheck(const icx_job *job, double f0xnew, double fxnew, int *conservative);

/* Stress constraint of the stress-constrained driver (constraint/opt-stress-constraint/
 * sensitivity.cpp:8-155: calConstraintValue and sensitivity) on the job's current design
 * and solve; 2D box jobs (the reference exits in 3D), young2 > young1.
 *   g = (sum_e (rho_e^qp E2 vm_e / (E_e stress_limit))^p_norm)^(1/p_norm) - 1
 * with vm_e the element-mean von Mises stress.  dgds (nelx doubles, NULL: g only) is
 * dg/drho_hat by one adjoint solve (the FP64 fused form of the Eigen CG loop at tol,
 * maxit); the job's
 * displacement stays.  adjoint_iterations may be NULL. */
icx_status icx_stress_constraint(icx_job *job, double stress_limit, double p_norm, double qp,
                                 double tol, int maxit, double *g, double *dgds,
                                 int *adjoint_iterations);

/* ---- device-resident design loops ------------------------------------------------
 * One design iteration of opt-macro-oc/opt_macro_oc.cpp:84-172 (driver 0: cone
 * filter + OC), opt-mma-HelmholtzFilter/opt_mma_Helmholtz.cpp:81-163 (driver 1:
 * Helmholtz + MMA) or opt_macro_gcmma.cpp:83-216 (driver 2: cone filter + GCMMA, with
 * beta = min(int(max((k-100)/40, 1)), 8), const_h = V - V0 and up to 15 inner loops,
 * each solving a trial design), with every vector kept on the device. */
typedef struct {
  int driver;          /* 0 = OC driver (A), 1 = MMA + Helmholtz driver (B), 2 = GCMMA */
  int filter;          /* ICX_FILTER_CONE / ICX_FILTER_HELMHOLTZ */
  double filter_radius;
  int projection;      /* 1: beta = min(2^(k/40), 16) with tanh projection */
  int form;            /* ICX_SOLVE_* for the structural solve */
  double tol;          /* structural solve tolerance */
  int maxit;
  double helm_tol;     /* Helmholtz solve tolerance */
  double xmin, xmax, move;
} icx_opt_params;

typedef struct {
  int step;
  double objective;    /* compliance */
  double constraint;   /* const_h[0] as the driver stores it */
  double volume;       /* sum rho_hat v */
  int cg_iterations;
  int update_steps;    /* OC bisections or MMA / GCMMA dual Newton steps */
  double seconds;      /* wall time of the design iteration */
  int inner_loops;     /* GCMMA inner updates this step */
  int trial_cg_iterations; /* GCMMA: PCG iterations of the trial-design solves */
} icx_iter_record;

icx_status icx_opt_init(icx_job *job, const icx_opt_params *par, const double *x0);
icx_status icx_opt_step(icx_job *job, icx_iter_record *rec);
/* icx_opt_step in two halves around a solve that the caller runs, e.g. through
 * icx_submit: begin = filter, projection, design -> moduli (opt_macro_oc.cpp:87-101,
 * opt_mma_Helmholtz.cpp:86-98); end = objective, volume, sensitivity, chain rule,
 * filter adjoint and the OC / MMA update (opt_macro_oc.cpp:104-165,
 * opt_mma_Helmholtz.cpp:100-158) on the job's current displacement. */
icx_status icx_opt_step_begin(icx_job *job);
icx_status icx_opt_step_end(icx_job *job, const icx_solve_info *solve, icx_iter_record *rec);
icx_status icx_opt_get_design(const icx_job *job, double *x);

/* ---- batched / serving -------------------------------------------------------------
 * No reference counterpart: the reference runs one process per job.  Submitted
 * solves of any jobs of a context advance together: each launch of a PCG kernel phase
 * covers every job whose solve is running in that precision, so jobs of uneven sizes
 * share the device; each job's result is bit-identical to its own icx_solve of the
 * same form.  The solve is the reference's Jacobi-PCG (solvers.cpp:64-92) in the mixed
 * or the fused FP64 form.
 *
 * icx_submit: start a mixed solve of the job's current design and load (set-up runs
 *   at once; the iterations run in icx_poll).  One submitted solve per job at a time.
 * icx_submit_form: the same with the form given: ICX_SOLVE_MIXED, or ICX_SOLVE_FUSED
 *   (FP64 throughout; the form for ill-conditioned jobs such as long quad4 grids, on
 *   which FP32 inner iterations need several times the FP64 iteration count or
 *   stall).  ICX_ERR_ARG for another form.
 * icx_poll: run up to `chunks` chunks of iterations of all submitted solves (with
 *   their FP64 residual replacements), returning as soon as at least one solve has
 *   finished; the finished jobs are written to finished[0 .. *nfinished).
 * icx_job_last_info: the info and status (ICX_OK / ICX_ERR_NOCONV) of the job's last
 *   finished submitted solve; its displacement is then the job's current one.
 * icx_solve_batch: submit all, poll until all have finished (forms ICX_SOLVE_MIXED and
 *   ICX_SOLVE_FUSED; the plain form, and fused solves of mesh or heat jobs, solve the
 *   jobs one after the other). */
icx_status icx_submit(icx_job *job, double tol, int maxit);
icx_status icx_submit_form(icx_job *job, int form, double tol, int maxit);
icx_status icx_poll(icx_ctx *ctx, int chunks, icx_job **finished, int cap, int *nfinished);
icx_status icx_job_last_info(const icx_job *job, icx_solve_info *info);
/* icx_cancel: stop the job's submitted solve and take it out of the batch (the other
 *   solves go on unchanged); its displacement is then undefined (no solution until the
 *   next solve) and icx_poll never returns it.  ICX_ERR_STATE if it has none running. */
icx_status icx_cancel(icx_job *job);
icx_status icx_solve_batch(icx_ctx *ctx, icx_job **jobs, int njobs, int form, double tol,
                           int maxit, icx_solve_info *infos);
/* device bytes a job holds (admission control) */
icx_status icx_job_footprint(const icx_job *job, size_t *bytes);

#ifdef __cplusplus
}
#endif
#endif
