// Internal declarations shared by the HIP kernels and the host side of libicx.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace icx {

// Structured node grid.  Nodes n = i + sx*(j + sy*k) (x fastest), elements
// e = ex + nx*(ey + ny*ez): the reference's MeshIcarat numbering.
struct Grid {
  int ndim;            // 2 (quad4) or 3 (hexa8)
  int nx, ny, nz;      // element divisions (nz = 1 for 2D)
  int sx, sy, sz;      // node counts per axis (sz = 1 for 2D)
  long long nn;        // nodes
  long long ne;        // elements
};

// Element operator in the per-element Haar basis.  For an axis-aligned box
// element the stiffness commutes with the three mirror reflections, so in the
// basis of sums/differences along each axis it is block diagonal with one
// 3x3 (hex8) or 2x2 (quad4) block per reflection character.  blk holds those
// blocks already scaled by 1/64 (hex8) or 1/16 (quad4), so that
// K0 u = H^T (blk (H u)) with H the unnormalised Haar butterfly.
template <typename T> struct ElasticOp {
  T blk[8][3][3];
  T kd[3];             // diagonal of K0 per displacement component
};

// Helmholtz element matrix (scalar), diagonal in the Haar basis.
template <typename T> struct HelmOp {
  T kf[8];             // Haar eigenvalues scaled by 1/64 (hex8) / 1/16 (quad4)
  T kdiag;             // diagonal entry of the element matrix
};

// Scalars of one PCG solve; written only by the last block of a kernel.
struct SolverState {
  double rhsNorm2;
  double threshold;    // max(tol^2 |b|^2, DBL_MIN)
  double absNew;       // r.z
  double absOld;
  double pAp;
  double alpha;        // step of the last matvec kernel
  double alpha_x;      // step still owed to x (mixed form, deferred update)
  double beta;
  double rr;           // |r|^2, recursive
  double rz_new;
  double true_rr;      // |b - K x|^2 in FP64 (mixed form)
  double rr_ref;       // residual norm^2 at the last replacement
  double delta2;       // replacement trigger (rr < delta2 * rr_ref)
  int done;            // 0 running, 1 converged, 2 maxit, 3 recursive converged, 4 replace
  int iter;            // Eigen iteration counter
  int maxit;
  int inner_since;     // inner iterations since the last replacement
  int replace_every;   // forced replacement period (0 = off)
  int pad;
  unsigned counters[8];
};

enum { RED_MAXBLOCKS = 4096 };

struct Reduce {
  double *partial;     // [2 * RED_MAXBLOCKS]
  unsigned *counter;
};

}  // namespace icx

// ---- kernel launchers (icx_kernels.hip) ------------------------------------------
namespace icx {

// y = K(E) x on the node grid (SoA, nc components).  mask: per-node bit c set =
// constrained (output forced to zero); may be null.  If red != null the dot
// x.y is reduced and, in the last block, st->pAp and st->alpha = absNew/pAp are
// written.
template <typename T>
void launch_elastic_matvec(const Grid &g, const ElasticOp<T> &op, const T *E, const uint8_t *mask,
                           const T *x, T *y, SolverState *st, Reduce red, hipStream_t s);

// Mixed/fused inner kernel A: p = invd*r + beta*p_old (beta from st), y = K p,
// x += alpha_x * p_old, dot p.y reduced into st->pAp / st->alpha.
template <typename T>
void launch_elastic_fusedA(const Grid &g, const ElasticOp<T> &op, const T *E, const uint8_t *mask,
                           const T *r, const T *invd, const T *pold, T *pnew, T *y, T *xacc,
                           SolverState *st, Reduce red, hipStream_t s);

void launch_helm_matvec(const Grid &g, const HelmOp<double> &op, const double *x, double *y,
                        SolverState *st, Reduce red, hipStream_t s);

// the Eigen loop after tmp = A p: p.tmp, x += a p, r -= a tmp, |r|^2, z = invd r,
// r.z, p = z + beta p -- one kernel each, in Eigen's order
void launch_eigen_iteration_tail(long long n, SolverState *st, double *x, double *r, double *p,
                                 const double *t, double *z, const double *invd, Reduce red,
                                 hipStream_t s);
// plain (Eigen-structured) vector kernels
template <typename T>
void launch_pcg_update(long long n, SolverState *st, T *x, T *r, const T *p, const T *t, T *z,
                       const T *invd, Reduce red, hipStream_t s);
template <typename T>
void launch_pcg_pupdate(long long n, SolverState *st, T *p, const T *z, hipStream_t s);
// fused/mixed kernel B: r -= alpha*Ap; rr, rz reduced; beta, absNew, iteration control.
template <typename T>
void launch_pcg_fusedB(long long n, SolverState *st, T *r, const T *Ap, const T *invd, Reduce red,
                       hipStream_t s);
// mixed replacement: xhi += xlo + alpha_x*p; xlo = 0
void launch_refine_x(long long n, SolverState *st, double *xhi, float *xlo, const float *p,
                     hipStream_t s);
// mixed replacement: rhi = b - Kxhi (given), r32 = rhi, true rr and rz reduced
void launch_refine_r(long long n, SolverState *st, const double *b, const double *Kx, float *r32,
                     const float *invd32, Reduce red, hipStream_t s);
// fused FP64 replacement-free variant keeps x in FP64: final flush x += alpha_x p
template <typename T>
void launch_flush_x(long long n, SolverState *st, T *x, const T *p, hipStream_t s);

// PCG set-up: r = b (copy), rr = b.b, threshold; z = invd*r; p = z; absNew = r.z
template <typename T>
void launch_pcg_init(long long n, SolverState *st, const T *b, T *r, T *p, const T *invd,
                     Reduce red, double tol, int maxit, int set_p, hipStream_t s);

// element modulus E = SIMP(design) in FP64 and FP32
void launch_simp(long long ne, double y1, double y2, double p, const double *design, double *E64,
                 float *E32, hipStream_t s);
// Jacobi inverse diagonal from E: invd[c][n] = 1/(kd_c * sum_{e ni n} E_e) (0 at constrained)
void launch_elastic_invdiag(const Grid &g, const double *kd, const double *E, const uint8_t *mask,
                            double *invd64, float *invd32, hipStream_t s);
void launch_helm_invdiag(const Grid &g, double kdiag, double *invd, hipStream_t s);

// layout conversions
void launch_dof_to_grid(const Grid &g, int nc, const int *dof, int numeq, const double *v,
                        const double *dirich /* nullable: per node*6 */, double *grid,
                        hipStream_t s);
void launch_grid_to_dof(const Grid &g, int nc, const int *dof, int numeq, const double *grid,
                        double *v, hipStream_t s);
void launch_node_values(const Grid &g, int nc, const int *dof, int numeq, const double *grid,
                        const double *dirich, double *nodeval, hipStream_t s);
void launch_convert_f64_f32(long long n, const double *a, float *b, hipStream_t s);
void launch_axpby(long long n, double a, const double *x, double b, double *y, hipStream_t s);
void launch_mul(long long n, const double *a, double *b, hipStream_t s);  // b *= a
void launch_mask_zero(const Grid &g, int nc, const uint8_t *mask, double *v, hipStream_t s);

// reductions producing a single double in out[0] (deterministic)
void launch_dot(long long n, const double *a, const double *b, double *out, Reduce red,
                hipStream_t s);

// sensitivity: dfds = -(dE/E) u_e^T K(E) u_e = -dE * u_e^T K0 u_e; dgds = vol/vscale
void launch_sensitivity(const Grid &g, const ElasticOp<double> &op, double y1, double y2, double p,
                        const double *design, const double *u, double vol, double vscale,
                        double *dfds, double *dgds, hipStream_t s);
// element energy sum_e E_e u_e^T K0 u_e (for the compliance identity)
void launch_energy(const Grid &g, const ElasticOp<double> &op, const double *E, const double *u,
                   double *energy, hipStream_t s);

// cone filter on the structured grid: stencil of offsets with weights
struct Stencil {
  int n;
  const int *off;      // n x 3
  const double *w;     // n
};
void launch_cone_aa(const Grid &g, Stencil st, double *aa, hipStream_t s);
void launch_cone_forward(const Grid &g, Stencil st, const double *aa, const double *in, double *out,
                         hipStream_t s);
void launch_cone_adjoint(const Grid &g, Stencil st, const double *aa, const double *in, double *out,
                         hipStream_t s);
void launch_project(long long n, double beta, double T, const double *in, double *out,
                    double *dout, hipStream_t s);

// Helmholtz rhs / element average
void launch_helm_rhs(const Grid &g, double scale, const double *rho, double *f, hipStream_t s);
void launch_helm_to_element(const Grid &g, double scale, const double *phi, double *out,
                            hipStream_t s);

// OC: x_new(lambda) and, for the cone filter, the fused filter+projection+volume
void launch_oc_candidate(long long n, double lambda, double move, double xmin, double xmax,
                         const double *x, const double *dfdx, const double *dgdx, double *xnew,
                         hipStream_t s);
void launch_volume(long long n, int project, double beta, double T, const double *rho,
                   const double *vol_or_null, double vol, double *out, Reduce red, hipStream_t s);
void launch_oc_volume_cone(const Grid &g, Stencil st, const double *aa, double lambda, double move,
                           double xmin, double xmax, const double *x, const double *dfdx,
                           const double *dgdx, int project, double beta, double T, double vol,
                           double *out, Reduce red, hipStream_t s);

// MMA (m = 1)
struct MMAVecs {
  double *low, *upp, *alpha, *beta, *p0, *q0, *p1, *q1, *xold1, *xold2;
};
void launch_mma_gensub(long long n, int iter, const MMAVecs &m, const double *xval,
                       const double *dfdx, const double *dgdx, double xmin, double xmax,
                       double move, double albefa, double asyminit, double asymdec, double asyminc,
                       double raa0, double xmamieps, double *bsum, Reduce red, hipStream_t s);
// x(lambda) and sums: out[0] = sum p1/(upp-x) + q1/(x-low), out[1] = sum PQ^2 df2
void launch_mma_dual(long long n, double lam, const MMAVecs &m, double *x, int want_hess,
                     double *out, Reduce red, hipStream_t s);

}  // namespace icx
