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
  int tile_depth;      // hexa8 operator: node layers per block along z (0 = automatic)
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
  T ikd[3];            // 1 / kd, rounded once to T (the per-node Jacobi factor)
};

// hexa8 scalar operators (heat and Helmholtz): the shared-element tile kernel, unless the
// context's tile depth (ICX_OPT_TILE_DEPTH) names a column depth of the per-node kernels
// (2, 4, 8 or 16)
inline bool scalar_tile_kernel(const Grid &g) {
  return g.ndim == 3 && g.tile_depth != 2 && g.tile_depth != 4 && g.tile_depth != 8 && g.tile_depth != 16;
}

// Jacobi factor of a box job's DOF from its node's inverse modulus sum s (s < 0 marks a
// zero diagonal, where the inverse is 1 as in Eigen's DiagonalPreconditioner) and the
// component's 1 / kd: one multiply, so every kernel that forms it gets the same bits.
template <typename T> __host__ __device__ inline T jacobi_inv(T s, T ik) { return s < T(0) ? T(1) : s * ik; }

// Helmholtz element matrix (scalar), diagonal in the Haar basis.
template <typename T> struct HelmOp {
  T kf[8];             // Haar eigenvalues scaled by 1/64 (hex8) / 1/16 (quad4)
  T kdiag;             // diagonal entry of the element matrix
};

enum { RED_MAXBLOCKS = 4096 };

// State of one PCG solve.  Written by the host between launches, or by block 0 of a
// kernel for fields no other block of that kernel reads.
struct SolverState {
  double rhsNorm2;
  double threshold;    // max(tol^2 |b|^2, DBL_MIN)
  double rr_ref;       // |r|^2 at the last residual replacement (mixed form)
  double delta2;       // replacement trigger: rr < delta2 * rr_ref (0 = off)
  double rr;           // |r|^2 at the stop
  double true_rr;      // |b - K x|^2 in FP64 at the last replacement
  double alpha_owed;   // step still owed to x at the stop (fused/mixed forms)
  int done;            // 0 running, 1 converged, 2 maxit, 3 recursive residual met, 4 replace
  int pending;         // fused/mixed forms: stop code found by kernel A at stop_iter, published
                       // as `done` by kernel B (no kernel writes a field that it reads itself:
                       // the waves of a block that starts late must not see a flag flip)
  int maxit;
  int iter_cur;        // iteration in flight: published by the first kernel of an iteration
  int iter;            // next iteration (= iterations completed), advanced by its last kernel
  int stop_iter;       // iteration index at which `done` was set
  int flush_iter;      // iteration whose owed x update was already applied
  double rz_hist[2];   // fused/mixed forms: r_i.z_i, stored by kernel A of iteration i in slot i & 1
  double alpha_hist[2];// fused/mixed forms: alpha_i, stored by kernel B of iteration i in slot i & 1
                       // (single-reduction form: r_i.z_i and alpha_i, stored by launch i + 1)
  // single-reduction form (FP32 hexa8 mixed solves, one kernel per iteration)
  int iter2[2];        // iteration index read by the launches of parity k from slot k; a launch
                       // of parity k that continues writes the next index to slot k ^ 1
  int resume_iter;     // first launch after a restart (set-up or residual replacement)
  int stop_slot;       // launch parity of the stopping launch
};

// Per-block partial sums of a PCG solve, double-buffered by iteration parity.
struct PcgPartials {
  double *PA;          // [2][RED_MAXBLOCKS]     p.Ap of kernel A (slot 0) / p.tmp of the plain loop
  double *PB;          // [2][RED_MAXBLOCKS][2]  (|r|^2, r.z) of kernel B / set-up / replacement (slot 0)
  double *PN;          // [2][RED_MAXBLOCKS]     plain loop |r|^2
  double *PZ;          // [2][RED_MAXBLOCKS]     plain loop r.z
  int nA, nB;          // partial counts written by the A-type and B-type kernels
};

// generic (non-iteration) reductions: partials + one-block finalisation
struct Reduce {
  double *partial;     // [NV][RED_MAXBLOCKS]
};

// Optional event pair bound to one kernel launch (hipExtLaunchKernelGGL): the events
// take the kernel's own start and end timestamps.  Null: plain launch.
struct LaunchEvents {
  hipEvent_t start, stop;
};

// Unsigned division by a run-time constant d >= 1 for dividends n < 2^29 (DOF, node and
// layer indices: at most 3 * 2^27 DOFs): q = (n * m) >> s with l = ceil(log2 d),
// s = 29 + l, m = ceil(2^s / d).  n m / 2^s = n/d + n (m d - 2^s) / (d 2^s), and
// 0 <= m d - 2^s < d with n < 2^29 bounds the second term below 2^-l <= 1/d, which
// the fractional part of n/d (at most 1 - 1/d) cannot absorb into the next integer:
// the floor is exact (Granlund-Montgomery).  d > 2^(l-1) gives m <= 2^30: a 32-bit
// multiplier, and a 32 x 32 -> 64-bit product.
struct Magic {
  uint32_t m;
  int s;
};
inline Magic magic_u29(long long d) {
  int l = 0;
  while ((1LL << l) < d) ++l;
  Magic r;
  r.s = 29 + l;
  r.m = (uint32_t)(((1ULL << r.s) + (unsigned long long)d - 1) / (unsigned long long)d);
  return r;
}
__host__ __device__ inline uint32_t udiv(uint32_t n, Magic g) {
  return (uint32_t)(((unsigned long long)n * (unsigned long long)g.m) >> g.s);
}

// Seam layers of the FP32 hexa8 kernel A (seam split along z): node layers k = TK s,
// 0 < k <= kmax, receive their lower half in y2.  kmax = 0: no seams.
struct SeamGeom {
  Magic nn, L, TK;              // divisors: nodes, nodes per layer, tile depth
  uint32_t nnodes, layer, tk, kmax;
};

// Seam flags of the nw consecutive DOFs i0 .. i0+nw-1 (SoA planes of sg.nnodes nodes,
// node layers of sg.layer >= 4 nodes, nw <= 4): bit w set when DOF i0+w lies on a seam
// layer k = TK s, 0 < k <= kmax.  The run crosses at most one layer boundary; a DOF in
// the next component plane is on layer 0.
__host__ __device__ inline uint32_t seam_flags(uint32_t i0, int nw, const SeamGeom &sg) {
  const uint32_t c = udiv(i0, sg.nn);
  const uint32_t n = i0 - c * sg.nnodes;
  const uint32_t k = udiv(n, sg.L);
  const uint32_t rem = n - k * sg.layer;
  const uint32_t kr = k - udiv(k, sg.TK) * sg.tk;
  const bool s0 = kr == 0 && k > 0 && k <= sg.kmax;
  const bool s1 = kr + 1 == sg.tk && k + 1 <= sg.kmax;
  uint32_t f = 0;
  for (int w = 0; w < nw; ++w) {
    if (n + w >= sg.nnodes) continue;
    f |= (uint32_t)(rem + w >= sg.layer ? s1 : s0) << w;
  }
  return f;
}

// Seam flags of the nw consecutive entries i0 .. i0+nw-1 (nw <= 4) of a node-interleaved
// vector (entry 3 n + c of node n): bit w set when the node of entry i0 + w lies on a seam
// layer.  The run spans at most two nodes.
__host__ __device__ inline uint32_t seam_flags_aos(uint32_t i0, int nw, const SeamGeom &sg) {
  const uint32_t n0 = i0 / 3, r = i0 - 3 * n0;
  const uint32_t f = seam_flags(n0, 2, sg);  // nodes n0 and n0 + 1 (node indices, plane 0)
  uint32_t out = 0;
  for (int w = 0; w < nw; ++w) out |= ((f >> ((r + (uint32_t)w) / 3)) & 1u) << w;
  return out;
}

// ---- structured operators (icx_operator.hip) -------------------------------------
// Arguments of one launch of an elastic operator kernel (one job).
template <typename T> struct OpArgs {
  Grid g;
  int TK, tiles_x, tiles_y, tiles_z;
  ElasticOp<T> op;
  const T *E;
  const uint8_t *mask;
  const T *in;        // MODE 0: x; MODE 1: r
  const T *invd;      // MODE 1 quad4: Jacobi inverse per DOF
  const T *invs;      // MODE 1 hexa8: per node 1 / (sum of adjacent moduli), -1 for 0
  T *P0, *P1;         // MODE 1: direction buffers, p_i lives in P[i & 1]
  T *y;               // K x (MODE 0) / K p_i (MODE 1)
  T *y2;              // FP32 hexa8 MODE 1: lower halves of the seam layers of K p_i
  T *x;               // MODE 1: solution accumulator (deferred update)
  SolverState *st;    // nullable in MODE 0
  PcgPartials P;      // MODE 1
};

// A batched launch runs one kernel phase for several jobs: global block b belongs to
// job j with off[j] <= b < off[j+1] (ranges padded to multiples of 8, so every job keeps
// the XCD placement of its own launch); each job computes exactly what its own launch
// would.
template <typename A> struct Batch {
  const A *args;
  const int *off;     // njobs + 1 entries
  int njobs;
};

// y = K(E) x on the node grid (SoA).  mask: per-node bit c = constrained (output
// forced to zero); may be null.  st may be null; otherwise the kernel exits at once
// when st->done is set.
template <typename T>
void launch_elastic_matvec(const Grid &g, const ElasticOp<T> &op, const T *E, const uint8_t *mask,
                           const T *x, T *y, const SolverState *st, hipStream_t s,
                           const LaunchEvents *ev = nullptr);

// y = K x on the node grid with the dense unit-modulus element stiffness K0 (24 x 24
// floats, row-major, DOF 3 q + c) contracted on the matrix cores; hexa8 only, FP32.
void launch_elastic_matvec_mfma(const Grid &g, const float *K0, const float *E, const uint8_t *mask, const float *x,
                                float *y, hipStream_t s, const LaunchEvents *ev = nullptr);

// Fused kernel A of iteration i = st->iter (published as st->iter_cur): scalars from the partials,
// p_i = invd r + beta_i p_{i-1} (P[i&1] <- ..., P[(i-1)&1] read), Ap_i = K p_i,
// x += alpha_{i-1} p_{i-1}, partial p_i.Ap_i -> PA[i&1].  FP32 hexa8 (seam split): the
// seam layers of Ap_i are completed by kernel B as Ap + Ap2 (fusedA_seams gives them);
// Ap2 is touched on the seam layers only.
template <typename T>
void launch_elastic_fusedA(const Grid &g, const ElasticOp<T> &op, const T *E, const uint8_t *mask,
                           const T *r, const T *invd, const T *invs, T *P0, T *P1, T *Ap, T *Ap2, T *x,
                           SolverState *st, PcgPartials P, hipStream_t s, const LaunchEvents *ev = nullptr);
// kernel A's arguments for one job; returns its block count
template <typename T>
int fusedA_args(const Grid &g, const ElasticOp<T> &op, const T *E, const uint8_t *mask, const T *r, const T *invd,
                const T *invs, T *P0, T *P1, T *Ap, T *Ap2, T *x, SolverState *st, PcgPartials P, OpArgs<T> &a);
// the seam layers of a kernel A launch (kmax = 0 when it has none)
template <typename T> SeamGeom fusedA_seams(const OpArgs<T> &a);
// kernel A for a batch of jobs of one dimension (args and off in device memory)
template <typename T>
void launch_fusedA_batch(int ndim, const OpArgs<T> *args, const int *off, int njobs, int total_blocks,
                         hipStream_t s);
// number of blocks of kernel A for this grid (its partial count)
template <typename T>
int elastic_fusedA_blocks(const Grid &g);

// ---- unstructured-mesh elastic operator (icx_mesh.hip) -------------------------------
// Mesh of one element type (ne = 4 quad4 / 8 hexa8, nc = 2 / 3 components): connectivity
// node-of-element-major ([a][e]), node coordinates ([d][node]), the unit-modulus material
// (normal diagonal d0, normal off-diagonal d1, shear d2), node adjacency in CSR (entries
// e * ne + a, increasing e), and a scratch of ne * nc doubles per element.  Element
// matrices are integrated at the Gauss points in every pass, never stored.
struct MeshOp {
  int ne, nc;
  long long nn, nelx;
  const int *conn;
  const double *xyz;
  double d0, d1, d2;
  // element outputs in node order: node n's (element, corner) outputs are the rows
  // adj_ptr[n] .. adj_ptr[n + 1] - 1 of ye (nc values each); slot[e * ne + a] is the row
  // of element e's corner a
  const int *adj_ptr, *slot;
  double *ye;
};
// y = K(E) x on the node vectors (SoA, nc components), zero on constrained components;
// st may be null, otherwise both passes exit at once when st->done is set
void launch_mesh_matvec(const MeshOp &m, const double *E, const uint8_t *mask, const double *x, double *y,
                        const SolverState *st, hipStream_t s, const LaunchEvents *ev = nullptr);
// diagonal of K(E), or its Jacobi inverse (1 where the diagonal is zero or constrained)
void launch_mesh_diag(const MeshOp &m, const double *E, const uint8_t *mask, double *out, bool inv,
                      hipStream_t s);
// compliance sensitivity -dE/drho u_e^T Ke0 u_e and dv = vol_e / vscale per element
void launch_mesh_sensitivity(const MeshOp &m, double y1, double y2, double p, const double *design,
                             const double *u, const double *vol, double vscale, double *dfds, double *dgds,
                             hipStream_t s);
// E_e u_e^T Ke0_e u_e
void launch_mesh_energy(const MeshOp &m, const double *E, const double *u, double *energy, hipStream_t s);
// element von Mises stress of the Gauss-point mean strain (u: node values, SoA)
void launch_mesh_mises(const MeshOp &m, const double *E, const double *u, double *out, hipStream_t s);
// Fused (T = double) and mixed (T = float, the inner iteration) forms on a mesh job: one
// PCG iteration (direction and deferred x step in place, element pass, gather with the
// p.Ap partials into P.PA) up to kernel B, which the caller launches; P.nA must be
// mesh_gather_blocks(m).  ev times the element pass.
int mesh_gather_blocks(const MeshOp &m);
template <typename T>
void launch_mesh_pcg_iteration(const MeshOp &m, SolverState *st, const T *E, const uint8_t *mask, const T *r,
                               const T *invd, T *p, T *Ap, T *x, const PcgPartials &P, hipStream_t s,
                               const LaunchEvents *ev = nullptr);
// Helmholtz filter on a mesh job, one scalar per node (rh2 = (R / (2 sqrt 3))^2):
// y = K x (both passes exit at once when st->done is set; st may be null), the Jacobi
// inverse of diag K (1 on a zero diagonal), the right-hand side sum_e rho_e int N / div,
// and the element values sum_ip N(ip) . phi_e / div
void launch_mesh_helm_matvec(const MeshOp &m, double rh2, const double *x, double *y, const SolverState *st,
                             hipStream_t s);
void launch_mesh_helm_invdiag(const MeshOp &m, double rh2, double *invd, hipStream_t s);
void launch_mesh_helm_rhs(const MeshOp &m, double div, const double *rho, double *f, hipStream_t s);
void launch_mesh_helm_to_element(const MeshOp &m, double div, const double *phi, double *out, hipStream_t s);

// ---- single-reduction PCG iteration, FP32 hexa8 (icx_sr.hip) -------------------------
// One kernel per iteration of the mixed form's Jacobi-PCG: launch k takes the step of
// iteration k-1 (alpha = rz/pi, r -= alpha Ap, x += alpha p), forms z = invd r and
// p = z + beta p with beta from the expansion of the new r.z in quantities the previous
// launch reduced (rz - 2 alpha Ap.z + alpha^2 Ap.invd Ap), applies the operator and
// writes the partials (r.z, |r|^2, p.Ap, Ap.z, Ap.invd Ap) of iteration k.  r, p and Ap
// alternate between two buffers by launch parity (neighbouring blocks read them at the
// halo); x is updated in place.
struct SrVectors {
  float *r[2], *p[2], *ap[2];
  float *x;
  const float *invd, *E;
  const uint8_t *mask;
  double *part[2];     // [RED_MAXBLOCKS][6]: the five partials per block, by launch parity
  const double *pb;    // (|r|^2, r.z) partials of the set-up or the replacement residual
  int npb;             // their count
};
// Arguments of one launch of the iteration kernel (one job, one launch parity).
struct SrArgs {
  Grid g;
  int TK, tiles_x, tiles_y, tiles_z;
  ElasticOp<float> op;
  const float *E;
  const uint8_t *mask;
  const float *invd;
  const float *r_in, *ap_in, *p_in;
  float *r_out, *ap_out, *p_out;
  float *x;
  SolverState *st;
  const double *part_in;
  double *part_out;
  const double *pb;
  int npb;
  int slot;
};
SrArgs sr_args(const Grid &g, const ElasticOp<float> &op, const SrVectors &v, SolverState *st, int slot);
// block count of the iteration kernel (its partial count)
int sr_blocks(const Grid &g);
// iteration of launch parity slot (chunks of iterations are an even number of launches)
void launch_sr_iteration(const Grid &g, const ElasticOp<float> &op, const SrVectors &v, SolverState *st,
                         int slot, hipStream_t s, const LaunchEvents *ev = nullptr);
// the iteration kernel for a batch of jobs (args of one launch parity and off in device memory)
void launch_sr_batch(const SrArgs *args, const int *off, int njobs, int total_blocks, hipStream_t s);
// restart after the set-up or a residual replacement (r[0] holds the new residual, pb its
// partials): iterate stop_iter formed again, p[0] = invd r + beta' p_{stop_iter-1},
// Ap[0] = K p[0], partials into part[0]; state: iter2[0] = resume_iter = stop_iter + 1.
// Does nothing when the state says done.
void launch_sr_restart(const Grid &g, const ElasticOp<float> &op, const SrVectors &v, SolverState *st,
                       hipStream_t s);

// hexa8 Haar-diagonal operator with the same eigenvalues kf[8] in every element (tile kernel)
void launch_helm_tile(const Grid &g, const double *kf, const double *x, double *y, const SolverState *st,
                      hipStream_t s);
void launch_helm_matvec(const Grid &g, const HelmOp<double> &op, const double *x, double *y,
                        const SolverState *st, hipStream_t s);

// ---- PCG vector kernels (icx_pcg.hip) -----------------------------------------------
// number of blocks of the B-type streaming kernels for n entries (their partial count)
int pcg_stream_blocks(long long n);
// set-up: r = b; partials (b.b, r.z) -> PB[1] (the slot of "iteration -1"); then a
// one-block kernel fills rhsNorm2, threshold, done
template <typename T>
void launch_pcg_setup(long long n, SolverState *st, const T *b, T *r, const T *invd, PcgPartials P,
                      double tol, int maxit, double delta2, hipStream_t s);
// Arguments of kernel B for one job (its grid is P.nB blocks).
template <typename T> struct BArgs {
  long long n;
  SolverState *st;
  T *r;
  const T *Ap;
  const T *invd;
  PcgPartials P;
  const T *Ap2;       // lower halves of kernel A's seam layers (read where sg says)
  SeamGeom sg;
  // box elastic jobs: the Jacobi factor per node (jacobi_inv(invs[node], ikd[c]), as
  // kernel A forms it) instead of invd; nn nodes, n = 3 nn.  Null: invd per entry.
  const T *invs;
  T ikd[3];
  long long nn;
};
// copies the states *src[j] of n jobs to dst[j]
void launch_gather_states(SolverState *const *src, SolverState *dst, int n, hipStream_t s);
// kernel B for a batch of jobs (args and off in device memory)
template <typename T>
void launch_pcg_fusedB_batch(const BArgs<T> *args, const int *off, int njobs, int total_blocks, hipStream_t s);
// fused kernel B of iteration i: alpha_i = rz_i / pAp_i; r -= alpha_i Ap; (|r|^2, r.z) -> PB[i&1]
// Ap2 / sg: seam layers of the FP32 hexa8 kernel A (sg.kmax = 0: none).
// invs / ikd (box elastic jobs, n = 3 nn): the Jacobi factor per node, read once for the
// three components of each node, instead of invd.  Worth it where the iteration's
// vectors outgrow the Infinity Cache and kernel B streams from HBM (measured: -1.6 us
// per launch at 192x96x96, +0.1 us at 96x48x48), so the callers pass invs for jobs of
// at least B_NODE_MIN_NODES nodes.
constexpr long long B_NODE_MIN_NODES = 1LL << 20;
template <typename T>
void launch_pcg_fusedB(long long n, SolverState *st, T *r, const T *Ap, const T *invd, PcgPartials P,
                       const T *Ap2, const SeamGeom &sg, hipStream_t s, const LaunchEvents *ev = nullptr,
                       const T *invs = nullptr, const T *ikd = nullptr);
// fused FP64 form at the stop: x += alpha_owed p_{stop-1}
void launch_pcg_flush(long long n, SolverState *st, double *x, const double *P0, const double *P1,
                      hipStream_t s);
// Mixed form replacement.  The FP64 vectors (xhi, b, Kx) are component planes; the FP32
// ones (xlo, P0, P1, r32, invd32) are too when nn_aos = 0, else node-interleaved over
// nn_aos nodes of nc components (entry nc n + c is FP64 entry c nn_aos + n).
// xhi += xlo + alpha_owed p_{stop-1}; xlo = 0
void launch_refine_x(long long n, SolverState *st, double *xhi, float *xlo, const float *P0,
                     const float *P1, long long nn_aos, int nc, hipStream_t s);
// r32 = b - Kx (FP64), (|r|^2, r.z) -> PB[(stop-1)&1], then a one-block kernel decides
// converged / continue and resets the recurrence
void launch_refine_r(long long n, SolverState *st, const double *b, const double *Kx, float *r32,
                     const float *invd32, PcgPartials P, long long nn_aos, int nc, hipStream_t s);

// the Eigen ConjugateGradient loop after tmp = A p, one kernel per vector expression
// (p.tmp, x += a p, r -= a tmp, |r|^2, z = M^-1 r, r.z, p = z + beta p)
void launch_eigen_iteration_tail(long long n, SolverState *st, double *x, double *r, double *p,
                                 const double *t, double *z, const double *invd, PcgPartials P,
                                 hipStream_t s);
// set-up of the Eigen loop: r = b, p = z = invd r; partials (b.b) -> PN[1], (r.z) -> PZ[1]
void launch_eigen_setup(long long n, SolverState *st, const double *b, double *r, double *p,
                        const double *invd, PcgPartials P, double tol, int maxit, hipStream_t s);

// ---- generic reductions ---------------------------------------------------------------
// out[k] = sum of partial[k][0..nblocks) (one block, fixed order)
void launch_finalize(const Reduce &red, int nblocks, int nv, double *out, hipStream_t s);
// dot product a.b -> out[0]
void launch_dot(long long n, const double *a, const double *b, double *out, Reduce red, hipStream_t s);

// ---- element moduli, diagonals, layout (icx_misc.hip) -------------------------------
void launch_simp(long long ne, double y1, double y2, double p, const double *design, double *E64,
                 float *E32, hipStream_t s);
// Jacobi inverse diagonal from E: invd[c][n] = 1/(kd_c * sum_{e ni n} E_e) (0 at constrained)
void launch_elastic_invdiag(const Grid &g, const double *kd, const double *E, const uint8_t *mask,
                            double *invd64, float *invd32, hipStream_t s);
// Jacobi preconditioner of the fused and mixed forms on a box job, in T: per node
// invs[n] = 1 / S_n (S_n = sum of the moduli of the elements around node n; -1 where
// S_n = 0) and per DOF invd[c][n] (aos: invd[3 n + c]) = jacobi_inv(invs[n], op.ikd[c])
// (0 at constrained DOFs).  Kernel A of a hexa8 job forms the same factor from invs (one
// word per node instead of one per DOF); kernel B and the set-up read invd.
template <typename T>
void launch_elastic_invdiag_node(const Grid &g, const ElasticOp<T> &op, const double *E, const uint8_t *mask,
                                 T *invs, T *invd, hipStream_t s, bool aos = false);
void launch_helm_invdiag(const Grid &g, double kdiag, double *invd, hipStream_t s);
void launch_dof_to_grid(const Grid &g, int nc, const int *dof, int numeq, const double *v,
                        const double *dirich /* nullable: per node*6 */, double *grid,
                        hipStream_t s);
void launch_grid_to_dof(const Grid &g, int nc, const int *dof, int numeq, const double *grid,
                        double *v, hipStream_t s);
void launch_node_values(const Grid &g, int nc, const int *dof, int numeq, const double *grid,
                        const double *dirich, double *nodeval, hipStream_t s);
void launch_convert_f64_f32(long long n, const double *a, float *b, hipStream_t s);
// b[nc n + c] = a[c nn + n] for nn nodes of nc components (component planes in FP64 to a
// node-interleaved FP32 vector)
void launch_convert_f64_f32_aos(long long nn, int nc, const double *a, float *b, hipStream_t s);
// b[nc n + c] = a[c nn + n] (to node-interleaved) and back, FP64, nc components per node
void launch_soa_to_aos_f64(long long nn, int nc, const double *a, double *b, hipStream_t s);
void launch_aos_to_soa_f64(long long nn, int nc, const double *a, double *b, hipStream_t s);
void launch_convert_f32_f64(long long n, const float *a, double *b, hipStream_t s);
void launch_axpby(long long n, double a, const double *x, double b, double *y, hipStream_t s);
void launch_mul(long long n, const double *a, double *b, hipStream_t s);  // b *= a
void launch_mask_zero(const Grid &g, int nc, const uint8_t *mask, double *v, hipStream_t s);

// sensitivity: dfds = -(dE/E) u_e^T K(E) u_e = -dE * u_e^T K0 u_e; dgds = vol/vscale
void launch_sensitivity(const Grid &g, const ElasticOp<double> &op, double y1, double y2, double p,
                        const double *design, const double *u, double vol, double vscale,
                        double *dfds, double *dgds, hipStream_t s);
// element energy E_e u_e^T K0 u_e (for the compliance identity)
void launch_energy(const Grid &g, const ElasticOp<double> &op, const double *E, const double *u,
                   double *energy, hipStream_t s);

// output step ---------------------------------------------------------------------
// node.val as a grid vector: u (free DOFs, zero where constrained) + the Dirichlet
// grid (prescribed values, zero on free DOFs; nullable), |v| < 1e-15 flushed
void launch_full_displacement(long long n, const double *u, const double *dirich, double *out,
                              hipStream_t s);
// material of the element-mean stress: state (ICX_PLANE_STRESS / _STRAIN / SOLID3D),
// Poisson ratio, element edge lengths
struct MisesMat {
  int state;
  double nu;
  double h[3];
};
// stress constraint of quad4 box jobs: qp relaxation, p-norm, SIMP (y1, y2, p)
struct StressPar {
  double qp, pn, slim, y1, y2, p;
};
// sig [3][ne], vm, sqp per element; psum = sum (sqp / slim)^pn
void launch_stress_eval(const Grid &g, const MisesMat &m, const StressPar &sp, const double *design,
                        const double *E, const double *u, double *sig, double *vm, double *sqp, double *psum,
                        Reduce red, hipStream_t s);
// f12 per element, fe [8][ne] scratch, load = the gathered adjoint load ([2][nn], masked)
void launch_stress_adjoint_load(const Grid &g, const MisesMat &m, const StressPar &sp, double dpndvm,
                                const double *design, const double *E, const double *sig, const double *vm,
                                const double *sqp, double *f12, double *fe, const uint8_t *mask, double *load,
                                hipStream_t s);
void launch_stress_sens(const Grid &g, const ElasticOp<double> &op, const StressPar &sp, const double *design,
                        const double *u, const double *lam, const double *f12, double *dgds, hipStream_t s);
// element-mean von Mises stress from the full displacement grid and E per element
void launch_mises(const Grid &g, const MisesMat &m, const double *E, const double *u, double *out,
                  hipStream_t s);
// node-major copy (3 per node, components >= nc zero) of a grid vector, flushed
void launch_grid_to_nodes(const Grid &g, int nc, const double *grid, double *out, hipStream_t s);

// heat conduction (src/problem/heat.hpp, heat.cpp): one DOF per node, element matrix
// K_e = w_e K1 + (1 - w_e) K2 with w_e = design_e^p (Heat::makeSIMPDe); K1, K2 are the
// conductivity matrices of materials 1 and 2, nq x nq row-major in the corner order
// q = a + 2b + 4c
struct HeatOp {
  int nq;
  double K1[64], K2[64];
  // haar = 1: K1 and K2 are diagonal in the element's Haar basis (diagonal tensors on a
  // box element), with eigenvalues kf1 / kf2 scaled as the Helmholtz operator's kf
  int haar;
  double kf1[8], kf2[8];
};
void launch_heat_weight(long long ne, double p, const double *design, double *w, hipStream_t s);
// y = K x (mask: bit 0 = constrained, output forced to zero; may be null); st may be null
void launch_heat_matvec(const Grid &g, const HeatOp &op, const double *w, const uint8_t *mask, const double *x,
                        double *y, const SolverState *st, hipStream_t s, const LaunchEvents *ev = nullptr);
// inverse diagonal (1 where the diagonal is 0, 0 on constrained DOFs)
void launch_heat_invdiag(const Grid &g, const HeatOp &op, const double *w, const uint8_t *mask, double *invd,
                         hipStream_t s);
// dfds = -p design^(p-1) T_e^T (K1 - K2) T_e (dK holds K1 - K2 in K1), dgds = vol / vscale
void launch_heat_sensitivity(const Grid &g, const HeatOp &dK, double p, const double *design, const double *u,
                             double vol, double vscale, double *dfds, double *dgds, hipStream_t s);

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

// OC bisection state on the device (OCSolver.hpp:20-68): the bracket [l0, l1], the
// lambda of the last evaluated step, the volume target g0, the step count and the stop
// flag.  The OC kernels below take it (nullable): with it, a kernel evaluates
// lambda = (l1 + l0) / 2 and leaves at once when done is set, so the host launches the
// steps in rounds and reads the state once per round instead of once per step.
struct OcBisect {
  double l0, l1, lambda, g0;
  int counter, done;
};
void launch_oc_init(OcBisect *b, double g0, hipStream_t s);
// one step's decision on the volume gk[0]: the reference's bracket update, counter and
// loop condition ((l1 - l0) / (l1 + l0) > 1e-5 and counter < 100)
void launch_oc_decide(OcBisect *b, const double *gk, hipStream_t s);

// OC: x_new(lambda) and, for the cone filter, the fused filter+projection+volume
// (lambda taken from bs when bs is given: the last evaluated step's, bs->lambda, for
// the candidate; the step's midpoint for the volume kernels)
void launch_oc_candidate(long long n, double lambda, double move, double xmin, double xmax,
                         const double *x, const double *dfdx, const double *dgdx, double *xnew,
                         hipStream_t s, const OcBisect *bs = nullptr, bool mid = false);
// sum_i proj(rho_i) * vol -> out[0]
// sum_e (projected) rho_e v_e with v_e = volv[e] when volv is given, else vol
void launch_volume(long long n, int project, double beta, double T, const double *rho, double vol,
                   const double *volv, double *out, Reduce red, hipStream_t s);
// cone filter over explicit neighbour lists (CSR rows in increasing neighbour id, w = r - d)
struct ConeCsr {
  const int *ptr, *id;
  const double *w;
};
void launch_cone_csr(long long n, ConeCsr f, const double *aa, const double *in, double *out, int adjoint,
                     hipStream_t s);
// Filter::sensitivityFilter over the cone stencil (box volume vol) or the CSR rows
// (per-element volumes volv)
void launch_cone_sensitivity(const Grid &g, Stencil st, const double *aa, const double *x, const double *df,
                             double vol, double *out, hipStream_t s);
void launch_cone_sensitivity_csr(long long n, ConeCsr f, const double *aa, const double *x, const double *df,
                                 const double *volv, double *out, hipStream_t s);
void launch_oc_volume_csr(long long n, ConeCsr f, const double *aa, double lambda, double move, double xmin,
                          double xmax, const double *x, const double *dfdx, const double *dgdx, int project,
                          double beta, double T, const double *vol, double *out, Reduce red, hipStream_t s,
                          const OcBisect *bs = nullptr);
void launch_oc_volume_cone(const Grid &g, Stencil st, const double *aa, double lambda, double move,
                           double xmin, double xmax, const double *x, const double *dfdx,
                           const double *dgdx, int project, double beta, double T, double vol,
                           double *out, Reduce red, hipStream_t s, const OcBisect *bs = nullptr);

// MMA (m = 1)
struct MMAVecs {
  double *low, *upp, *alpha, *beta, *p0, *q0, *p1, *q1, *xold1, *xold2;
};
// asymp = 1: new asymptotes and x history (MMA, GCMMA outer); 0: keep them (GCMMA inner).
// clamp = 1: MMA's alpha <= xmax, beta >= xmin. sums[0..1] = objective / constraint sums.
void launch_mma_gensub(long long n, int iter, int asymp, int clamp, const MMAVecs &m, const double *xval,
                       const double *dfdx, const double *dgdx, double xmin, double xmax,
                       double move, double albefa, double asyminit, double asymdec, double asyminc,
                       double raa0, double raa1, double xmamieps, double *sums, Reduce red, hipStream_t s);
// GCMMA approximations at xmma and raacof: out[0..2]
void launch_gcmma_approx(long long n, const MMAVecs &m, const double *xmma, const double *xval, double xmami,
                         double *out, Reduce red, hipStream_t s);
// x(lambda) and sums: out[0] = sum p1/(upp-x) + q1/(x-low), out[1] = sum PQ^2 df2.
// With d given, lambda is d->lam and the kernel leaves at once when d->done is set.
void launch_mma_dual(long long n, double lam, const MMAVecs &m, double *x, double *out, Reduce red,
                     hipStream_t s, const struct MmaDual *d = nullptr);

// The dual's interior-point Newton loop (MMASolver::DualSolve with m = 1, a = 0,
// MMASolver.cpp:400-512) kept on the device: launch_mma_dual evaluates x(lambda) and the
// two sums at d->lam, launch_mma_dual_advance takes them and runs the loop's control flow
// up to its next Newton step (or to the end: done), so the host queues (advance, eval)
// pairs in rounds and reads the state once per round instead of once per Newton step.
struct MmaDual {
  double lam, mu, epsi, err, b, cc, epsimin, y, z;
  int outloop, loop, total, in_inner, need_err, done;
};
void launch_mma_dual_init(MmaDual *d, double b, double cc, double epsimin, hipStream_t s);
void launch_mma_dual_advance(MmaDual *d, const double *sums, hipStream_t s);

}  // namespace icx
