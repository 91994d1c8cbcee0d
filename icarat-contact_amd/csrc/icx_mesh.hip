// Unstructured-mesh elastic operator of libicx (gfx950): explicit connectivity, element
// matrices never stored (MeshMDPA meshes, src/io/mesh_mdpa.hpp; the element of
// LinearElastic::makeKeFint, src/problem/elasticity.cpp:210-254).
//
// Every element pass integrates at the Gauss points from the element's node coordinates:
//   y_e = E_e sum_ip B_ip^T D (B_ip x_e) det J_ip
// (2 x 2 (x 2) points +-1/sqrt(3), weight 1), which is Ke0_e x_e without forming Ke0_e:
// per element it reads 8 node coordinates and values instead of a 300-word matrix, and
// spends ~5k FP64 operations on them (the part has FP64 arithmetic to spare, not
// bandwidth).  y = K(E) x is then two passes, deterministic (no atomics):
//   element pass: one thread per element, y_e written element-major ([e][i]);
//   node pass: y_n = sum over the (element, local node) pairs of node n, in increasing
//     element order (node adjacency in CSR), zero on constrained components.
// The Jacobi diagonal (the element pass writes diag Ke0_e) and the element energies
// u_e^T Ke0_e u_e = sum_ip eps^T D eps det J use the same integration.
#include "icx_device.h"

namespace icx {

namespace {

// the reference's local node order: corner signs of node a along x, y, z
__host__ __device__ constexpr int csign(int a, int d) {
  return d == 0 ? ((((a + 1) >> 1) & 1) ? 1 : -1) : d == 1 ? (((a >> 1) & 1) ? 1 : -1) : (((a >> 2) & 1) ? 1 : -1);
}

// Gauss point ip of the element and the Jacobian there: inv(J) and det(J).  X[a][k] are
// the node coordinates.  dN_a/dr_d = s_ad / 2^NC prod_{e != d} (1 + s_ae r_e).
template <int NE, int NC, typename T = double>
struct GaussPoint {
  T r[NC];
  T inv[NC][NC];
  T det;
  __device__ __forceinline__ T dr(int a, int d) const {
    T v = T(NC == 3 ? 0.125 : 0.25) * T(csign(a, d));
#pragma unroll
    for (int e = 0; e < NC; ++e)
      if (e != d) v *= T(1) + T(csign(a, e)) * r[e];
    return v;
  }
  // N_a = prod_d (1 + s_ad r_d) / 2^NC
  __device__ __forceinline__ T N(int a) const {
    T v = T(NC == 3 ? 0.125 : 0.25);
#pragma unroll
    for (int d = 0; d < NC; ++d) v *= T(1) + T(csign(a, d)) * r[d];
    return v;
  }
  // dN_a / dX_k
  __device__ __forceinline__ T dx(int a, int k) const {
    T v = T(0);
#pragma unroll
    for (int d = 0; d < NC; ++d) v = fma_t(inv[k][d], dr(a, d), v);
    return v;
  }
  __device__ __forceinline__ GaussPoint(int ip, const T (&X)[NE][NC]) {
    const T g = T(0.57735026918962576451);  // 1 / sqrt(3)
#pragma unroll
    for (int d = 0; d < NC; ++d) r[d] = T(csign(ip, d)) * g;
    T J[NC][NC];
#pragma unroll
    for (int d = 0; d < NC; ++d)
#pragma unroll
      for (int k = 0; k < NC; ++k) {
        T v = T(0);
#pragma unroll
        for (int a = 0; a < NE; ++a) v = fma_t(dr(a, d), X[a][k], v);
        J[d][k] = v;
      }
    if (NC == 2) {
      det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
      const T id = T(1) / det;
      inv[0][0] = J[1][1] * id; inv[0][1] = -J[0][1] * id;
      inv[1][0] = -J[1][0] * id; inv[1][1] = J[0][0] * id;
    } else {
      const T c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
      const T c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
      const T c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
      det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
      const T id = T(1) / det;
      inv[0][0] = c00 * id;
      inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
      inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
      inv[1][0] = c01 * id;
      inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
      inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
      inv[2][0] = c02 * id;
      inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
      inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
    }
  }
};

constexpr int voigt(int nc) { return nc == 3 ? 6 : 3; }

// engineering strain B u (xx, yy, zz, xy, yz, zx / xx, yy, xy)
template <int NE, int NC, typename T = double>
__device__ __forceinline__ void strain(const GaussPoint<NE, NC, T> &gp, const T (&u)[NE * NC],
                                       T (&eps)[voigt(NC)]) {
#pragma unroll
  for (int v = 0; v < voigt(NC); ++v) eps[v] = T(0);
#pragma unroll
  for (int a = 0; a < NE; ++a) {
    const T gx = gp.dx(a, 0), gy = gp.dx(a, 1);
    if (NC == 2) {
      const T ux = u[2 * a], uy = u[2 * a + 1];
      eps[0] = fma_t(gx, ux, eps[0]);
      eps[1] = fma_t(gy, uy, eps[1]);
      eps[2] = fma_t(gy, ux, fma_t(gx, uy, eps[2]));
    } else {
      const T gz = gp.dx(a, 2);
      const T ux = u[3 * a], uy = u[3 * a + 1], uz = u[3 * a + 2];
      eps[0] = fma_t(gx, ux, eps[0]);
      eps[1] = fma_t(gy, uy, eps[1]);
      eps[2] = fma_t(gz, uz, eps[2]);
      eps[3] = fma_t(gy, ux, fma_t(gx, uy, eps[3]));
      eps[4] = fma_t(gz, uy, fma_t(gy, uz, eps[4]));
      eps[5] = fma_t(gz, ux, fma_t(gx, uz, eps[5]));
    }
  }
}

// stress D eps of the unit-modulus material: normal diagonal d0, normal off-diagonal d1,
// shear d2 (isotropic 3D, plane stress and plane strain all have this pattern)
template <int NC, typename T = double>
__device__ __forceinline__ void stress(const MeshOp &m, const T (&eps)[voigt(NC)], T (&sig)[voigt(NC)]) {
  const T d0 = T(m.d0), d1 = T(m.d1), d2 = T(m.d2);
  if (NC == 2) {
    sig[0] = fma_t(d0, eps[0], d1 * eps[1]);
    sig[1] = fma_t(d1, eps[0], d0 * eps[1]);
    sig[2] = d2 * eps[2];
  } else {
    const T t = d1 * (eps[0] + eps[1] + eps[2]);
    const T dd = d0 - d1;
    sig[0] = fma_t(dd, eps[0], t);
    sig[1] = fma_t(dd, eps[1], t);
    sig[2] = fma_t(dd, eps[2], t);
    sig[3] = d2 * eps[3];
    sig[4] = d2 * eps[4];
    sig[5] = d2 * eps[5];
  }
}

// node coordinates of element e; in FP32 relative to its first node (formed in FP64), so
// the Jacobian, which depends on coordinate differences only, keeps FP32 precision
template <int NE, int NC, typename T = double>
__device__ __forceinline__ void gather_coords(const MeshOp &m, long long e, T (&X)[NE][NC]) {
  double base[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) base[c] = sizeof(T) == 4 ? m.xyz[c * m.nn + m.conn[e]] : 0.0;
#pragma unroll
  for (int a = 0; a < NE; ++a) {
    const long long n = m.conn[a * m.nelx + e];
#pragma unroll
    for (int c = 0; c < NC; ++c) X[a][c] = sizeof(T) == 4 ? T(m.xyz[c * m.nn + n] - base[c]) : T(m.xyz[c * m.nn + n]);
  }
}
template <int NE, int NC, typename T = double>
__device__ __forceinline__ void gather_values(const MeshOp &m, long long e, const T *x, T (&u)[NE * NC], bool flush) {
#pragma unroll
  for (int a = 0; a < NE; ++a) {
    const long long n = m.conn[a * m.nelx + e];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const T v = x[c * m.nn + n];
      u[a * NC + c] = flush ? T(flush15(v)) : v;
    }
  }
}

// y_e = E_e Ke0_e x_e (T = float: the mixed form's FP32 operator)
template <int NE, int NC, typename T = double>
__global__ void __launch_bounds__(256) k_mesh_element(MeshOp m, const T *E, const T *x, T *ye,
                                                      const SolverState *st) {
  constexpr int ND = NE * NC;
  if (st && st->done) return;
  for (long long e = blockIdx.x * 256LL + threadIdx.x; e < m.nelx; e += (long long)gridDim.x * 256) {
    T X[NE][NC], u[ND], f[ND];
    gather_coords<NE, NC>(m, e, X);
    gather_values<NE, NC>(m, e, x, u, false);
#pragma unroll
    for (int i = 0; i < ND; ++i) f[i] = T(0);
#pragma unroll 1  // one Gauss point at a time: the element's values stay in registers
    for (int ip = 0; ip < NE; ++ip) {
      const GaussPoint<NE, NC, T> gp(ip, X);
      T eps[voigt(NC)], sig[voigt(NC)];
      strain<NE, NC>(gp, u, eps);
      stress<NC>(m, eps, sig);
#pragma unroll
      for (int v = 0; v < voigt(NC); ++v) sig[v] *= gp.det;
      // B^T sigma
#pragma unroll
      for (int a = 0; a < NE; ++a) {
        const T gx = gp.dx(a, 0), gy = gp.dx(a, 1);
        if (NC == 2) {
          f[2 * a] = fma_t(gx, sig[0], fma_t(gy, sig[2], f[2 * a]));
          f[2 * a + 1] = fma_t(gy, sig[1], fma_t(gx, sig[2], f[2 * a + 1]));
        } else {
          const T gz = gp.dx(a, 2);
          f[3 * a] = fma_t(gx, sig[0], fma_t(gy, sig[3], fma_t(gz, sig[5], f[3 * a])));
          f[3 * a + 1] = fma_t(gy, sig[1], fma_t(gx, sig[3], fma_t(gz, sig[4], f[3 * a + 1])));
          f[3 * a + 2] = fma_t(gz, sig[2], fma_t(gy, sig[4], fma_t(gx, sig[5], f[3 * a + 2])));
        }
      }
    }
    const T Ee = E[e];
#pragma unroll
    for (int a = 0; a < NE; ++a) {
      T *out = ye + (long long)m.slot[e * NE + a] * NC;
#pragma unroll
      for (int c = 0; c < NC; ++c) out[c] = Ee * f[a * NC + c];
    }
  }
}

// E_e diag(Ke0_e) into ye: DOF (a, c) has B column entries g_c in the normal row of c and
// the other two g in the shear rows of c, so Ke_ii = sum_ip det J (d0 g_c^2 + d2 (|g|^2 - g_c^2))
template <int NE, int NC>
__global__ void __launch_bounds__(256) k_mesh_elem_diag(MeshOp m, const double *E, double *ye) {
  constexpr int ND = NE * NC;
  for (long long e = blockIdx.x * 256LL + threadIdx.x; e < m.nelx; e += (long long)gridDim.x * 256) {
    double X[NE][NC], dg[ND];
    gather_coords<NE, NC>(m, e, X);
#pragma unroll
    for (int i = 0; i < ND; ++i) dg[i] = 0.0;
#pragma unroll 1  // one Gauss point at a time: the element's values stay in registers
    for (int ip = 0; ip < NE; ++ip) {
      const GaussPoint<NE, NC> gp(ip, X);
#pragma unroll
      for (int a = 0; a < NE; ++a) {
        double g[NC], s2 = 0.0;
#pragma unroll
        for (int k = 0; k < NC; ++k) {
          g[k] = gp.dx(a, k);
          s2 = fma_t(g[k], g[k], s2);
        }
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          const double gc2 = g[c] * g[c];
          dg[a * NC + c] = fma_t(gp.det, fma_t(m.d0, gc2, m.d2 * (s2 - gc2)), dg[a * NC + c]);
        }
      }
    }
    const double Ee = E[e];
#pragma unroll
    for (int a = 0; a < NE; ++a) {
      double *out = ye + (long long)m.slot[e * NE + a] * NC;
#pragma unroll
      for (int c = 0; c < NC; ++c) out[c] = Ee * dg[a * NC + c];
    }
  }
}

template <int NE, int NC>
__global__ void __launch_bounds__(256) k_mesh_gather(MeshOp m, const double *ye, const uint8_t *mask, double *y,
                                                     const SolverState *st, int invert) {
  if (st && st->done) return;
  for (long long n = blockIdx.x * 256LL + threadIdx.x; n < m.nn; n += (long long)gridDim.x * 256) {
    double acc[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] = 0.0;
    for (int k = m.adj_ptr[n]; k < m.adj_ptr[n + 1]; ++k)
#pragma unroll
      for (int c = 0; c < NC; ++c) acc[c] += ye[(long long)k * NC + c];
    const uint32_t msk = mask ? mask[n] : 0u;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const double d = ((msk >> c) & 1) ? 0.0 : acc[c];
      y[c * m.nn + n] = invert ? (d != 0.0 ? 1.0 / d : 1.0) : d;
    }
  }
}

// ---- fused and mixed forms on mesh jobs (FP64 Jacobi-PCG; FP32 inner Jacobi-PCG with
// FP64 residual replacement) ------
// Per iteration: k_mesh_dir (the iteration scalars from kernel B's partials and the stop
// protocol of the structured kernel A, then p = invd r + beta p and x += alpha_{i-1}
// p_{i-1} in place: one thread per vector entry reads and writes it), the FP32 element
// pass, k_mesh_gather_dot (Ap = node sums, zero on constrained components, and the
// block partials of p.Ap), then kernel B (icx_pcg.hip).  A stop found by k_mesh_dir
// leaves p unchanged (p_{i-1}, the direction the replacement's owed x step reads).
template <typename T>
__global__ void __launch_bounds__(256) k_mesh_dir(long long n, SolverState *st, const T *r, const T *invd, T *p,
                                                  T *x, PcgPartials P) {
  __shared__ double s_sc[2 * 4];
  const SolverState ss = load_state(st);
  if (ss.done) return;  // uniform
  double rr, rz;
  block_total2<256>(P.PB, P.nB, s_sc, rr, rz);
  const int it = ss.iter;
  const IterScalars sc = iter_scalars(ss, it, rr, rz);
  const bool owed = it > 0 && ss.flush_iter != it;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->iter_cur = it;
    st->rz_hist[it & 1] = sc.rz;
    if (sc.stop) {
      st->pending = sc.stop;  // kernel B publishes it as `done`
      st->stop_iter = it;
      st->rr = sc.rr;
      st->alpha_owed = owed ? sc.alpha_prev : 0.0;
    }
  }
  if (sc.stop) return;
  const T beta = (T)sc.beta, ax = owed ? (T)sc.alpha_prev : T(0);
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const T po = p[i];
    x[i] = fma_t(ax, po, x[i]);
    p[i] = fma_t(beta, po, invd[i] * r[i]);
  }
}

template <int NC, typename T>
__global__ void __launch_bounds__(256) k_mesh_gather_dot(MeshOp m, const T *ye, const uint8_t *mask, const T *p, T *y,
                                                         const SolverState *st, double *PA) {
  __shared__ double s_red[4];
  if (st->done) return;  // uniform
  double dot = 0.0;
  for (long long n = blockIdx.x * 256LL + threadIdx.x; n < m.nn; n += (long long)gridDim.x * 256) {
    T acc[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] = T(0);
    for (int k = m.adj_ptr[n]; k < m.adj_ptr[n + 1]; ++k)
#pragma unroll
      for (int c = 0; c < NC; ++c) acc[c] += ye[(long long)k * NC + c];
    const uint32_t msk = mask ? mask[n] : 0u;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const T d = ((msk >> c) & 1) ? T(0) : acc[c];
      y[c * m.nn + n] = d;
      dot = fma_t((double)p[c * m.nn + n], (double)d, dot);
    }
  }
  dot = block_sum<256>(dot, s_red);
  if (threadIdx.x == 0) PA[blockIdx.x] = dot;
}

// u_e^T Ke0_e u_e = sum_ip eps^T D eps det J with |u| < 1e-15 flushed (the displacement
// the drivers' sensitivity reads), then the compliance sensitivity or E_e times it
template <int NE, int NC, bool SENS>
__global__ void __launch_bounds__(256) k_mesh_energy(MeshOp m, double y1, double y2, double p, const double *design,
                                                     const double *E, const double *u, const double *vol,
                                                     double vscale, double *dfds, double *dgds) {
  constexpr int ND = NE * NC;
  for (long long e = blockIdx.x * 256LL + threadIdx.x; e < m.nelx; e += (long long)gridDim.x * 256) {
    double scale;
    if (SENS) {
      const double dens = design[e];
      if (y2 >= y1) scale = -(p * pow(dens, p - 1.0) * (y2 - y1));
      else scale = -((y2 - y1) * p * pow(1.0 - dens, p - 1.0));
      if (dgds) dgds[e] = vol[e] / vscale;
    } else {
      scale = E[e];
    }
    double X[NE][NC], ue[ND];
    gather_coords<NE, NC>(m, e, X);
    gather_values<NE, NC>(m, e, u, ue, true);
    double en = 0.0;
#pragma unroll 1  // one Gauss point at a time: the element's values stay in registers
    for (int ip = 0; ip < NE; ++ip) {
      const GaussPoint<NE, NC> gp(ip, X);
      double eps[voigt(NC)], sig[voigt(NC)];
      strain<NE, NC>(gp, ue, eps);
      stress<NC>(m, eps, sig);
      double w = 0.0;
#pragma unroll
      for (int v = 0; v < voigt(NC); ++v) w = fma_t(eps[v], sig[v], w);
      en = fma_t(w, gp.det, en);
    }
    if (dfds) dfds[e] = scale * en;
  }
}

// Elastic::makeMisesStress (src/problem/elasticity.cpp:165-208): strain averaged over the
// Gauss points (weight 1 / ipmax), stress E_e D of it, mises = sqrt(s^T V s) with V the
// deviatoric form (shear terms weighted 3).  u: node values (flushed, Dirichlet values in).
template <int NE, int NC>
__global__ void __launch_bounds__(256) k_mesh_mises(MeshOp m, const double *E, const double *u, double *out) {
  constexpr int ND = NE * NC;
  for (long long e = blockIdx.x * 256LL + threadIdx.x; e < m.nelx; e += (long long)gridDim.x * 256) {
    double X[NE][NC], ue[ND], eb[voigt(NC)];
    gather_coords<NE, NC>(m, e, X);
    gather_values<NE, NC>(m, e, u, ue, false);
#pragma unroll
    for (int v = 0; v < voigt(NC); ++v) eb[v] = 0.0;
#pragma unroll 1
    for (int ip = 0; ip < NE; ++ip) {
      const GaussPoint<NE, NC> gp(ip, X);
      double eps[voigt(NC)];
      strain<NE, NC>(gp, ue, eps);
#pragma unroll
      for (int v = 0; v < voigt(NC); ++v) eb[v] = fma_t(eps[v], 1.0 / NE, eb[v]);
    }
    double sg[voigt(NC)];
    stress<NC>(m, eb, sg);
    const double Ee = E[e];
#pragma unroll
    for (int v = 0; v < voigt(NC); ++v) sg[v] *= Ee;
    double q;
    if (NC == 2) {
      q = sg[0] * sg[0] + sg[1] * sg[1] - sg[0] * sg[1] + 3.0 * sg[2] * sg[2];
    } else {
      q = sg[0] * sg[0] + sg[1] * sg[1] + sg[2] * sg[2] - sg[0] * sg[1] - sg[1] * sg[2] - sg[0] * sg[2] +
          3.0 * (sg[3] * sg[3] + sg[4] * sg[4] + sg[5] * sg[5]);
    }
    out[e] = sqrt(q);
  }
}

// ---- Helmholtz filter on the mesh (FilterHelmholtz, src/optimization/filter_helm.hpp:75-256):
// one scalar per node, element matrix sum_ip (N N^T + rh2 dN^T dN) det J with
// rh2 = (R / (2 sqrt 3))^2, never stored.  The outputs go through ye in node order and
// the scalar gather, as the elastic operator's.

// y_e = Ke_e x_e
template <int NE, int NC>
__global__ void __launch_bounds__(256) k_mesh_helm_element(MeshOp m, double rh2, const double *x, double *ye,
                                                           const SolverState *st) {
  if (st && st->done) return;
  for (long long e = blockIdx.x * 256LL + threadIdx.x; e < m.nelx; e += (long long)gridDim.x * 256) {
    double X[NE][NC], v[NE], f[NE];
    gather_coords<NE, NC>(m, e, X);
#pragma unroll
    for (int a = 0; a < NE; ++a) {
      v[a] = x[m.conn[a * m.nelx + e]];
      f[a] = 0.0;
    }
#pragma unroll 1
    for (int ip = 0; ip < NE; ++ip) {
      const GaussPoint<NE, NC> gp(ip, X);
      double nv = 0.0, gv[NC];
#pragma unroll
      for (int k = 0; k < NC; ++k) gv[k] = 0.0;
#pragma unroll
      for (int a = 0; a < NE; ++a) {
        nv = fma_t(gp.N(a), v[a], nv);
#pragma unroll
        for (int k = 0; k < NC; ++k) gv[k] = fma_t(gp.dx(a, k), v[a], gv[k]);
      }
#pragma unroll
      for (int k = 0; k < NC; ++k) gv[k] *= rh2;
#pragma unroll
      for (int a = 0; a < NE; ++a) {
        double t = gp.N(a) * nv;
#pragma unroll
        for (int k = 0; k < NC; ++k) t = fma_t(gp.dx(a, k), gv[k], t);
        f[a] = fma_t(t, gp.det, f[a]);
      }
    }
#pragma unroll
    for (int a = 0; a < NE; ++a) ye[m.slot[e * NE + a]] = f[a];
  }
}

// diag Ke_e = sum_ip (N_a^2 + rh2 |dN_a|^2) det J
template <int NE, int NC>
__global__ void __launch_bounds__(256) k_mesh_helm_diag(MeshOp m, double rh2, double *ye) {
  for (long long e = blockIdx.x * 256LL + threadIdx.x; e < m.nelx; e += (long long)gridDim.x * 256) {
    double X[NE][NC], f[NE];
    gather_coords<NE, NC>(m, e, X);
#pragma unroll
    for (int a = 0; a < NE; ++a) f[a] = 0.0;
#pragma unroll 1
    for (int ip = 0; ip < NE; ++ip) {
      const GaussPoint<NE, NC> gp(ip, X);
#pragma unroll
      for (int a = 0; a < NE; ++a) {
        double g2 = 0.0;
#pragma unroll
        for (int k = 0; k < NC; ++k) g2 = fma_t(gp.dx(a, k), gp.dx(a, k), g2);
        const double n = gp.N(a);
        f[a] = fma_t(fma_t(rh2, g2, n * n), gp.det, f[a]);
      }
    }
#pragma unroll
    for (int a = 0; a < NE; ++a) ye[m.slot[e * NE + a]] = f[a];
  }
}

// right-hand side f_n = sum over the elements of n of rho_e int N_a det J (divided by
// ipmax in the sensitivity mode: div = NE, else 1), filter_helm.hpp:216-219
template <int NE, int NC>
__global__ void __launch_bounds__(256) k_mesh_helm_rhs(MeshOp m, double div, const double *rho, double *ye) {
  for (long long e = blockIdx.x * 256LL + threadIdx.x; e < m.nelx; e += (long long)gridDim.x * 256) {
    double X[NE][NC], f[NE];
    gather_coords<NE, NC>(m, e, X);
#pragma unroll
    for (int a = 0; a < NE; ++a) f[a] = 0.0;
#pragma unroll 1
    for (int ip = 0; ip < NE; ++ip) {
      const GaussPoint<NE, NC> gp(ip, X);
#pragma unroll
      for (int a = 0; a < NE; ++a) f[a] = fma_t(gp.N(a), gp.det, f[a]);
    }
    const double r = rho[e];
#pragma unroll
    for (int a = 0; a < NE; ++a) ye[m.slot[e * NE + a]] = (r * f[a]) / div;
  }
}

// divideToElement (filter_helm.hpp:224-255): out_e = sum_ip N(ip) . phi_e, divided by
// ipmax (div = NE) in the forward mode
template <int NE, int NC>
__global__ void __launch_bounds__(256) k_mesh_helm_to_element(MeshOp m, double div, const double *phi, double *out) {
  for (long long e = blockIdx.x * 256LL + threadIdx.x; e < m.nelx; e += (long long)gridDim.x * 256) {
    double v[NE];
#pragma unroll
    for (int a = 0; a < NE; ++a) v[a] = phi[m.conn[a * m.nelx + e]];
    const double g = 0.57735026918962576451;
    double s = 0.0;
#pragma unroll 1
    for (int ip = 0; ip < NE; ++ip) {
      double t = 0.0;
#pragma unroll
      for (int a = 0; a < NE; ++a) {
        double n = NC == 3 ? 0.125 : 0.25;
#pragma unroll
        for (int d = 0; d < NC; ++d) n *= 1.0 + csign(a, d) * csign(ip, d) * g;
        t = fma_t(n, v[a], t);
      }
      s += t;
    }
    out[e] = s / div;
  }
}

int mesh_blocks(long long n) { return ew_blocks(n, 1, 4096); }

template <int NE, int NC, class K, class... A>
void mesh_scalar_pass(const MeshOp &m, K kern, double *y, bool invert, const SolverState *st, hipStream_t s,
                      const LaunchEvents *ev, A... args) {
  launch_ev(kern, mesh_blocks(m.nelx), 256, s, ev, m, args...);
  hipLaunchKernelGGL((k_mesh_gather<NE, 1>), dim3(mesh_blocks(m.nn)), dim3(256), 0, s, m, (const double *)m.ye,
                     (const uint8_t *)nullptr, y, st, (int)invert);
}

}  // namespace

void launch_mesh_matvec(const MeshOp &m, const double *E, const uint8_t *mask, const double *x, double *y,
                        const SolverState *st, hipStream_t s, const LaunchEvents *ev) {
  if (m.ne == 8) {
    launch_ev(k_mesh_element<8, 3>, mesh_blocks(m.nelx), 256, s, ev, m, E, x, m.ye, st);
    hipLaunchKernelGGL((k_mesh_gather<8, 3>), dim3(mesh_blocks(m.nn)), dim3(256), 0, s, m, (const double *)m.ye,
                       mask, y, st, 0);
  } else {
    launch_ev(k_mesh_element<4, 2>, mesh_blocks(m.nelx), 256, s, ev, m, E, x, m.ye, st);
    hipLaunchKernelGGL((k_mesh_gather<4, 2>), dim3(mesh_blocks(m.nn)), dim3(256), 0, s, m, (const double *)m.ye,
                       mask, y, st, 0);
  }
  ICX_CHECK_LAUNCH();
}

void launch_mesh_diag(const MeshOp &m, const double *E, const uint8_t *mask, double *out, bool inv,
                      hipStream_t s) {
  if (m.ne == 8) {
    hipLaunchKernelGGL((k_mesh_elem_diag<8, 3>), dim3(mesh_blocks(m.nelx)), dim3(256), 0, s, m, E, m.ye);
    hipLaunchKernelGGL((k_mesh_gather<8, 3>), dim3(mesh_blocks(m.nn)), dim3(256), 0, s, m, (const double *)m.ye,
                       mask, out, nullptr, (int)inv);
  } else {
    hipLaunchKernelGGL((k_mesh_elem_diag<4, 2>), dim3(mesh_blocks(m.nelx)), dim3(256), 0, s, m, E, m.ye);
    hipLaunchKernelGGL((k_mesh_gather<4, 2>), dim3(mesh_blocks(m.nn)), dim3(256), 0, s, m, (const double *)m.ye,
                       mask, out, nullptr, (int)inv);
  }
  ICX_CHECK_LAUNCH();
}

void launch_mesh_sensitivity(const MeshOp &m, double y1, double y2, double p, const double *design,
                             const double *u, const double *vol, double vscale, double *dfds, double *dgds,
                             hipStream_t s) {
  if (m.ne == 8)
    hipLaunchKernelGGL((k_mesh_energy<8, 3, true>), dim3(mesh_blocks(m.nelx)), dim3(256), 0, s, m, y1, y2, p,
                       design, nullptr, u, vol, vscale, dfds, dgds);
  else
    hipLaunchKernelGGL((k_mesh_energy<4, 2, true>), dim3(mesh_blocks(m.nelx)), dim3(256), 0, s, m, y1, y2, p,
                       design, nullptr, u, vol, vscale, dfds, dgds);
  ICX_CHECK_LAUNCH();
}

void launch_mesh_energy(const MeshOp &m, const double *E, const double *u, double *energy, hipStream_t s) {
  if (m.ne == 8)
    hipLaunchKernelGGL((k_mesh_energy<8, 3, false>), dim3(mesh_blocks(m.nelx)), dim3(256), 0, s, m, 0.0, 0.0,
                       0.0, nullptr, E, u, nullptr, 1.0, energy, nullptr);
  else
    hipLaunchKernelGGL((k_mesh_energy<4, 2, false>), dim3(mesh_blocks(m.nelx)), dim3(256), 0, s, m, 0.0, 0.0,
                       0.0, nullptr, E, u, nullptr, 1.0, energy, nullptr);
  ICX_CHECK_LAUNCH();
}

void launch_mesh_mises(const MeshOp &m, const double *E, const double *u, double *out, hipStream_t s) {
  if (m.ne == 8)
    hipLaunchKernelGGL((k_mesh_mises<8, 3>), dim3(mesh_blocks(m.nelx)), dim3(256), 0, s, m, E, u, out);
  else
    hipLaunchKernelGGL((k_mesh_mises<4, 2>), dim3(mesh_blocks(m.nelx)), dim3(256), 0, s, m, E, u, out);
  ICX_CHECK_LAUNCH();
}

int mesh_gather_blocks(const MeshOp &m) { return mesh_blocks(m.nn); }

template <typename T>
void launch_mesh_pcg_iteration(const MeshOp &m, SolverState *st, const T *E, const uint8_t *mask, const T *r,
                               const T *invd, T *p, T *Ap, T *x, const PcgPartials &P, hipStream_t s,
                               const LaunchEvents *ev) {
  const long long n = (long long)m.nc * m.nn;
  hipLaunchKernelGGL(k_mesh_dir<T>, dim3(ew_blocks(n, 4, 4096)), dim3(256), 0, s, n, st, r, invd, p, x, P);
  T *ye = reinterpret_cast<T *>(m.ye);
  if (m.ne == 8) {
    launch_ev(k_mesh_element<8, 3, T>, mesh_blocks(m.nelx), 256, s, ev, m, E, (const T *)p, ye,
              (const SolverState *)st);
    hipLaunchKernelGGL((k_mesh_gather_dot<3, T>), dim3(mesh_gather_blocks(m)), dim3(256), 0, s, m, (const T *)ye,
                       mask, (const T *)p, Ap, (const SolverState *)st, P.PA);
  } else {
    launch_ev(k_mesh_element<4, 2, T>, mesh_blocks(m.nelx), 256, s, ev, m, E, (const T *)p, ye,
              (const SolverState *)st);
    hipLaunchKernelGGL((k_mesh_gather_dot<2, T>), dim3(mesh_gather_blocks(m)), dim3(256), 0, s, m, (const T *)ye,
                       mask, (const T *)p, Ap, (const SolverState *)st, P.PA);
  }
  ICX_CHECK_LAUNCH();
}
template void launch_mesh_pcg_iteration<float>(const MeshOp &, SolverState *, const float *, const uint8_t *,
                                               const float *, const float *, float *, float *, float *,
                                               const PcgPartials &, hipStream_t, const LaunchEvents *);
template void launch_mesh_pcg_iteration<double>(const MeshOp &, SolverState *, const double *, const uint8_t *,
                                                const double *, const double *, double *, double *, double *,
                                                const PcgPartials &, hipStream_t, const LaunchEvents *);

void launch_mesh_helm_matvec(const MeshOp &m, double rh2, const double *x, double *y, const SolverState *st,
                             hipStream_t s) {
  if (m.ne == 8) mesh_scalar_pass<8, 3>(m, k_mesh_helm_element<8, 3>, y, false, st, s, nullptr, rh2, x, m.ye, st);
  else mesh_scalar_pass<4, 2>(m, k_mesh_helm_element<4, 2>, y, false, st, s, nullptr, rh2, x, m.ye, st);
  ICX_CHECK_LAUNCH();
}

void launch_mesh_helm_invdiag(const MeshOp &m, double rh2, double *invd, hipStream_t s) {
  if (m.ne == 8) mesh_scalar_pass<8, 3>(m, k_mesh_helm_diag<8, 3>, invd, true, nullptr, s, nullptr, rh2, m.ye);
  else mesh_scalar_pass<4, 2>(m, k_mesh_helm_diag<4, 2>, invd, true, nullptr, s, nullptr, rh2, m.ye);
  ICX_CHECK_LAUNCH();
}

void launch_mesh_helm_rhs(const MeshOp &m, double div, const double *rho, double *f, hipStream_t s) {
  if (m.ne == 8) mesh_scalar_pass<8, 3>(m, k_mesh_helm_rhs<8, 3>, f, false, nullptr, s, nullptr, div, rho, m.ye);
  else mesh_scalar_pass<4, 2>(m, k_mesh_helm_rhs<4, 2>, f, false, nullptr, s, nullptr, div, rho, m.ye);
  ICX_CHECK_LAUNCH();
}

void launch_mesh_helm_to_element(const MeshOp &m, double div, const double *phi, double *out, hipStream_t s) {
  if (m.ne == 8)
    hipLaunchKernelGGL((k_mesh_helm_to_element<8, 3>), dim3(mesh_blocks(m.nelx)), dim3(256), 0, s, m, div, phi, out);
  else
    hipLaunchKernelGGL((k_mesh_helm_to_element<4, 2>), dim3(mesh_blocks(m.nelx)), dim3(256), 0, s, m, div, phi, out);
  ICX_CHECK_LAUNCH();
}

}  // namespace icx
