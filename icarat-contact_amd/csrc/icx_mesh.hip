// Unstructured-mesh elastic operator of libicx (gfx950): explicit connectivity and one
// unit-modulus element matrix per element (MeshMDPA meshes, src/io/mesh_mdpa.hpp; the
// reference stores Ke per element as well, src/problem/elasticity.cpp:210-254).
//
// y = K(E) x in two passes, deterministic (no atomics):
//   element pass: y_e = E_e Ke0_e x_e, one thread per element; Ke0 is symmetric and its
//     upper triangle is stored entry-major ([p][e], p the row-major packed index of
//     i <= j), so the threads of a wave stream consecutive addresses; y_e is written
//     element-major ([e][i]) for the node pass;
//   node pass: y_n = sum over the (element, local node) pairs of node n, in increasing
//     element order (node adjacency in CSR), zero on constrained components.
// The Jacobi diagonal and the element energies u_e^T Ke0_e u_e use the same layouts.
#include "icx_common.cuh"

namespace icx {

namespace {

// packed index of (i, j), i <= j, in the row-major upper triangle of an nd x nd matrix
__host__ __device__ constexpr int tri(int nd, int i, int j) { return i * nd - i * (i - 1) / 2 + (j - i); }

// y = Ke0_e x over the packed upper triangle (each off-diagonal entry used twice)
template <int ND>
__device__ __forceinline__ void sym_apply(const double *ke, long long ne, long long e, const double (&x)[ND],
                                          double (&y)[ND]) {
#pragma unroll
  for (int i = 0; i < ND; ++i) y[i] = 0.0;
  // running pointer over the packed entries (one stride register, not one offset per entry)
  const double *kp = ke + e;
#pragma unroll
  for (int i = 0; i < ND; ++i)
#pragma unroll
    for (int j = i; j < ND; ++j) {
      const double k = *kp;
      kp += ne;
      y[i] = fma_t(k, x[j], y[i]);
      if (j != i) y[j] = fma_t(k, x[i], y[j]);
    }
}

template <int NE, int NC>
__global__ void __launch_bounds__(256) k_mesh_element(MeshOp m, const double *E, const double *x, double *ye,
                                                      const SolverState *st) {
  constexpr int ND = NE * NC;
  if (st && st->done) return;
  const long long ne = m.nelx;
  for (long long e = blockIdx.x * 256LL + threadIdx.x; e < ne; e += (long long)gridDim.x * 256) {
    double xe[ND];
#pragma unroll
    for (int a = 0; a < NE; ++a) {
      const long long n = m.conn[a * ne + e];
#pragma unroll
      for (int c = 0; c < NC; ++c) xe[a * NC + c] = x[c * m.nn + n];
    }
    double y[ND];
    sym_apply<ND>(m.ke, ne, e, xe, y);
    const double Ee = E[e];
    double *out = ye + e * ND;
#pragma unroll
    for (int i = 0; i < ND; ++i) out[i] = Ee * y[i];
  }
}

template <int NE, int NC>
__global__ void __launch_bounds__(256) k_mesh_gather(MeshOp m, const double *ye, const uint8_t *mask, double *y,
                                                     const SolverState *st) {
  constexpr int ND = NE * NC;
  if (st && st->done) return;
  for (long long n = blockIdx.x * 256LL + threadIdx.x; n < m.nn; n += (long long)gridDim.x * 256) {
    double acc[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] = 0.0;
    for (int k = m.adj_ptr[n]; k < m.adj_ptr[n + 1]; ++k) {
      const long long v = m.adj[k];
      const long long e = v / NE;
      const int a = (int)(v - e * NE);
#pragma unroll
      for (int c = 0; c < NC; ++c) acc[c] += ye[e * ND + a * NC + c];
    }
    const uint32_t msk = mask ? mask[n] : 0u;
#pragma unroll
    for (int c = 0; c < NC; ++c) y[c * m.nn + n] = ((msk >> c) & 1) ? 0.0 : acc[c];
  }
}

// diag (inv = false) or Jacobi inverse (inv = true: 1/d, 1 where d = 0 or constrained)
template <int NE, int NC>
__global__ void __launch_bounds__(256) k_mesh_diag(MeshOp m, const double *E, const uint8_t *mask, double *out,
                                                   bool inv) {
  constexpr int ND = NE * NC;
  const long long ne = m.nelx;
  for (long long n = blockIdx.x * 256LL + threadIdx.x; n < m.nn; n += (long long)gridDim.x * 256) {
    double acc[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] = 0.0;
    for (int k = m.adj_ptr[n]; k < m.adj_ptr[n + 1]; ++k) {
      const long long v = m.adj[k];
      const long long e = v / NE;
      const int a = (int)(v - e * NE);
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const int i = a * NC + c;
        acc[c] += E[e] * m.ke[(long long)tri(ND, i, i) * ne + e];
      }
    }
    const uint32_t msk = mask ? mask[n] : 0u;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const double d = ((msk >> c) & 1) ? 0.0 : acc[c];
      out[c * m.nn + n] = inv ? (d != 0.0 ? 1.0 / d : 1.0) : d;
    }
  }
}

// u_e^T Ke0_e u_e with |u| < 1e-15 flushed (the displacement the drivers' sensitivity
// reads), then the compliance sensitivity or E_e times the energy
template <int NE, int NC, bool SENS>
__global__ void __launch_bounds__(256) k_mesh_energy(MeshOp m, double y1, double y2, double p, const double *design,
                                                     const double *E, const double *u, const double *vol,
                                                     double vscale, double *dfds, double *dgds) {
  constexpr int ND = NE * NC;
  const long long ne = m.nelx;
  for (long long e = blockIdx.x * 256LL + threadIdx.x; e < ne; e += (long long)gridDim.x * 256) {
    // the SIMP derivative first: pow's registers are free again before the element loads
    double scale;
    if (SENS) {
      const double dens = design[e];
      if (y2 >= y1) scale = -(p * pow(dens, p - 1.0) * (y2 - y1));
      else scale = -((y2 - y1) * p * pow(1.0 - dens, p - 1.0));
      if (dgds) dgds[e] = vol[e] / vscale;
    } else {
      scale = E[e];
    }
    double ue[ND];
#pragma unroll
    for (int a = 0; a < NE; ++a) {
      const long long n = m.conn[a * ne + e];
#pragma unroll
      for (int c = 0; c < NC; ++c) ue[a * NC + c] = flush15(u[c * m.nn + n]);
    }
    double y[ND];
    sym_apply<ND>(m.ke, ne, e, ue, y);
    double en = 0.0;
#pragma unroll
    for (int i = 0; i < ND; ++i) en = fma_t(ue[i], y[i], en);
    if (dfds) dfds[e] = scale * en;
  }
}

int mesh_blocks(long long n) { return ew_blocks(n, 1, 4096); }

}  // namespace

void launch_mesh_matvec(const MeshOp &m, const double *E, const uint8_t *mask, const double *x, double *y,
                        const SolverState *st, hipStream_t s, const LaunchEvents *ev) {
  if (m.ne == 8) {
    launch_ev(k_mesh_element<8, 3>, mesh_blocks(m.nelx), 256, s, ev, m, E, x, m.ye, st);
    hipLaunchKernelGGL((k_mesh_gather<8, 3>), dim3(mesh_blocks(m.nn)), dim3(256), 0, s, m, (const double *)m.ye,
                       mask, y, st);
  } else {
    launch_ev(k_mesh_element<4, 2>, mesh_blocks(m.nelx), 256, s, ev, m, E, x, m.ye, st);
    hipLaunchKernelGGL((k_mesh_gather<4, 2>), dim3(mesh_blocks(m.nn)), dim3(256), 0, s, m, (const double *)m.ye,
                       mask, y, st);
  }
  ICX_CHECK_LAUNCH();
}

void launch_mesh_diag(const MeshOp &m, const double *E, const uint8_t *mask, double *out, bool inv,
                      hipStream_t s) {
  if (m.ne == 8)
    hipLaunchKernelGGL((k_mesh_diag<8, 3>), dim3(mesh_blocks(m.nn)), dim3(256), 0, s, m, E, mask, out, inv);
  else
    hipLaunchKernelGGL((k_mesh_diag<4, 2>), dim3(mesh_blocks(m.nn)), dim3(256), 0, s, m, E, mask, out, inv);
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

}  // namespace icx
