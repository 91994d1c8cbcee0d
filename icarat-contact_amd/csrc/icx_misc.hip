// Element-level and streaming kernels of libicx (gfx950): SIMP moduli, Jacobi
// diagonals, layout conversions, sensitivity, cone filter, projection, Helmholtz
// right-hand side, OC and MMA pieces.
#include "icx_device.h"

#include <math.h>

namespace icx {

// ---------------------------------------------------------------------------------
// element moduli, diagonals, layout conversions
// ---------------------------------------------------------------------------------
// SIMP, src/problem/elasticity.hpp:15-22
__device__ __forceinline__ double simp(double y1, double y2, double s, double p) {
  if (y2 - y1 >= 0.0) return (1.0 - pow(s, p)) * y1 + pow(s, p) * y2;
  return pow(1.0 - s, p) * y1 + (1.0 - pow(1.0 - s, p)) * y2;
}

__global__ void k_simp(long long ne, double y1, double y2, double p, const double *design, double *E64,
                       float *E32) {
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < ne;
       e += (long long)gridDim.x * blockDim.x) {
    const double E = simp(y1, y2, design[e], p);
    E64[e] = E;
    if (E32) E32[e] = (float)E;
  }
}
void launch_simp(long long ne, double y1, double y2, double p, const double *design, double *E64,
                 float *E32, hipStream_t s) {
  hipLaunchKernelGGL(k_simp, dim3(ew_blocks(ne)), dim3(256), 0, s, ne, y1, y2, p, design, E64, E32);
  ICX_CHECK_LAUNCH();
}

struct KD { double v[3]; };
__global__ void k_elastic_invdiag(Grid g, KD kd, const double *E, const uint8_t *mask, double *invd64,
                                  float *invd32) {
  const int nc = g.ndim;
  for (long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x; n < g.nn;
       n += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(n % g.sx), j = (int)((n / g.sx) % g.sy), k = (int)(n / ((long long)g.sx * g.sy));
    double S = 0.0;
    for (int dz = (g.ndim == 3 ? -1 : 0); dz <= 0; ++dz)
      for (int dy = -1; dy <= 0; ++dy)
        for (int dx = -1; dx <= 0; ++dx) {
          const int ex = i + dx, ey = j + dy, ez = k + dz;
          if (ex < 0 || ey < 0 || ez < 0 || ex >= g.nx || ey >= g.ny || ez >= g.nz) continue;
          S += E[ex + (long long)g.nx * (ey + (long long)g.ny * ez)];
        }
    const uint8_t m = mask ? mask[n] : 0;
    for (int c = 0; c < nc; ++c) {
      const double d = kd.v[c] * S;
      double inv = (d != 0.0) ? 1.0 / d : 1.0;
      if ((m >> c) & 1) inv = 0.0;
      if (invd64) invd64[c * g.nn + n] = inv;
      if (invd32) invd32[c * g.nn + n] = (float)inv;
    }
  }
}
void launch_elastic_invdiag(const Grid &g, const double *kd, const double *E, const uint8_t *mask,
                            double *invd64, float *invd32, hipStream_t s) {
  KD k{{kd[0], kd[1], kd[2]}};
  hipLaunchKernelGGL(k_elastic_invdiag, dim3(ew_blocks(g.nn)), dim3(256), 0, s, g, k, E, mask, invd64,
                     invd32);
  ICX_CHECK_LAUNCH();
}

template <typename T>
__global__ void k_elastic_invdiag_node(Grid g, T ik0, T ik1, T ik2, const double *E, const uint8_t *mask, T *invs,
                                       T *invd, bool aos) {
  const int nc = g.ndim;
  const T ik[3] = {ik0, ik1, ik2};
  for (long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x; n < g.nn;
       n += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(n % g.sx), j = (int)((n / g.sx) % g.sy), k = (int)(n / ((long long)g.sx * g.sy));
    double S = 0.0;
    for (int dz = (g.ndim == 3 ? -1 : 0); dz <= 0; ++dz)
      for (int dy = -1; dy <= 0; ++dy)
        for (int dx = -1; dx <= 0; ++dx) {
          const int ex = i + dx, ey = j + dy, ez = k + dz;
          if (ex < 0 || ey < 0 || ez < 0 || ex >= g.nx || ey >= g.ny || ez >= g.nz) continue;
          S += E[ex + (long long)g.nx * (ey + (long long)g.ny * ez)];
        }
    const T s = S != 0.0 ? (T)(1.0 / S) : T(-1);
    invs[n] = s;
    const uint8_t m = mask ? mask[n] : 0;
    for (int c = 0; c < nc; ++c)
      invd[aos ? nc * n + c : c * g.nn + n] = ((m >> c) & 1) ? T(0) : jacobi_inv(s, ik[c]);
  }
}
template <typename T>
void launch_elastic_invdiag_node(const Grid &g, const ElasticOp<T> &op, const double *E, const uint8_t *mask,
                                 T *invs, T *invd, hipStream_t s, bool aos) {
  hipLaunchKernelGGL(k_elastic_invdiag_node<T>, dim3(ew_blocks(g.nn)), dim3(256), 0, s, g, op.ikd[0], op.ikd[1],
                     op.ikd[2], E, mask, invs, invd, aos);
  ICX_CHECK_LAUNCH();
}
template void launch_elastic_invdiag_node<float>(const Grid &, const ElasticOp<float> &, const double *,
                                                 const uint8_t *, float *, float *, hipStream_t, bool);
template void launch_elastic_invdiag_node<double>(const Grid &, const ElasticOp<double> &, const double *,
                                                  const uint8_t *, double *, double *, hipStream_t, bool);

__global__ void k_helm_invdiag(Grid g, double kdiag, double *invd) {
  for (long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x; n < g.nn;
       n += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(n % g.sx), j = (int)((n / g.sx) % g.sy), k = (int)(n / ((long long)g.sx * g.sy));
    int cnt = 0;
    for (int dz = (g.ndim == 3 ? -1 : 0); dz <= 0; ++dz)
      for (int dy = -1; dy <= 0; ++dy)
        for (int dx = -1; dx <= 0; ++dx) {
          const int ex = i + dx, ey = j + dy, ez = k + dz;
          cnt += !(ex < 0 || ey < 0 || ez < 0 || ex >= g.nx || ey >= g.ny || ez >= g.nz);
        }
    const double d = kdiag * cnt;
    invd[n] = d != 0.0 ? 1.0 / d : 1.0;
  }
}
void launch_helm_invdiag(const Grid &g, double kdiag, double *invd, hipStream_t s) {
  hipLaunchKernelGGL(k_helm_invdiag, dim3(ew_blocks(g.nn)), dim3(256), 0, s, g, kdiag, invd);
  ICX_CHECK_LAUNCH();
}

__global__ void k_dof_to_grid(Grid g, int nc, const int *dof, int numeq, const double *v,
                              const double *dirich, double *grid) {
  for (long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x; n < g.nn;
       n += (long long)gridDim.x * blockDim.x)
    for (int c = 0; c < nc; ++c) {
      const int d = dof[n * nc + c];
      grid[c * g.nn + n] = d < numeq ? v[d] : (dirich ? dirich[6 * n + c] : 0.0);
    }
}
void launch_dof_to_grid(const Grid &g, int nc, const int *dof, int numeq, const double *v,
                        const double *dirich, double *grid, hipStream_t s) {
  hipLaunchKernelGGL(k_dof_to_grid, dim3(ew_blocks(g.nn)), dim3(256), 0, s, g, nc, dof, numeq, v, dirich,
                     grid);
  ICX_CHECK_LAUNCH();
}

__global__ void k_grid_to_dof(Grid g, int nc, const int *dof, int numeq, const double *grid, double *v) {
  for (long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x; n < g.nn;
       n += (long long)gridDim.x * blockDim.x)
    for (int c = 0; c < nc; ++c) {
      const int d = dof[n * nc + c];
      if (d < numeq) v[d] = grid[c * g.nn + n];
    }
}
void launch_grid_to_dof(const Grid &g, int nc, const int *dof, int numeq, const double *grid,
                        double *v, hipStream_t s) {
  hipLaunchKernelGGL(k_grid_to_dof, dim3(ew_blocks(g.nn)), dim3(256), 0, s, g, nc, dof, numeq, grid, v);
  ICX_CHECK_LAUNCH();
}

// divideValIntoNode, src/analysis/analysis_utility.hpp:8-31 (|v| < 1e-15 -> 0)
__global__ void k_node_values(Grid g, int nc, const int *dof, int numeq, const double *grid,
                              const double *dirich, double *nodeval) {
  for (long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x; n < g.nn;
       n += (long long)gridDim.x * blockDim.x)
    for (int c = 0; c < 3; ++c) {
      double v = 0.0;
      if (c < nc) {
        const int d = dof[n * nc + c];
        v = d < numeq ? grid[c * g.nn + n] : (dirich ? dirich[6 * n + c] : 0.0);
        if (fabs(v) < 1.0e-15) v = 0.0;
      }
      nodeval[3 * n + c] = v;
    }
}
void launch_node_values(const Grid &g, int nc, const int *dof, int numeq, const double *grid,
                        const double *dirich, double *nodeval, hipStream_t s) {
  hipLaunchKernelGGL(k_node_values, dim3(ew_blocks(g.nn)), dim3(256), 0, s, g, nc, dof, numeq, grid,
                     dirich, nodeval);
  ICX_CHECK_LAUNCH();
}

__global__ void k_f64_f32(long long n, const double *a, float *b) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}
void launch_convert_f64_f32(long long n, const double *a, float *b, hipStream_t s) {
  hipLaunchKernelGGL(k_f64_f32, dim3(ew_blocks(n, 4)), dim3(256), 0, s, n, a, b);
  ICX_CHECK_LAUNCH();
}

// component planes <-> node-interleaved FP64 vectors (nn nodes of nc components)
__global__ void k_soa_aos_f64(long long nn, int nc, const double *a, double *b, bool to_aos) {
  for (long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x; n < nn;
       n += (long long)gridDim.x * blockDim.x)
    for (int c = 0; c < nc; ++c) {
      if (to_aos) b[nc * n + c] = a[c * nn + n];
      else b[c * nn + n] = a[nc * n + c];
    }
}
void launch_soa_to_aos_f64(long long nn, int nc, const double *a, double *b, hipStream_t s) {
  hipLaunchKernelGGL(k_soa_aos_f64, dim3(ew_blocks(nn, 4)), dim3(256), 0, s, nn, nc, a, b, true);
  ICX_CHECK_LAUNCH();
}
void launch_aos_to_soa_f64(long long nn, int nc, const double *a, double *b, hipStream_t s) {
  hipLaunchKernelGGL(k_soa_aos_f64, dim3(ew_blocks(nn, 4)), dim3(256), 0, s, nn, nc, a, b, false);
  ICX_CHECK_LAUNCH();
}

__global__ void k_f64_f32_aos(long long nn, int nc, const double *a, float *b) {
  for (long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x; n < nn;
       n += (long long)gridDim.x * blockDim.x)
    for (int c = 0; c < nc; ++c) b[nc * n + c] = (float)a[c * nn + n];
}
void launch_convert_f64_f32_aos(long long nn, int nc, const double *a, float *b, hipStream_t s) {
  hipLaunchKernelGGL(k_f64_f32_aos, dim3(ew_blocks(nn, 4)), dim3(256), 0, s, nn, nc, a, b);
  ICX_CHECK_LAUNCH();
}

__global__ void k_f32_f64(long long n, const float *a, double *b) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    b[i] = (double)a[i];
}
void launch_convert_f32_f64(long long n, const float *a, double *b, hipStream_t s) {
  hipLaunchKernelGGL(k_f32_f64, dim3(ew_blocks(n, 4)), dim3(256), 0, s, n, a, b);
  ICX_CHECK_LAUNCH();
}

__global__ void k_axpby(long long n, double a, const double *x, double b, double *y) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    y[i] = a * x[i] + b * y[i];
}
void launch_axpby(long long n, double a, const double *x, double b, double *y, hipStream_t s) {
  hipLaunchKernelGGL(k_axpby, dim3(ew_blocks(n, 4)), dim3(256), 0, s, n, a, x, b, y);
  ICX_CHECK_LAUNCH();
}

__global__ void k_mul(long long n, const double *a, double *b) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    b[i] *= a[i];
}
void launch_mul(long long n, const double *a, double *b, hipStream_t s) {
  hipLaunchKernelGGL(k_mul, dim3(ew_blocks(n)), dim3(256), 0, s, n, a, b);
  ICX_CHECK_LAUNCH();
}

__global__ void k_mask_zero(Grid g, int nc, const uint8_t *mask, double *v) {
  for (long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x; n < g.nn;
       n += (long long)gridDim.x * blockDim.x) {
    const uint8_t m = mask[n];
    for (int c = 0; c < nc; ++c)
      if ((m >> c) & 1) v[c * g.nn + n] = 0.0;
  }
}
void launch_mask_zero(const Grid &g, int nc, const uint8_t *mask, double *v, hipStream_t s) {
  hipLaunchKernelGGL(k_mask_zero, dim3(ew_blocks(g.nn)), dim3(256), 0, s, g, nc, mask, v);
  ICX_CHECK_LAUNCH();
}

// ---------------------------------------------------------------------------------
// sensitivity (opt_macro_oc.cpp:182-212, opt_mma_Helmholtz.cpp:173-205)
// ---------------------------------------------------------------------------------

template <bool SENS>
__global__ void k_element_energy(Grid g, ElasticOp<double> op, double y1, double y2, double p,
                                 const double *design, const double *E, const double *u, double vol,
                                 double vscale, double *dfds, double *dgds) {
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < g.ne;
       e += (long long)gridDim.x * blockDim.x) {
    const int ex = (int)(e % g.nx), ey = (int)((e / g.nx) % g.ny), ez = (int)(e / ((long long)g.nx * g.ny));
    double en;
    if (g.ndim == 3) {
      double v[3][8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const long long m = (ex + (q & 1)) + (long long)g.sx * ((ey + ((q >> 1) & 1)) + (long long)g.sy * (ez + (q >> 2)));
#pragma unroll
        for (int c = 0; c < 3; ++c) v[c][q] = flush15(u[c * g.nn + m]);
      }
      en = energy_hex8<double>(v, op);
    } else {
      double v[2][4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const long long m = (ex + (q & 1)) + (long long)g.sx * (ey + (q >> 1));
#pragma unroll
        for (int c = 0; c < 2; ++c) v[c][q] = flush15(u[c * g.nn + m]);
      }
      en = energy_quad4<double>(v, op);
    }
    if (SENS) {
      const double dens = design[e];
      double dyoung;
      if (y2 >= y1) dyoung = p * pow(dens, p - 1.0) * (y2 - y1);
      else dyoung = (y2 - y1) * p * pow(1.0 - dens, p - 1.0);
      if (dfds) dfds[e] = -dyoung * en;
      if (dgds) dgds[e] = vol / vscale;
    } else {
      dfds[e] = E[e] * en;
    }
  }
}
void launch_sensitivity(const Grid &g, const ElasticOp<double> &op, double y1, double y2, double p,
                        const double *design, const double *u, double vol, double vscale,
                        double *dfds, double *dgds, hipStream_t s) {
  hipLaunchKernelGGL(k_element_energy<true>, dim3(ew_blocks(g.ne)), dim3(256), 0, s, g, op, y1, y2, p,
                     design, nullptr, u, vol, vscale, dfds, dgds);
  ICX_CHECK_LAUNCH();
}
void launch_energy(const Grid &g, const ElasticOp<double> &op, const double *E, const double *u,
                   double *energy, hipStream_t s) {
  hipLaunchKernelGGL(k_element_energy<false>, dim3(ew_blocks(g.ne)), dim3(256), 0, s, g, op, 0.0, 0.0,
                     0.0, nullptr, E, u, 0.0, 1.0, energy, nullptr);
  ICX_CHECK_LAUNCH();
}

// ---------------------------------------------------------------------------------
// output step: node.val, element-mean von Mises stress, reactions
// ---------------------------------------------------------------------------------
__global__ void k_full_disp(long long n, const double *u, const double *dirich, double *out) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    out[i] = flush15(u[i] + (dirich ? dirich[i] : 0.0));
}
void launch_full_displacement(long long n, const double *u, const double *dirich, double *out,
                              hipStream_t s) {
  hipLaunchKernelGGL(k_full_disp, dim3(ew_blocks(n, 4)), dim3(256), 0, s, n, u, dirich, out);
  ICX_CHECK_LAUNCH();
}

// Elastic::makeMisesStress (src/problem/elasticity.cpp:165-208): the strain and stress
// averaged over the Gauss points, mises = sqrt(s^T V s).  On an axis-aligned box the
// Gauss-point average of B is B at the element centre (the rule is symmetric and B is
// linear in each coordinate it depends on), where dN_q/dx = xi_q / (n_e/2 h_x): the
// strains are face differences, e.g. e_xx = (sum of u_x on the +x face - on the -x
// face) / (n_e/2 h_x).  De as in src/problem/elasticity.cpp:34-98 with E_e.
__global__ void k_mises(Grid g, MisesMat m, const double *E, const double *u, double *out) {
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < g.ne;
       e += (long long)gridDim.x * blockDim.x) {
    const int ex = (int)(e % g.nx), ey = (int)((e / g.nx) % g.ny), ez = (int)(e / ((long long)g.nx * g.ny));
    const double Ee = E[e], nu = m.nu;
    double q;
    if (g.ndim == 3) {
      double dx[3] = {0, 0, 0}, dy[3] = {0, 0, 0}, dz[3] = {0, 0, 0};
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int a = k & 1, b = (k >> 1) & 1, c3 = k >> 2;
        const long long n = (ex + a) + (long long)g.sx * ((ey + b) + (long long)g.sy * (ez + c3));
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const double v = u[c * g.nn + n];
          dx[c] += a ? v : -v;
          dy[c] += b ? v : -v;
          dz[c] += c3 ? v : -v;
        }
      }
      const double fx = 0.25 / m.h[0], fy = 0.25 / m.h[1], fz = 0.25 / m.h[2];
      const double exx = dx[0] * fx, eyy = dy[1] * fy, ezz = dz[2] * fz;
      const double gxy = dy[0] * fy + dx[1] * fx, gyz = dz[1] * fz + dy[2] * fy, gzx = dz[0] * fz + dx[2] * fx;
      const double lambda = Ee * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)), mu = Ee / (2 * (1.0 + nu));
      const double sxx = (lambda + 2 * mu) * exx + lambda * (eyy + ezz);
      const double syy = (lambda + 2 * mu) * eyy + lambda * (exx + ezz);
      const double szz = (lambda + 2 * mu) * ezz + lambda * (exx + eyy);
      const double sxy = mu * gxy, syz = mu * gyz, szx = mu * gzx;
      q = sxx * sxx + syy * syy + szz * szz - sxx * syy - syy * szz - sxx * szz +
          3.0 * (sxy * sxy + syz * syz + szx * szx);
    } else {
      double dx[2] = {0, 0}, dy[2] = {0, 0};
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int a = k & 1, b = k >> 1;
        const long long n = (ex + a) + (long long)g.sx * (ey + b);
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          const double v = u[c * g.nn + n];
          dx[c] += a ? v : -v;
          dy[c] += b ? v : -v;
        }
      }
      const double fx = 0.5 / m.h[0], fy = 0.5 / m.h[1];
      const double exx = dx[0] * fx, eyy = dy[1] * fy, gxy = dy[0] * fy + dx[1] * fx;
      double sxx, syy, sxy;
      if (m.state == 0) {  // plane stress
        const double f = Ee / (1.0 - nu * nu);
        sxx = f * (exx + nu * eyy);
        syy = f * (nu * exx + eyy);
        sxy = f * (0.5 * (1.0 - nu)) * gxy;
      } else {             // plane strain
        const double f = Ee * ((1.0 - nu) / ((1.0 - 2.0 * nu) * (1.0 + nu))), a = nu / (1.0 - nu);
        sxx = f * (exx + a * eyy);
        syy = f * (a * exx + eyy);
        sxy = f * ((1.0 - 2.0 * nu) / (2.0 * (1.0 - nu))) * gxy;
      }
      q = sxx * sxx - sxx * syy + syy * syy + 3.0 * sxy * sxy;
    }
    out[e] = sqrt(q);
  }
}
void launch_mises(const Grid &g, const MisesMat &m, const double *E, const double *u, double *out,
                  hipStream_t s) {
  hipLaunchKernelGGL(k_mises, dim3(ew_blocks(g.ne)), dim3(256), 0, s, g, m, E, u, out);
  ICX_CHECK_LAUNCH();
}

// ---------------------------------------------------------------------------------
// stress constraint, quad4 box jobs (constraint/opt-stress-constraint/sensitivity.cpp:8-155).
// Element-mean strain of a rectangle = strain at its centre; dN_k/dx = +-1/(2 hx),
// dN_k/dy = +-1/(2 hy) for local node k = a + 2b at (ex + a, ey + b).
// ---------------------------------------------------------------------------------
__device__ __forceinline__ void q4_mean_stress(const Grid &g, const MisesMat &m, double Ee, const double *u,
                                               int ex, int ey, double (&sg)[3]) {
  double dx[2] = {0, 0}, dy[2] = {0, 0};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int a = k & 1, b = k >> 1;
    const long long n = (ex + a) + (long long)g.sx * (ey + b);
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const double v = flush15(u[c * g.nn + n]);
      dx[c] += a ? v : -v;
      dy[c] += b ? v : -v;
    }
  }
  const double fx = 0.5 / m.h[0], fy = 0.5 / m.h[1], nu = m.nu;
  const double exx = dx[0] * fx, eyy = dy[1] * fy, gxy = dy[0] * fy + dx[1] * fx;
  if (m.state == 0) {
    const double f = Ee / (1.0 - nu * nu);
    sg[0] = f * (exx + nu * eyy); sg[1] = f * (nu * exx + eyy); sg[2] = f * (0.5 * (1.0 - nu)) * gxy;
  } else {
    const double f = Ee * ((1.0 - nu) / ((1.0 - 2.0 * nu) * (1.0 + nu))), a = nu / (1.0 - nu);
    sg[0] = f * (exx + a * eyy); sg[1] = f * (a * exx + eyy); sg[2] = f * ((1.0 - 2.0 * nu) / (2.0 * (1.0 - nu))) * gxy;
  }
}

// sigma_e, vm_e, s_e = rho^qp E2 vm_e / E_e; partials of sum (s_e / slim)^P
__global__ void __launch_bounds__(256) k_stress_eval(Grid g, MisesMat m, StressPar sp, const double *design,
                                                     const double *E, const double *u, double *sig, double *vm,
                                                     double *sqp, Reduce red) {
  __shared__ double scr[4];
  double acc = 0.0;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < g.ne;
       e += (long long)gridDim.x * blockDim.x) {
    const int ex = (int)(e % g.nx), ey = (int)(e / g.nx);
    const double Ee = E[e];
    double sg[3];
    q4_mean_stress(g, m, Ee, u, ex, ey, sg);
    const double v = sqrt(sg[0] * sg[0] - sg[0] * sg[1] + sg[1] * sg[1] + 3.0 * sg[2] * sg[2]);
    const double s = pow(design[e], sp.qp) * sp.y2 * v / Ee;
    sig[e] = sg[0]; sig[g.ne + e] = sg[1]; sig[2 * g.ne + e] = sg[2];
    vm[e] = v;
    sqp[e] = s;
    acc += pow(s / sp.slim, sp.pn);
  }
  acc = block_sum<256>(acc, scr);
  if (threadIdx.x == 0) red.partial[blockIdx.x] = acc;
}

// factor1 . factor2 into f12, and the element loads of the adjoint problem
// Fe = (rho^qp E2 / E_e) factor1^T De B_mean into fe[(2k + c) ne + e]
__global__ void __launch_bounds__(256) k_stress_dual(Grid g, MisesMat m, StressPar sp, double dpndvm,
                                                     const double *design, const double *E, const double *sig,
                                                     const double *vm, const double *sqp, double *f12, double *fe) {
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < g.ne;
       e += (long long)gridDim.x * blockDim.x) {
    const double Ee = E[e], rho = design[e], v = vm[e], nu = m.nu;
    const double s0 = sig[e], s1 = sig[g.ne + e], s2 = sig[2 * g.ne + e];
    const double dvm[3] = {(2.0 * s0 - s1) / (2.0 * v), (2.0 * s1 - s0) / (2.0 * v), 3.0 * s2 / v};
    const double c1 = dpndvm * pow(sqp[e] / sp.slim, sp.pn - 1);
    const double dvds = sp.qp * pow(rho, sp.qp - 1.0) * sp.y2;
    double f1[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) f1[i] = c1 * dvm[i] / sp.slim;
    f12[e] = f1[0] * (dvds * s0 / Ee) + f1[1] * (dvds * s1 / Ee) + f1[2] * (dvds * s2 / Ee);
    // t = De f1 (De = E_e D symmetric), scaled by rho^qp E2 / E_e
    double d00, d01, d22;
    if (m.state == 0) {
      const double f = Ee / (1.0 - nu * nu);
      d00 = f; d01 = f * nu; d22 = f * (0.5 * (1.0 - nu));
    } else {
      const double f = Ee * ((1.0 - nu) / ((1.0 - 2.0 * nu) * (1.0 + nu)));
      d00 = f; d01 = f * (nu / (1.0 - nu)); d22 = f * ((1.0 - 2.0 * nu) / (2.0 * (1.0 - nu)));
    }
    const double coef = pow(rho, sp.qp) * sp.y2 / Ee;
    const double t0 = coef * (d00 * f1[0] + d01 * f1[1]), t1 = coef * (d01 * f1[0] + d00 * f1[1]),
                 t2 = coef * d22 * f1[2];
    const double hx = 0.5 / m.h[0], hy = 0.5 / m.h[1];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const double nx = (k & 1) ? hx : -hx, ny = (k >> 1) ? hy : -hy;
      fe[(2 * k) * g.ne + e] = nx * t0 + ny * t2;
      fe[(2 * k + 1) * g.ne + e] = ny * t1 + nx * t2;
    }
  }
}

// node gather of the element loads (free DOFs; constrained ones are zero)
__global__ void k_q4_gather(Grid g, const double *fe, const uint8_t *mask, double *out) {
  for (long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x; n < g.nn;
       n += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(n % g.sx), j = (int)(n / g.sx);
    double acc[2] = {0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int a = k & 1, b = k >> 1, ex = i - a, ey = j - b;
      if (ex < 0 || ey < 0 || ex >= g.nx || ey >= g.ny) continue;
      const long long e = ex + (long long)g.nx * ey;
#pragma unroll
      for (int c = 0; c < 2; ++c) acc[c] += fe[(2 * k + c) * g.ne + e];
    }
    const uint8_t mk = mask[n];
#pragma unroll
    for (int c = 0; c < 2; ++c) out[c * g.nn + n] = ((mk >> c) & 1) ? 0.0 : acc[c];
  }
}

// dg_e = factor1 . factor2 - p rho^(p-1) (E2 - E1) u_e^T K0 lambda_e
__global__ void k_stress_sens(Grid g, ElasticOp<double> op, StressPar sp, const double *design, const double *u,
                              const double *lam, const double *f12, double *dgds) {
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < g.ne;
       e += (long long)gridDim.x * blockDim.x) {
    const int ex = (int)(e % g.nx), ey = (int)(e / g.nx);
    double vu[2][4], vl[2][4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const long long n = (ex + (q & 1)) + (long long)g.sx * (ey + (q >> 1));
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        vu[c][q] = flush15(u[c * g.nn + n]);
        vl[c][q] = lam[c * g.nn + n];
      }
    }
    const double b = bilinear_quad4<double>(vu, vl, op);
    const double dvpds = sp.p * pow(design[e], sp.p - 1.0) * (sp.y2 - sp.y1);
    dgds[e] = f12[e] - dvpds * b;
  }
}

void launch_stress_eval(const Grid &g, const MisesMat &m, const StressPar &sp, const double *design,
                        const double *E, const double *u, double *sig, double *vm, double *sqp, double *psum,
                        Reduce red, hipStream_t s) {
  const int nb = ew_blocks(g.ne, 4, RED_BLOCKS);
  hipLaunchKernelGGL(k_stress_eval, dim3(nb), dim3(256), 0, s, g, m, sp, design, E, u, sig, vm, sqp, red);
  launch_finalize(red, nb, 1, psum, s);
}
void launch_stress_adjoint_load(const Grid &g, const MisesMat &m, const StressPar &sp, double dpndvm,
                                const double *design, const double *E, const double *sig, const double *vm,
                                const double *sqp, double *f12, double *fe, const uint8_t *mask, double *load,
                                hipStream_t s) {
  hipLaunchKernelGGL(k_stress_dual, dim3(ew_blocks(g.ne)), dim3(256), 0, s, g, m, sp, dpndvm, design, E, sig, vm,
                     sqp, f12, fe);
  ICX_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_q4_gather, dim3(ew_blocks(g.nn)), dim3(256), 0, s, g, fe, mask, load);
  ICX_CHECK_LAUNCH();
}
void launch_stress_sens(const Grid &g, const ElasticOp<double> &op, const StressPar &sp, const double *design,
                        const double *u, const double *lam, const double *f12, double *dgds, hipStream_t s) {
  hipLaunchKernelGGL(k_stress_sens, dim3(ew_blocks(g.ne)), dim3(256), 0, s, g, op, sp, design, u, lam, f12, dgds);
  ICX_CHECK_LAUNCH();
}

__global__ void k_grid_to_nodes(long long nn, int nc, const double *grid, double *out) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nn;
       i += (long long)gridDim.x * blockDim.x)
#pragma unroll
    for (int c = 0; c < 3; ++c) out[3 * i + c] = c < nc ? flush15(grid[c * nn + i]) : 0.0;
}
void launch_grid_to_nodes(const Grid &g, int nc, const double *grid, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_grid_to_nodes, dim3(ew_blocks(g.nn)), dim3(256), 0, s, g.nn, nc, grid, out);
  ICX_CHECK_LAUNCH();
}

// ---------------------------------------------------------------------------------
// heat conduction (src/problem/heat.cpp:62-109): node-parallel gather over the 2^d
// elements of a node, dense element rows (the anisotropic conductivity tensor has no
// mirror symmetry to exploit), FP64
// ---------------------------------------------------------------------------------
__global__ void k_heat_weight(long long ne, double p, const double *design, double *w) {
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < ne;
       e += (long long)gridDim.x * blockDim.x)
    w[e] = pow(design[e], p);
}
void launch_heat_weight(long long ne, double p, const double *design, double *w, hipStream_t s) {
  hipLaunchKernelGGL(k_heat_weight, dim3(ew_blocks(ne)), dim3(256), 0, s, ne, p, design, w);
  ICX_CHECK_LAUNCH();
}

// corners of element (ex, ey, ez) in the order q = a + 2b + 4c
__device__ __forceinline__ long long corner(const Grid &g, int ex, int ey, int ez, int q) {
  return (ex + (q & 1)) + (long long)g.sx * ((ey + ((q >> 1) & 1)) + (long long)g.sy * (ez + (q >> 2)));
}

// y_n = sum over the elements e of node n (local index qn) of row qn of K_e times x_e;
// DIAG: the diagonal entry instead
template <bool DIAG>
__global__ void k_heat_op(Grid g, HeatOp op, const double *w, const uint8_t *mask, const double *x, double *y,
                          const SolverState *st) {
  if (st && st->done) return;
  const int nq = op.nq;
  for (long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x; n < g.nn;
       n += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(n % g.sx), j = (int)((n / g.sx) % g.sy), k = (int)(n / ((long long)g.sx * g.sy));
    double acc = 0.0;
    const int nz = g.ndim == 3 ? 2 : 1;
    for (int dz = 0; dz < nz; ++dz)
      for (int dy = 0; dy < 2; ++dy)
        for (int dx = 0; dx < 2; ++dx) {
          const int ex = i - 1 + dx, ey = j - 1 + dy, ez = g.ndim == 3 ? k - 1 + dz : 0;
          if (ex < 0 || ey < 0 || ez < 0 || ex >= g.nx || ey >= g.ny || ez >= g.nz) continue;
          const long long e = ex + (long long)g.nx * (ey + (long long)g.ny * ez);
          const int qn = (1 - dx) + 2 * (1 - dy) + (g.ndim == 3 ? 4 * (1 - dz) : 0);
          const double we = w[e];
          double a1 = 0.0, a2 = 0.0;
          if (DIAG) {
            a1 = op.K1[qn * nq + qn];
            a2 = op.K2[qn * nq + qn];
          } else {
            for (int q = 0; q < nq; ++q) {
              const double xq = x[corner(g, ex, ey, ez, q)];
              a1 += op.K1[qn * nq + q] * xq;
              a2 += op.K2[qn * nq + q] * xq;
            }
          }
          acc += we * a1 + (1.0 - we) * a2;
        }
    const bool fixed = mask && (mask[n] & 1);
    if (DIAG) y[n] = fixed ? 0.0 : (acc != 0.0 ? 1.0 / acc : 1.0);
    else y[n] = fixed ? 0.0 : acc;
  }
}
// The same product when K1 and K2 are Haar-diagonal: each element's 2^d values go
// through the Haar transform once, are scaled by w kf1 + (1 - w) kf2 and transformed
// back; the node keeps its own corner (as the Helmholtz operator).
template <int ND>
__global__ void __launch_bounds__(256) k_heat_haar(Grid g, HeatOp op, const double *w, const uint8_t *mask,
                                                   const double *x, double *y, const SolverState *st) {
  if (st && st->done) return;
  constexpr int NQ = ND == 3 ? 8 : 4;
  for (long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x; n < g.nn;
       n += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(n % g.sx), j = (int)((n / g.sx) % g.sy), k = ND == 3 ? (int)(n / ((long long)g.sx * g.sy)) : 0;
    double acc = 0.0;
#pragma unroll
    for (int dz = 0; dz < (ND == 3 ? 2 : 1); ++dz)
#pragma unroll
      for (int dy = 0; dy < 2; ++dy)
#pragma unroll
        for (int dx = 0; dx < 2; ++dx) {
          const int ex = i - 1 + dx, ey = j - 1 + dy, ez = ND == 3 ? k - 1 + dz : 0;
          if (ex < 0 || ey < 0 || ez < 0 || ex >= g.nx || ey >= g.ny || ez >= g.nz) continue;
          const long long e = ex + (long long)g.nx * (ey + (long long)g.ny * ez);
          const double we = w[e];
          double v[NQ];
#pragma unroll
          for (int q = 0; q < NQ; ++q) v[q] = x[corner(g, ex, ey, ez, q)];
          if (ND == 3) haar_fwd8(reinterpret_cast<double(&)[8]>(v));
          else haar_fwd4(reinterpret_cast<double(&)[4]>(v));
#pragma unroll
          for (int q = 0; q < NQ; ++q) v[q] *= we * op.kf1[q] + (1.0 - we) * op.kf2[q];
          if (ND == 3) haar_inv8(reinterpret_cast<double(&)[8]>(v));
          else haar_inv4(reinterpret_cast<double(&)[4]>(v));
          acc += v[(1 - dx) + 2 * (1 - dy) + (ND == 3 ? 4 * (1 - dz) : 0)];
        }
    y[n] = (mask && (mask[n] & 1)) ? 0.0 : acc;
  }
}

// hexa8, Haar-diagonal: a thread marches up to ZC nodes of one (i, j) column along z,
// keeping the 3 x 3 x 3 x values and 2 x 2 x 2 element weights around its node in
// registers (9 x loads and 4 weight loads per node instead of 64 and 8); consecutive
// threads take consecutive i.
template <int ZC>
__global__ void __launch_bounds__(256) k_heat_haar_col(Grid g, HeatOp op, const double *w, const uint8_t *mask,
                                                       const double *x, double *y, const SolverState *st) {
  if (st && st->done) return;
  const long long cols = (long long)g.sx * g.sy;
  const long long total = cols * ((g.sz + ZC - 1) / ZC);
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    const long long col = t % cols;
    const int i = (int)(col % g.sx), j = (int)(col / g.sx);
    const int k0 = (int)(t / cols) * ZC, k1 = min(k0 + ZC, g.sz);
    double P[3][3][3], W[2][2][2];
    auto plane = [&](int kk, double (&Q)[3][3]) {
#pragma unroll
      for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          const int ii = i - 1 + a, jj = j - 1 + b;
          const bool in = ii >= 0 && jj >= 0 && kk >= 0 && ii < g.sx && jj < g.sy && kk < g.sz;
          Q[b][a] = in ? x[ii + (long long)g.sx * (jj + (long long)g.sy * kk)] : 0.0;
        }
    };
    auto weights = [&](int ez, double (&V)[2][2]) {
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int a = 0; a < 2; ++a) {
          const int ex = i - 1 + a, ey = j - 1 + b;
          const bool in = ex >= 0 && ey >= 0 && ez >= 0 && ex < g.nx && ey < g.ny && ez < g.nz;
          V[b][a] = in ? w[ex + (long long)g.nx * (ey + (long long)g.ny * ez)] : -1.0;  // -1: no element
        }
    };
    plane(k0 - 1, P[0]);
    plane(k0, P[1]);
    weights(k0 - 1, W[0]);
    for (int k = k0; k < k1; ++k) {
      plane(k + 1, P[2]);
      weights(k, W[1]);
      double acc = 0.0;
#pragma unroll
      for (int dz = 0; dz < 2; ++dz)
#pragma unroll
        for (int dy = 0; dy < 2; ++dy)
#pragma unroll
          for (int dx = 0; dx < 2; ++dx) {
            const double we = W[dz][dy][dx];
            if (we < 0.0) continue;
            double v[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) v[q] = P[dz + (q >> 2)][dy + ((q >> 1) & 1)][dx + (q & 1)];
            haar_fwd8(v);
#pragma unroll
            for (int q = 0; q < 8; ++q) v[q] *= we * op.kf1[q] + (1.0 - we) * op.kf2[q];
            haar_inv8(v);
            acc += v[(1 - dx) + 2 * (1 - dy) + 4 * (1 - dz)];
          }
      const long long n = i + (long long)g.sx * (j + (long long)g.sy * k);
      y[n] = (mask && (mask[n] & 1)) ? 0.0 : acc;
#pragma unroll
      for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int a = 0; a < 3; ++a) { P[0][b][a] = P[1][b][a]; P[1][b][a] = P[2][b][a]; }
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int a = 0; a < 2; ++a) W[0][b][a] = W[1][b][a];
    }
  }
}

// hexa8, Haar-diagonal, elements transformed once: a 16 x 16 block owns the element
// columns (ex, ey) = (i0 - 1 + tx, j0 - 1 + ty) around its 15 x 15 node tile and marches
// the node layers k0 .. k0 + HZ - 1.  At layer k each thread transforms its element
// (ex, ey, k) once; the element's lower-face outputs plus the upper-face outputs kept
// from (ex, ey, k - 1) are its column's share of the four nodes at layer k, exchanged
// through LDS (double-buffered: one barrier per layer).  w = null: every element has the
// eigenvalues kf2 (the Helmholtz filter's operator).
constexpr int HT = 15, HZ = 8;
__global__ void __launch_bounds__(256) k_heat_haar_tile(Grid g, HeatOp op, const double *w, const uint8_t *mask,
                                                        const double *x, double *y, const SolverState *st) {
  if (st && st->done) return;
  __shared__ double sh[2][4][256];
  const int tiles_x = (g.sx + HT - 1) / HT, tiles_y = (g.sy + HT - 1) / HT;
  const int b = blockIdx.x;
  const int ti = b % tiles_x, tj = (b / tiles_x) % tiles_y, tk = b / (tiles_x * tiles_y);
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int i0 = ti * HT, j0 = tj * HT, k0 = tk * HZ, k1 = min(k0 + HZ, g.sz);
  const int ex = i0 - 1 + tx, ey = j0 - 1 + ty;
  const bool col = ex >= 0 && ey >= 0 && ex < g.nx && ey < g.ny;
  double dkf[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) dkf[q] = op.kf1[q] - op.kf2[q];
  // x on the element column's four corner lines at node layer kk (0 outside the grid)
  auto face = [&](int kk, double (&f)[4]) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int ii = ex + (c & 1), jj = ey + (c >> 1);
      const bool in = col && kk >= 0 && kk < g.sz;
      f[c] = in ? x[ii + (long long)g.sx * (jj + (long long)g.sy * kk)] : 0.0;
    }
  };
  // y_e of element (ex, ey, ez) from its lower / upper faces; false: no element there
  auto element = [&](int ez, const double (&lo)[4], const double (&hi)[4], double (&v)[8]) -> bool {
    if (!col || ez < 0 || ez >= g.nz) return false;
    const double we = w ? w[ex + (long long)g.nx * (ey + (long long)g.ny * ez)] : 0.0;
#pragma unroll
    for (int c = 0; c < 4; ++c) { v[c] = lo[c]; v[c + 4] = hi[c]; }
    haar_fwd8(v);
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] *= fma(we, dkf[q], op.kf2[q]);
    haar_inv8(v);
    return true;
  };
  double lo[4], hi[4], up[4] = {0.0, 0.0, 0.0, 0.0}, v[8];
  face(k0 - 1, lo);
  face(k0, hi);
  if (element(k0 - 1, lo, hi, v))
#pragma unroll
    for (int c = 0; c < 4; ++c) up[c] = v[c + 4];
  for (int k = k0; k < k1; ++k) {
#pragma unroll
    for (int c = 0; c < 4; ++c) lo[c] = hi[c];
    face(k + 1, hi);
    double cur[4] = {0.0, 0.0, 0.0, 0.0};
    if (element(k, lo, hi, v)) {
#pragma unroll
      for (int c = 0; c < 4; ++c) cur[c] = v[c];
    }
    const int buf = (k - k0) & 1;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      sh[buf][c][threadIdx.x] = cur[c] + up[c];
      up[c] = (col && k < g.nz) ? v[c + 4] : 0.0;
    }
    __syncthreads();
    // node (i0 + tx, j0 + ty): corner 3 of column (tx, ty), 2 of (tx+1, ty), 1 of (tx, ty+1), 0 of (tx+1, ty+1)
    const int i = i0 + tx, j = j0 + ty;
    if (tx < HT && ty < HT && i < g.sx && j < g.sy) {
      const int t = threadIdx.x;
      const double acc = sh[buf][3][t] + sh[buf][2][t + 1] + sh[buf][1][t + 16] + sh[buf][0][t + 17];
      const long long n = i + (long long)g.sx * (j + (long long)g.sy * k);
      y[n] = (mask && (mask[n] & 1)) ? 0.0 : acc;
    }
  }
}

void launch_helm_tile(const Grid &g, const double *kf, const double *x, double *y, const SolverState *st,
                      hipStream_t s) {
  HeatOp op{};
  for (int q = 0; q < 8; ++q) op.kf1[q] = op.kf2[q] = kf[q];
  op.haar = 1;
  const int nb = ((g.sx + HT - 1) / HT) * ((g.sy + HT - 1) / HT) * ((g.sz + HZ - 1) / HZ);
  hipLaunchKernelGGL(k_heat_haar_tile, dim3(nb), dim3(256), 0, s, g, op, nullptr, nullptr, x, y, st);
  ICX_CHECK_LAUNCH();
}

void launch_heat_matvec(const Grid &g, const HeatOp &op, const double *w, const uint8_t *mask, const double *x,
                        double *y, const SolverState *st, hipStream_t s, const LaunchEvents *ev) {
  if (op.haar && scalar_tile_kernel(g)) {
    const int nb = ((g.sx + HT - 1) / HT) * ((g.sy + HT - 1) / HT) * ((g.sz + HZ - 1) / HZ);
    launch_ev(k_heat_haar_tile, nb, 256, s, ev, g, op, w, mask, x, y, st);
    ICX_CHECK_LAUNCH();
    return;
  }
  // node layers per thread: the context's tile depth when it is 2, 4, 8 or 16, else 4
  // (measured at 100^3: 30.7, 28.2, 29.9, 36.7 us for 2, 4, 8, 16)
  const int zc = (g.tile_depth == 2 || g.tile_depth == 8 || g.tile_depth == 16) ? g.tile_depth : 4;
  const long long threads = (long long)g.sx * g.sy * ((g.sz + zc - 1) / zc);
  const int nb = (int)std::min<long long>((threads + 255) / 256, 65535);
  if (op.haar && g.ndim == 3) {
    if (zc == 2) launch_ev(k_heat_haar_col<2>, nb, 256, s, ev, g, op, w, mask, x, y, st);
    else if (zc == 8) launch_ev(k_heat_haar_col<8>, nb, 256, s, ev, g, op, w, mask, x, y, st);
    else if (zc == 16) launch_ev(k_heat_haar_col<16>, nb, 256, s, ev, g, op, w, mask, x, y, st);
    else launch_ev(k_heat_haar_col<4>, nb, 256, s, ev, g, op, w, mask, x, y, st);
  }
  else if (op.haar) launch_ev(k_heat_haar<2>, ew_blocks(g.nn), 256, s, ev, g, op, w, mask, x, y, st);
  else launch_ev(k_heat_op<false>, ew_blocks(g.nn), 256, s, ev, g, op, w, mask, x, y, st);
  ICX_CHECK_LAUNCH();
}
void launch_heat_invdiag(const Grid &g, const HeatOp &op, const double *w, const uint8_t *mask, double *invd,
                         hipStream_t s) {
  hipLaunchKernelGGL(k_heat_op<true>, dim3(ew_blocks(g.nn)), dim3(256), 0, s, g, op, w, mask, nullptr, invd,
                     nullptr);
  ICX_CHECK_LAUNCH();
}

// opt-heat/opt_heat.cpp:158-194 (the projection factor is applied by the caller)
__global__ void k_heat_sens(Grid g, HeatOp dK, double p, const double *design, const double *u, double vol,
                            double vscale, double *dfds, double *dgds) {
  const int nq = dK.nq;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < g.ne;
       e += (long long)gridDim.x * blockDim.x) {
    const int ex = (int)(e % g.nx), ey = (int)((e / g.nx) % g.ny), ez = (int)(e / ((long long)g.nx * g.ny));
    double t[8];
    for (int q = 0; q < nq; ++q) t[q] = flush15(u[corner(g, ex, ey, ez, q)]);
    double en = 0.0;
    for (int a = 0; a < nq; ++a) {
      double r = 0.0;
      for (int b = 0; b < nq; ++b) r += dK.K1[a * nq + b] * t[b];
      en += t[a] * r;
    }
    dfds[e] = -(p * pow(design[e], p - 1.0)) * en;
    if (dgds) dgds[e] = vol / vscale;
  }
}
void launch_heat_sensitivity(const Grid &g, const HeatOp &dK, double p, const double *design, const double *u,
                             double vol, double vscale, double *dfds, double *dgds, hipStream_t s) {
  hipLaunchKernelGGL(k_heat_sens, dim3(ew_blocks(g.ne)), dim3(256), 0, s, g, dK, p, design, u, vol, vscale, dfds,
                     dgds);
  ICX_CHECK_LAUNCH();
}

// ---------------------------------------------------------------------------------
// cone filter (src/optimization/filter.hpp:102-151); stencil offsets are sorted in
// increasing neighbour id, the order of the reference's neighbour lists.
// ---------------------------------------------------------------------------------
__device__ __forceinline__ bool nb_elem(const Grid &g, long long e, const int *off, long long &j) {
  const int ex = (int)(e % g.nx), ey = (int)((e / g.nx) % g.ny), ez = (int)(e / ((long long)g.nx * g.ny));
  const int X = ex + off[0], Y = ey + off[1], Z = ez + off[2];
  if (X < 0 || Y < 0 || Z < 0 || X >= g.nx || Y >= g.ny || Z >= g.nz) return false;
  j = X + (long long)g.nx * (Y + (long long)g.ny * Z);
  return true;
}

__global__ void k_cone_aa(Grid g, Stencil st, double *aa) {
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < g.ne;
       e += (long long)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int q = 0; q < st.n; ++q) {
      long long j;
      if (nb_elem(g, e, st.off + 3 * q, j)) s += st.w[q];
    }
    aa[e] = s;
  }
}
__global__ void k_cone_fwd(Grid g, Stencil st, const double *aa, const double *in, double *out) {
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < g.ne;
       e += (long long)gridDim.x * blockDim.x) {
    double bb = 0.0;
    for (int q = 0; q < st.n; ++q) {
      long long j;
      if (nb_elem(g, e, st.off + 3 * q, j)) bb += st.w[q] * in[j];
    }
    out[e] = bb / aa[e];
  }
}
__global__ void k_cone_adj(Grid g, Stencil st, const double *aa, const double *in, double *out) {
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < g.ne;
       e += (long long)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int q = 0; q < st.n; ++q) {
      long long j;
      if (nb_elem(g, e, st.off + 3 * q, j)) s += st.w[q] * in[j] / aa[j];
    }
    out[e] = s;
  }
}
void launch_cone_aa(const Grid &g, Stencil st, double *aa, hipStream_t s) {
  hipLaunchKernelGGL(k_cone_aa, dim3(ew_blocks(g.ne)), dim3(256), 0, s, g, st, aa);
  ICX_CHECK_LAUNCH();
}
void launch_cone_forward(const Grid &g, Stencil st, const double *aa, const double *in, double *out,
                         hipStream_t s) {
  hipLaunchKernelGGL(k_cone_fwd, dim3(ew_blocks(g.ne)), dim3(256), 0, s, g, st, aa, in, out);
  ICX_CHECK_LAUNCH();
}
void launch_cone_adjoint(const Grid &g, Stencil st, const double *aa, const double *in, double *out,
                         hipStream_t s) {
  hipLaunchKernelGGL(k_cone_adj, dim3(ew_blocks(g.ne)), dim3(256), 0, s, g, st, aa, in, out);
  ICX_CHECK_LAUNCH();
}

// Filter::sensitivityFilter (src/optimization/filter.hpp:72-100), the classic sensitivity
// filter, as the reference writes it: bb = sum_j (w x_j df_j) / (x_j v_j) in neighbour
// order, out = ((x_i v_i) / (x_i aa_i)) bb (an element with x = 0 gives 0 / 0, as there).
// v: per-element volumes (volv) or the box volume (vol).
__device__ __forceinline__ double sens_term(double w, double x, double df, double v) {
  return (w * x * df) / (x * v);
}
__global__ void k_cone_sens(Grid g, Stencil st, const double *aa, const double *x, const double *df, double vol,
                            double *out) {
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < g.ne;
       e += (long long)gridDim.x * blockDim.x) {
    double bb = 0.0;
    for (int q = 0; q < st.n; ++q) {
      long long j;
      if (nb_elem(g, e, st.off + 3 * q, j)) bb += sens_term(st.w[q], x[j], df[j], vol);
    }
    out[e] = ((x[e] * vol) / (x[e] * aa[e])) * bb;
  }
}
__global__ void k_cone_sens_csr(long long n, ConeCsr f, const double *aa, const double *x, const double *df,
                                const double *volv, double *out) {
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
       e += (long long)gridDim.x * blockDim.x) {
    double bb = 0.0;
    for (int k = f.ptr[e]; k < f.ptr[e + 1]; ++k) {
      const int j = f.id[k];
      bb += sens_term(f.w[k], x[j], df[j], volv[j]);
    }
    out[e] = ((x[e] * volv[e]) / (x[e] * aa[e])) * bb;
  }
}
void launch_cone_sensitivity(const Grid &g, Stencil st, const double *aa, const double *x, const double *df,
                             double vol, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_cone_sens, dim3(ew_blocks(g.ne)), dim3(256), 0, s, g, st, aa, x, df, vol, out);
  ICX_CHECK_LAUNCH();
}
void launch_cone_sensitivity_csr(long long n, ConeCsr f, const double *aa, const double *x, const double *df,
                                 const double *volv, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_cone_sens_csr, dim3(ew_blocks(n)), dim3(256), 0, s, n, f, aa, x, df, volv, out);
  ICX_CHECK_LAUNCH();
}

// threshold / thresholdDerivative, src/optimization/filter.hpp:153-185
__device__ __forceinline__ double proj(double x, double beta, double T) {
  return 0.5 * (tanh(T * beta) + tanh(beta * (x - T))) / tanh(T * beta);
}
__device__ __forceinline__ double dproj(double x, double beta, double T) {
  const double th = tanh(beta * (x - T));
  return (0.5 * (beta * (1.0 - th * th))) / tanh(T * beta);
}
__global__ void k_project(long long n, double beta, double T, const double *in, double *out, double *dout) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double x = in[i];
    if (out) out[i] = proj(x, beta, T);
    if (dout) dout[i] = dproj(x, beta, T);
  }
}
void launch_project(long long n, double beta, double T, const double *in, double *out, double *dout,
                    hipStream_t s) {
  hipLaunchKernelGGL(k_project, dim3(ew_blocks(n)), dim3(256), 0, s, n, beta, T, in, out, dout);
  ICX_CHECK_LAUNCH();
}

// ---------------------------------------------------------------------------------
// Helmholtz rhs f_n = scale * sum_{e ni n} rho_e, element value scale * sum_{n in e} phi_n
// ---------------------------------------------------------------------------------
__global__ void k_helm_rhs(Grid g, double scale, const double *rho, double *f) {
  for (long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x; n < g.nn;
       n += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(n % g.sx), j = (int)((n / g.sx) % g.sy), k = (int)(n / ((long long)g.sx * g.sy));
    double s = 0.0;
    for (int dz = (g.ndim == 3 ? -1 : 0); dz <= 0; ++dz)
      for (int dy = -1; dy <= 0; ++dy)
        for (int dx = -1; dx <= 0; ++dx) {
          const int ex = i + dx, ey = j + dy, ez = k + dz;
          if (ex < 0 || ey < 0 || ez < 0 || ex >= g.nx || ey >= g.ny || ez >= g.nz) continue;
          s += rho[ex + (long long)g.nx * (ey + (long long)g.ny * ez)] * scale;
        }
    f[n] = s;
  }
}
void launch_helm_rhs(const Grid &g, double scale, const double *rho, double *f, hipStream_t s) {
  hipLaunchKernelGGL(k_helm_rhs, dim3(ew_blocks(g.nn)), dim3(256), 0, s, g, scale, rho, f);
  ICX_CHECK_LAUNCH();
}
__global__ void k_helm_to_element(Grid g, double scale, const double *phi, double *out) {
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < g.ne;
       e += (long long)gridDim.x * blockDim.x) {
    const int ex = (int)(e % g.nx), ey = (int)((e / g.nx) % g.ny), ez = (int)(e / ((long long)g.nx * g.ny));
    const int nq = g.ndim == 3 ? 8 : 4;
    double s = 0.0;
    for (int q = 0; q < nq; ++q)
      s += phi[(ex + (q & 1)) + (long long)g.sx * ((ey + ((q >> 1) & 1)) + (long long)g.sy * (ez + (q >> 2)))];
    out[e] = s * scale;
  }
}
void launch_helm_to_element(const Grid &g, double scale, const double *phi, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_helm_to_element, dim3(ew_blocks(g.ne)), dim3(256), 0, s, g, scale, phi, out);
  ICX_CHECK_LAUNCH();
}

// ---------------------------------------------------------------------------------
// OC, src/optimization/opt-solver/OCSolver.hpp:33-48
// ---------------------------------------------------------------------------------
__device__ __forceinline__ double oc_x(double lambda, double move, double xmin, double xmax, double x,
                                       double dfdx, double dgdx) {
  const double t = -dfdx / (dgdx * lambda);
  double v = pow(t > 0.0 ? t : 0.0, 0.5) * x;
  const double lo = fmax(xmin, (1.0 - move) * x);
  const double hi = fmin(xmax, (1.0 + move) * x);
  if (v < lo) v = lo;
  else if (v > hi) v = hi;
  return v;
}
__global__ void k_oc_init(OcBisect *b, double g0) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    b->l0 = 0.0;
    b->l1 = 1.0e4;
    b->lambda = 0.0;
    b->g0 = g0;
    b->counter = 0;
    b->done = 0;
  }
}
void launch_oc_init(OcBisect *b, double g0, hipStream_t s) {
  hipLaunchKernelGGL(k_oc_init, dim3(1), dim3(64), 0, s, b, g0);
  ICX_CHECK_LAUNCH();
}
__global__ void k_oc_decide(OcBisect *b, const double *gk) {
  if (threadIdx.x != 0 || blockIdx.x != 0 || b->done) return;
  const double lambda = 0.5 * (b->l1 + b->l0);
  if (gk[0] - b->g0 > 0.0) b->l0 = lambda;
  else b->l1 = lambda;
  b->lambda = lambda;
  b->counter += 1;
  if (!((b->l1 - b->l0) / (b->l1 + b->l0) > 1.0e-5 && 100 > b->counter)) b->done = 1;
}
void launch_oc_decide(OcBisect *b, const double *gk, hipStream_t s) {
  hipLaunchKernelGGL(k_oc_decide, dim3(1), dim3(64), 0, s, b, gk);
  ICX_CHECK_LAUNCH();
}
// lambda of an OC kernel: the argument, or from the device bisection state (the step's
// midpoint, or the last evaluated step's lambda for the final candidate)
__device__ __forceinline__ double oc_lambda(double lambda, const OcBisect *bs, bool mid) {
  if (!bs) return lambda;
  return mid ? 0.5 * (bs->l1 + bs->l0) : bs->lambda;
}

__global__ void k_oc_candidate(long long n, double lambda_arg, double move, double xmin, double xmax,
                               const double *x, const double *dfdx, const double *dgdx, double *xnew,
                               const OcBisect *bs, bool mid) {
  if (bs && mid && bs->done) return;  // a bisection step queued past the stop
  const double lambda = oc_lambda(lambda_arg, bs, mid);
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    xnew[i] = oc_x(lambda, move, xmin, xmax, x[i], dfdx[i], dgdx[i]);
}
void launch_oc_candidate(long long n, double lambda, double move, double xmin, double xmax,
                         const double *x, const double *dfdx, const double *dgdx, double *xnew,
                         hipStream_t s, const OcBisect *bs, bool mid) {
  hipLaunchKernelGGL(k_oc_candidate, dim3(ew_blocks(n)), dim3(256), 0, s, n, lambda, move, xmin, xmax, x,
                     dfdx, dgdx, xnew, bs, mid);
  ICX_CHECK_LAUNCH();
}

__global__ void __launch_bounds__(256) k_volume(long long n, int project, double beta, double T,
                                                const double *rho, double vol, const double *volv, Reduce red) {
  __shared__ double scr[4];
  double s = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    s += (project ? proj(rho[i], beta, T) : rho[i]) * (volv ? volv[i] : vol);
  s = block_sum<256>(s, scr);
  if (threadIdx.x == 0) red.partial[blockIdx.x] = s;
}
void launch_volume(long long n, int project, double beta, double T, const double *rho, double vol,
                   const double *volv, double *out, Reduce red, hipStream_t s) {
  const int nb = ew_blocks(n, 4, RED_BLOCKS);
  hipLaunchKernelGGL(k_volume, dim3(nb), dim3(256), 0, s, n, project, beta, T, rho, vol, volv, red);
  launch_finalize(red, nb, 1, out, s);
}

// Cone filter on explicit neighbour lists (unstructured meshes): the same sums as the
// stencil kernels, over the CSR rows (neighbours in increasing element id)
__global__ void k_cone_csr(long long n, ConeCsr f, const double *aa, const double *in, double *out, int adjoint) {
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
       e += (long long)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int k = f.ptr[e]; k < f.ptr[e + 1]; ++k) {
      const int j = f.id[k];
      if (adjoint) s += f.w[k] * in[j] / aa[j];
      else s += f.w[k] * in[j];
    }
    out[e] = adjoint ? s : s / aa[e];
  }
}
void launch_cone_csr(long long n, ConeCsr f, const double *aa, const double *in, double *out, int adjoint,
                     hipStream_t s) {
  hipLaunchKernelGGL(k_cone_csr, dim3(ew_blocks(n)), dim3(256), 0, s, n, f, aa, in, out, adjoint);
  ICX_CHECK_LAUNCH();
}

// one bisection step of the OC volume lambda with the cone filter, fused:
// sum_i threshold( sum_j w_ij x_new_j(lambda) / aa_i ) v
__global__ void __launch_bounds__(256) k_oc_volume_cone(Grid g, Stencil st, const double *aa, double lambda_arg,
                                                        double move, double xmin, double xmax, const double *x,
                                                        const double *dfdx, const double *dgdx, int project,
                                                        double beta, double T, double vol, Reduce red,
                                                        const OcBisect *bs) {
  __shared__ double scr[4];
  if (bs && bs->done) return;  // a bisection step queued past the stop (uniform exit)
  const double lambda = oc_lambda(lambda_arg, bs, true);
  double s = 0.0;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < g.ne;
       e += (long long)gridDim.x * blockDim.x) {
    double bb = 0.0;
    for (int q = 0; q < st.n; ++q) {
      long long j;
      if (nb_elem(g, e, st.off + 3 * q, j))
        bb += st.w[q] * oc_x(lambda, move, xmin, xmax, x[j], dfdx[j], dgdx[j]);
    }
    const double f = bb / aa[e];
    s += (project ? proj(f, beta, T) : f) * vol;
  }
  s = block_sum<256>(s, scr);
  if (threadIdx.x == 0) red.partial[blockIdx.x] = s;
}
// the fused OC bisection step on explicit neighbour lists, per-element volumes
__global__ void __launch_bounds__(256) k_oc_volume_csr(long long n, ConeCsr f, const double *aa,
                                                       double lambda_arg, double move, double xmin, double xmax,
                                                       const double *x, const double *dfdx, const double *dgdx,
                                                       int project, double beta, double T, const double *vol,
                                                       Reduce red, const OcBisect *bs) {
  __shared__ double scr[4];
  if (bs && bs->done) return;  // a bisection step queued past the stop (uniform exit)
  const double lambda = oc_lambda(lambda_arg, bs, true);
  double s = 0.0;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
       e += (long long)gridDim.x * blockDim.x) {
    double bb = 0.0;
    for (int k = f.ptr[e]; k < f.ptr[e + 1]; ++k) {
      const int j = f.id[k];
      bb += f.w[k] * oc_x(lambda, move, xmin, xmax, x[j], dfdx[j], dgdx[j]);
    }
    const double v = bb / aa[e];
    s += (project ? proj(v, beta, T) : v) * vol[e];
  }
  s = block_sum<256>(s, scr);
  if (threadIdx.x == 0) red.partial[blockIdx.x] = s;
}
void launch_oc_volume_csr(long long n, ConeCsr f, const double *aa, double lambda, double move, double xmin,
                          double xmax, const double *x, const double *dfdx, const double *dgdx, int project,
                          double beta, double T, const double *vol, double *out, Reduce red, hipStream_t s,
                          const OcBisect *bs) {
  const int nb = ew_blocks(n, 1, RED_BLOCKS);
  hipLaunchKernelGGL(k_oc_volume_csr, dim3(nb), dim3(256), 0, s, n, f, aa, lambda, move, xmin, xmax, x, dfdx,
                     dgdx, project, beta, T, vol, red, bs);
  launch_finalize(red, nb, 1, out, s);
}

void launch_oc_volume_cone(const Grid &g, Stencil st, const double *aa, double lambda, double move,
                           double xmin, double xmax, const double *x, const double *dfdx,
                           const double *dgdx, int project, double beta, double T, double vol,
                           double *out, Reduce red, hipStream_t s, const OcBisect *bs) {
  const int nb = ew_blocks(g.ne, 1, RED_BLOCKS);
  hipLaunchKernelGGL(k_oc_volume_cone, dim3(nb), dim3(256), 0, s, g, st, aa, lambda, move, xmin, xmax, x,
                     dfdx, dgdx, project, beta, T, vol, red, bs);
  launch_finalize(red, nb, 1, out, s);
}

// ---------------------------------------------------------------------------------
// MMA (m = 1), src/optimization/opt-solver/MMASolver.cpp:400-512 (GenSub) and
// 184-398 (XYZofLAMBDA, DualGrad, DualHess, DualResidual sums); GCMMA's Asymp and GenSub
// (GCMMASolver.cpp:432-489, 546-596) are the same pass with raa0 / raa in place of MMA's
// one raa0, no clamp of alpha to xmax and of beta to xmin, and, in inner updates, the
// asymptotes and x history left as they are (asymp = 0).
// sums: [0] = sum p0/(U-x) + q0/(x-L), [1] = sum p1/(U-x) + q1/(x-L), both at xval.
// ---------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_mma_gensub(long long n, int iter, int asymp, int clamp, MMAVecs m,
                                                    const double *xval, const double *dfdx, const double *dgdx,
                                                    double xmin, double xmax, double move, double albefa,
                                                    double asyminit, double asymdec, double asyminc, double raa0,
                                                    double raa1, double xmamieps, Reduce red) {
  __shared__ double scr[4];
  double s = 0.0, s0 = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double xv = xval[i];
    double low, upp;
    if (!asymp) {
      low = m.low[i]; upp = m.upp[i];
    } else if (iter < 3) {
      low = xv - asyminit * (xmax - xmin);
      upp = xv + asyminit * (xmax - xmin);
    } else {
      low = m.low[i]; upp = m.upp[i];
      const double xo1 = m.xold1[i], xo2 = m.xold2[i];
      const double zzz = (xv - xo1) * (xo1 - xo2);
      const double gamma = zzz < 0.0 ? asymdec : (zzz > 0.0 ? asyminc : 1.0);
      low = xv - gamma * (xo1 - low);
      upp = xv + gamma * (upp - xo1);
      const double xmami = fmax(xmamieps, xmax - xmin);
      low = fmax(low, xv - 100.0 * xmami);
      low = fmin(low, xv - 1.0e-5 * xmami);
      upp = fmax(upp, xv + 1.0e-5 * xmami);
      upp = fmin(upp, xv + 100.0 * xmami);
      const double xmi = xmin - 1.0e-6, xma = xmax + 1.0e-6;
      if (xv < xmi) { low = xv - (xma - xv) / 0.9; upp = xv + (xma - xv) / 0.9; }
      if (xv > xma) { low = xv - (xv - xmi) / 0.9; upp = xv + (xv - xmi) / 0.9; }
    }
    if (asymp) {
      m.low[i] = low; m.upp[i] = upp;
      m.xold2[i] = m.xold1[i];
      m.xold1[i] = xv;
    }
    double al = fmax(xmin, low + albefa * (xv - low));
    al = fmax(al, xv - move * (xmax - xmin));
    m.alpha[i] = clamp ? fmin(al, xmax) : al;
    double be = fmin(xmax, upp - albefa * (upp - xv));
    be = fmin(be, xv + move * (xmax - xmin));
    m.beta[i] = clamp ? fmax(be, xmin) : be;
    const double xmamiinv = 1.0 / fmax(xmamieps, xmax - xmin);
    const double df = dfdx[i];
    const double pq = 0.001 * fabs(df) + raa0 * xmamiinv;
    const double u2 = pow(upp - xv, 2.0), l2 = pow(xv - low, 2.0);
    const double p0 = u2 * (fmax(0.0, df) + pq), q0 = l2 * (fmax(0.0, -1.0 * df) + pq);
    m.p0[i] = p0; m.q0[i] = q0;
    const double dg = dgdx[i];
    const double pqj = 0.001 * fabs(dg) + raa1 * xmamiinv;
    const double p1 = u2 * (fmax(0.0, dg) + pqj), q1 = l2 * (fmax(0.0, -1.0 * dg) + pqj);
    m.p1[i] = p1; m.q1[i] = q1;
    s0 += p0 / (upp - xv) + q0 / (xv - low);
    s += p1 / (upp - xv) + q1 / (xv - low);
  }
  s0 = block_sum<256>(s0, scr);
  s = block_sum<256>(s, scr);
  if (threadIdx.x == 0) {
    red.partial[blockIdx.x] = s0;
    red.partial[RED_MAXBLOCKS + blockIdx.x] = s;
  }
}
void launch_mma_gensub(long long n, int iter, int asymp, int clamp, const MMAVecs &m, const double *xval,
                       const double *dfdx, const double *dgdx, double xmin, double xmax,
                       double move, double albefa, double asyminit, double asymdec, double asyminc,
                       double raa0, double raa1, double xmamieps, double *sums, Reduce red, hipStream_t s) {
  const int nb = ew_blocks(n, 4, RED_BLOCKS);
  hipLaunchKernelGGL(k_mma_gensub, dim3(nb), dim3(256), 0, s, n, iter, asymp, clamp, m, xval, dfdx, dgdx, xmin,
                     xmax, move, albefa, asyminit, asymdec, asyminc, raa0, raa1, xmamieps, red);
  launch_finalize(red, nb, 2, sums, s);
}

// GCMMA ComputeApprox and RaaUpdate's raacof (GCMMASolver.cpp:598-620, 510-525) at xmma:
// out[0] = sum p0/(U-x) + q0/(x-L), out[1] = the same with p1 / q1,
// out[2] = sum (x-xval)^2 / ((U-x)(x-L)) * (U-L) / max(xmamieps, xmax-xmin)
__global__ void __launch_bounds__(256) k_gcmma_approx(long long n, MMAVecs m, const double *xmma,
                                                      const double *xval, double xmami, Reduce red) {
  __shared__ double scr[4];
  double a0 = 0.0, a1 = 0.0, rc = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double x = xmma[i], xv = xval[i], low = m.low[i], upp = m.upp[i];
    const double uxinv = 1.0 / (upp - x), xlinv = 1.0 / (x - low);
    a0 += m.p0[i] * uxinv + m.q0[i] * xlinv;
    a1 += m.p1[i] * uxinv + m.q1[i] * xlinv;
    const double xxux = (x - xv) / (upp - x), xxxl = (x - xv) / (x - low);
    rc += xxux * xxxl * ((upp - low) / xmami);
  }
  a0 = block_sum<256>(a0, scr);
  a1 = block_sum<256>(a1, scr);
  rc = block_sum<256>(rc, scr);
  if (threadIdx.x == 0) {
    red.partial[blockIdx.x] = a0;
    red.partial[RED_MAXBLOCKS + blockIdx.x] = a1;
    red.partial[2 * RED_MAXBLOCKS + blockIdx.x] = rc;
  }
}
void launch_gcmma_approx(long long n, const MMAVecs &m, const double *xmma, const double *xval, double xmami,
                         double *out, Reduce red, hipStream_t s) {
  const int nb = ew_blocks(n, 4, RED_BLOCKS);
  hipLaunchKernelGGL(k_gcmma_approx, dim3(nb), dim3(256), 0, s, n, m, xmma, xval, xmami, red);
  launch_finalize(red, nb, 3, out, s);
}

__global__ void __launch_bounds__(256) k_mma_dual(long long n, double lam, MMAVecs m, double *x, Reduce red,
                                                  const MmaDual *d) {
  __shared__ double scr[4];
  if (d) {
    if (d->done) return;
    lam = d->lam;
  }
  double s1 = 0.0, s2 = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double low = m.low[i], upp = m.upp[i];
    const double pj = m.p0[i] + m.p1[i] * lam, qj = m.q0[i] + m.q1[i] * lam;
    const double xp = (sqrt(pj) * low + sqrt(qj) * upp) / (sqrt(pj) + sqrt(qj));
    double xi = xp;
    if (xi < m.alpha[i]) xi = m.alpha[i];
    if (xi > m.beta[i]) xi = m.beta[i];
    x[i] = xi;
    s1 += m.p1[i] / (upp - xi) + m.q1[i] / (xi - low);
    const double PQ = m.p1[i] / pow(upp - xi, 2.0) - m.q1[i] / pow(xi - low, 2.0);
    double df2 = -1.0 / (2.0 * pj / pow(upp - xi, 3.0) + 2.0 * qj / pow(xi - low, 3.0));
    if (xp < m.alpha[i]) df2 = 0.0;
    if (xp > m.beta[i]) df2 = 0.0;
    s2 += (0.0 + PQ * df2) * PQ;
  }
  s1 = block_sum<256>(s1, scr);
  s2 = block_sum<256>(s2, scr);
  if (threadIdx.x == 0) {
    red.partial[blockIdx.x] = s1;
    red.partial[RED_MAXBLOCKS + blockIdx.x] = s2;
  }
}
void launch_mma_dual(long long n, double lam, const MMAVecs &m, double *x, double *out, Reduce red,
                     hipStream_t s, const MmaDual *d) {
  const int nb = ew_blocks(n, 4, RED_BLOCKS);
  hipLaunchKernelGGL(k_mma_dual, dim3(nb), dim3(256), 0, s, n, lam, m, x, red, d);
  launch_finalize(red, nb, 2, out, s);
}

// The dual loop's scalar steps, in the reference's operation order (this library is
// built without contraction), so that every value is the one the host loop computes
// from the same sums.  xyz's prologue: lambda >= 0, y and z of lambda.
__device__ __forceinline__ void mma_dual_yz(MmaDual &s) {
  const double a = 0.0;
  if (s.lam < 0.0) s.lam = 0;
  const double ly = s.lam - s.cc;
  s.y = 0.0 < ly ? ly : 0.0;
  const double lz = 10.0 * (s.lam * a - 1.0);
  s.z = 0.0 < lz ? lz : 0.0;
}
__global__ void k_mma_dual_init(MmaDual *d, double b, double cc, double epsimin) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  MmaDual s;
  s.lam = cc / 2.0;
  s.mu = 1.0;
  s.epsi = 1.0;
  s.err = 1.0;
  s.b = b;
  s.cc = cc;
  s.epsimin = epsimin;
  s.outloop = s.loop = s.total = 0;
  s.in_inner = s.need_err = s.done = 0;
  mma_dual_yz(s);
  *d = s;
}
void launch_mma_dual_init(MmaDual *d, double b, double cc, double epsimin, hipStream_t s) {
  hipLaunchKernelGGL(k_mma_dual_init, dim3(1), dim3(64), 0, s, d, b, cc, epsimin);
  ICX_CHECK_LAUNCH();
}
// Takes the sums at d->lam: the residual of the last Newton step if one is pending, then
// the loops' tests (inner: err > 0.9 epsi and loop < 100; outer: epsi > epsimin and
// outloop < 100, epsi /= 10 after each inner loop) until the next Newton step, which it
// takes, or the end.  Every pass of the for loop either takes a step and leaves, or
// ends an inner loop, or starts one of at most 100 outer loops, so it ends.
__global__ void k_mma_dual_advance(MmaDual *d, const double *sums) {
  if (threadIdx.x != 0 || blockIdx.x != 0 || d->done) return;
  MmaDual s = *d;
  const double a = 0.0;
  const double S1 = sums[0], S2 = sums[1];
  if (s.need_err) {
    const double r0 = -s.b - a * s.z - s.y + s.mu + S1;
    const double r1 = s.mu * s.lam - s.epsi;
    const double e0 = fabs(r0), e1 = fabs(r1);
    s.err = e0 < e1 ? e1 : e0;
    s.need_err = 0;
  }
  for (;;) {
    if (s.in_inner) {
      if (s.err > 0.9 * s.epsi && s.loop < 100) {
        s.loop++;
        s.total++;
        double grad = -s.b - a * s.z - s.y + S1;
        grad = -1.0 * grad - s.epsi / s.lam;
        if (s.lam < 0.0) s.lam = 0.0;
        double hess = S2;
        if (s.lam > s.cc) hess += -1.0;
        hess += -s.mu / s.lam;
        double corr = 1e-4 * hess / 1;
        if (-1.0 * corr < 1.0e-7) corr = -1.0e-7;
        hess += corr;
        const double s0 = grad / hess;
        const double s1 = -s.mu + s.epsi / s.lam - s0 * s.mu / s.lam;
        double theta = 1.005;
        if (theta < -1.01 * s0 / s.lam) theta = -1.01 * s0 / s.lam;
        if (theta < -1.01 * s1 / s.mu) theta = -1.01 * s1 / s.mu;
        theta = 1.0 / theta;
        s.lam = s.lam + theta * s0;
        s.mu = s.mu + theta * s1;
        mma_dual_yz(s);
        s.need_err = 1;
        break;
      }
      s.in_inner = 0;
      s.epsi = s.epsi * 0.1;
    }
    if (s.epsi > s.epsimin && s.outloop < 100) {
      s.outloop++;
      s.loop = 0;
      s.in_inner = 1;
      continue;
    }
    s.done = 1;
    break;
  }
  *d = s;
}
void launch_mma_dual_advance(MmaDual *d, const double *sums, hipStream_t s) {
  hipLaunchKernelGGL(k_mma_dual_advance, dim3(1), dim3(64), 0, s, d, sums);
  ICX_CHECK_LAUNCH();
}

}  // namespace icx
