/*
 * icx_oracle.c -- CPU restatement (FP64, OpenMP) of ICARAT's topology
 * optimisation hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the HIP library and the
 * timed CPU baseline.  Nothing in the product library links or calls this.
 *
 * Each function restates the reference source named in its comment (paths
 * relative to the reference repository root).  The reference itself cannot be
 * built offline (Eigen 3.4.0 and toml11 are fetched at configure time), so
 * this file is written from the reference's semantics; where the reference
 * delegates to Eigen (ConjugateGradient with DiagonalPreconditioner,
 * SimplicialLDLT) the published Eigen 3.4 algorithm is restated instead.
 */
#include "icx_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXNE 8
#define MAXDOF 24

/* ------------------------------------------------------------------------ */
/* mesh: src/io/mesh_icarat.hpp:93-110 (counts), 325-355 (quad4),            */
/*       440-494 (hexa8)                                                      */
/* ------------------------------------------------------------------------ */
int icxo_mesh_box(int ne, const int *div, const double *len, double *x, int *conn)
{
  if (ne == 4) {
    for (int i = 0; i < div[1] + 1; i++)
      for (int j = 0; j < div[0] + 1; j++) {
        int n = j + i + div[0] * i;
        x[3 * n + 0] = (len[0] / div[0]) * j;
        x[3 * n + 1] = (len[1] / div[1]) * i;
        x[3 * n + 2] = 0.0;
      }
    for (int i = 0; i < div[1]; i++)
      for (int j = 0; j < div[0]; j++) {
        int e = div[0] * i + j;
        int n0 = div[0] * i + j + i;
        conn[4 * e + 0] = n0;
        conn[4 * e + 1] = n0 + 1;
        conn[4 * e + 2] = n0 + 1 + (div[0] + 1);
        conn[4 * e + 3] = n0 + (div[0] + 1);
      }
    return 0;
  }
  if (ne == 8) {
    const int sx = div[0] + 1, sxy = (div[0] + 1) * (div[1] + 1);
    for (int i = 0; i < div[2] + 1; i++)
      for (int j = 0; j < div[1] + 1; j++)
        for (int k = 0; k < div[0] + 1; k++) {
          int n = k + j * sx + i * sxy;
          x[3 * n + 0] = (len[0] / div[0]) * k;
          x[3 * n + 1] = (len[1] / div[1]) * j;
          x[3 * n + 2] = (len[2] / div[2]) * i;
        }
    for (int i = 0; i < div[2]; i++)
      for (int j = 0; j < div[1]; j++)
        for (int k = 0; k < div[0]; k++) {
          int e = div[0] * div[1] * i + div[0] * j + k;
          int b = sxy * i + sx * j + k;
          int *c = conn + 8 * e;
          c[0] = b;           c[1] = b + 1;
          c[2] = b + 1 + sx;  c[3] = b + sx;
          c[4] = b + sxy;     c[5] = b + 1 + sxy;
          c[6] = b + 1 + sx + sxy; c[7] = b + sx + sxy;
        }
    return 0;
  }
  return -1;
}

/* ------------------------------------------------------------------------ */
/* Dirichlet box ranges: src/io/boundary_condition_icarat.hpp:68-121 and     */
/* setDirichValue src/io/boundary_condition_base.hpp:18-33 (flags OR-ed, the */
/* last matching entry's values win).                                        */
/* ------------------------------------------------------------------------ */
void icxo_dirichlet(int ndim, int dofnp, int numnp, const double *x,
                    const int *flag6, const double *val6, const double *range6,
                    int *has, int *node_flag, double *node_val)
{
  const double tol = 1.0e-7;
  double xmin = range6[0] - tol, xmax = range6[1] + tol;
  double ymin = range6[2] - tol, ymax = range6[3] + tol;
  double zmin = range6[4] - tol, zmax = range6[5] + tol;
  for (int j = 0; j < numnp; j++) {
    const double *p = x + 3 * j;
    int hit = p[0] >= xmin && p[0] <= xmax && p[1] >= ymin && p[1] <= ymax;
    if (ndim == 3) hit = hit && p[2] >= zmin && p[2] <= zmax;
    if (!hit) continue;
    if (!has[j]) {
      has[j] = 1;
      for (int k = 0; k < 6; k++) { node_flag[6 * j + k] = 0; node_val[6 * j + k] = 0.0; }
    }
    for (int k = 0; k < dofnp; k++) {
      if (node_flag[6 * j + k] != 1) node_flag[6 * j + k] = flag6[k];
      node_val[6 * j + k] = val6[k];
    }
  }
}

/* countDOF: src/io/boundary_condition_base.hpp:35-88.  Free DOFs first, in
 * node order; constrained DOFs numeq..neq-1 afterwards, also in node order. */
int icxo_count_dof(int numnp, int dofnp, const int *has, const int *node_flag, int *dof)
{
  int counter = 0;
  for (int n = 0; n < numnp; n++)
    for (int j = 0; j < dofnp; j++)
      if (!has[n] || node_flag[6 * n + j] == 0) dof[dofnp * n + j] = counter++;
  int numeq = counter;
  for (int n = 0; n < numnp; n++)
    if (has[n])
      for (int j = 0; j < dofnp; j++)
        if (node_flag[6 * n + j] != 0) dof[dofnp * n + j] = counter++;
  return numeq;
}

/* ------------------------------------------------------------------------ */
/* shape functions / Gauss rules: src/basis/shape_function.cpp:12-26,        */
/* 97-119, 403-447; src/basis/gauss_integration.cpp:12-22, 72-88, 268-289    */
/* ------------------------------------------------------------------------ */
static void gauss_point(int ne, int ip, double *r)
{
  const double g = 1.0 / sqrt(3.0);
  if (ne == 4) {
    static const int s[4][2] = {{-1, -1}, {1, -1}, {-1, 1}, {1, 1}};
    r[0] = s[ip][0] * g; r[1] = s[ip][1] * g;
  } else {
    static const int s[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
    r[0] = s[ip][0] * g; r[1] = s[ip][1] * g; r[2] = s[ip][2] * g;
  }
}

static void shape(int ne, const double *r, double *N, double *dNdr /* d x ne */)
{
  if (ne == 4) {
    N[0] = 0.25 * (1.0 - r[0]) * (1.0 - r[1]);
    N[1] = 0.25 * (1.0 + r[0]) * (1.0 - r[1]);
    N[2] = 0.25 * (1.0 + r[0]) * (1.0 + r[1]);
    N[3] = 0.25 * (1.0 - r[0]) * (1.0 + r[1]);
    dNdr[0] = -0.25 * (1.0 - r[1]); dNdr[1] = 0.25 * (1.0 - r[1]);
    dNdr[2] = 0.25 * (1.0 + r[1]);  dNdr[3] = -0.25 * (1.0 + r[1]);
    dNdr[4] = -0.25 * (1.0 - r[0]); dNdr[5] = -0.25 * (1.0 + r[0]);
    dNdr[6] = 0.25 * (1.0 + r[0]);  dNdr[7] = 0.25 * (1.0 - r[0]);
    return;
  }
  static const int sg[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                               {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
  for (int n = 0; n < 8; n++) {
    double a = 1.0 + sg[n][0] * r[0], b = 1.0 + sg[n][1] * r[1], c = 1.0 + sg[n][2] * r[2];
    N[n] = 0.125 * a * b * c;
    dNdr[0 * 8 + n] = 0.125 * sg[n][0] * b * c;
    dNdr[1 * 8 + n] = 0.125 * sg[n][1] * c * a;
    dNdr[2 * 8 + n] = 0.125 * sg[n][2] * a * b;
  }
}

/* J = dNdr * X, dNdX = J^-1 dNdr, jac = det J * w: src/basis/bmatrix.hpp:28-87, 89-162 */
static double kinematics(int ne, int d, int ip, const double *X, double *N, double *dNdX)
{
  double r[3], dNdr[3 * MAXNE], J[9] = {0}, Ji[9], det;
  gauss_point(ne, ip, r);
  shape(ne, r, N, dNdr);
  for (int a = 0; a < d; a++)
    for (int b = 0; b < d; b++) {
      double s = 0.0;
      for (int n = 0; n < ne; n++) s += dNdr[a * ne + n] * X[n * d + b];
      J[a * d + b] = s;
    }
  if (d == 2) {
    det = J[0] * J[3] - J[1] * J[2];
    Ji[0] = J[3] / det; Ji[1] = -J[1] / det; Ji[2] = -J[2] / det; Ji[3] = J[0] / det;
  } else {
    det = J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) +
          J[2] * (J[3] * J[7] - J[4] * J[6]);
    Ji[0] = (J[4] * J[8] - J[5] * J[7]) / det; Ji[1] = (J[2] * J[7] - J[1] * J[8]) / det;
    Ji[2] = (J[1] * J[5] - J[2] * J[4]) / det; Ji[3] = (J[5] * J[6] - J[3] * J[8]) / det;
    Ji[4] = (J[0] * J[8] - J[2] * J[6]) / det; Ji[5] = (J[2] * J[3] - J[0] * J[5]) / det;
    Ji[6] = (J[3] * J[7] - J[4] * J[6]) / det; Ji[7] = (J[1] * J[6] - J[0] * J[7]) / det;
    Ji[8] = (J[0] * J[4] - J[1] * J[3]) / det;
  }
  for (int a = 0; a < d; a++)
    for (int n = 0; n < ne; n++) {
      double s = 0.0;
      for (int b = 0; b < d; b++) s += Ji[a * d + b] * dNdr[b * ne + n];
      dNdX[a * ne + n] = s;
    }
  return det * 1.0; /* Gauss weights are 1 for both rules */
}

/* SIMP: src/problem/elasticity.hpp:15-22 */
double icxo_simp(double y1, double y2, double s, double p)
{
  if (y2 - y1 >= 0.0) return (1.0 - pow(s, p)) * y1 + pow(s, p) * y2;
  return pow(1.0 - s, p) * y1 + (1.0 - pow(1.0 - s, p)) * y2;
}

/* De: src/problem/elasticity.cpp:34-98 */
static int make_De(int state, double E, double nu, double *D, int *voigt)
{
  if (state == ICXO_PLANE_STRESS || state == ICXO_PLANE_STRAIN) {
    *voigt = 3;
    memset(D, 0, 9 * sizeof(double));
    double f;
    if (state == ICXO_PLANE_STRESS) {
      D[0] = 1.0; D[1] = nu; D[4] = 1.0; D[8] = 0.5 * (1.0 - nu);
      f = E / (1.0 - nu * nu);
    } else {
      D[0] = 1.0; D[1] = nu / (1.0 - nu); D[4] = 1.0;
      D[8] = (1.0 - 2.0 * nu) / (2.0 * (1.0 - nu));
      f = E * ((1.0 - nu) / ((1.0 - 2.0 * nu) * (1.0 + nu)));
    }
    D[3] = D[1];
    for (int i = 0; i < 9; i++) D[i] *= f;
    return 0;
  }
  *voigt = 6;
  memset(D, 0, 36 * sizeof(double));
  double lambda = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
  double mu = E / (2 * (1.0 + nu));
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) D[i * 6 + j] = (i == j) ? lambda + 2 * mu : lambda;
  D[21] = mu; D[28] = mu; D[35] = mu;
  return 0;
}

/* structural B: src/basis/bmatrix.hpp:62-69 (2D), 124-140 (3D) */
static void make_B(int ne, int d, const double *dNdX, double *B /* voigt x ne*d */)
{
  int nd = ne * d;
  if (d == 2) {
    memset(B, 0, 3 * nd * sizeof(double));
    for (int n = 0; n < ne; n++) {
      B[0 * nd + 2 * n] = dNdX[0 * ne + n];
      B[1 * nd + 2 * n + 1] = dNdX[1 * ne + n];
      B[2 * nd + 2 * n] = dNdX[1 * ne + n];
      B[2 * nd + 2 * n + 1] = dNdX[0 * ne + n];
    }
    return;
  }
  memset(B, 0, 6 * nd * sizeof(double));
  for (int n = 0; n < ne; n++) {
    double dx = dNdX[0 * ne + n], dy = dNdX[1 * ne + n], dz = dNdX[2 * ne + n];
    B[0 * nd + 3 * n] = dx;
    B[1 * nd + 3 * n + 1] = dy;
    B[2 * nd + 3 * n + 2] = dz;
    B[3 * nd + 3 * n] = dy; B[3 * nd + 3 * n + 1] = dx;
    B[4 * nd + 3 * n + 1] = dz; B[4 * nd + 3 * n + 2] = dy;
    B[5 * nd + 3 * n] = dz; B[5 * nd + 3 * n + 2] = dx;
  }
}

/* Ke = sum_ip B^T De B jac, volume = sum_ip jac:
 * src/problem/elasticity.cpp:210-254 (makeKeFint) */
void icxo_element_stiffness(int ne, int ndim, int state, double E, double nu,
                            const double *X, double *Ke, double *volume)
{
  double D[36], N[MAXNE], dNdX[3 * MAXNE], B[6 * MAXDOF], T[MAXDOF * 6];
  int voigt, nd = ne * ndim;
  make_De(state, E, nu, D, &voigt);
  memset(Ke, 0, nd * nd * sizeof(double));
  double vol = 0.0;
  for (int ip = 0; ip < ne; ip++) {
    double jac = kinematics(ne, ndim, ip, X, N, dNdX);
    make_B(ne, ndim, dNdX, B);
    for (int i = 0; i < nd; i++)
      for (int v = 0; v < voigt; v++) {
        double s = 0.0;
        for (int w = 0; w < voigt; w++) s += B[w * nd + i] * D[w * voigt + v];
        T[i * voigt + v] = s;
      }
    for (int i = 0; i < nd; i++)
      for (int j = 0; j < nd; j++) {
        double s = 0.0;
        for (int v = 0; v < voigt; v++) s += T[i * voigt + v] * B[v * nd + j];
        Ke[i * nd + j] += s * jac;
      }
    vol += jac;
  }
  *volume = vol;
}

/* Heat: conductivity tensor from the 6 entries (xx, yy, zz, xy, yz, xz) of a
 * [material] conductivity row, Heat::setParameter (src/problem/heat.cpp:19-60).  The
 * reference's 3D branch assigns material 2 to De1_ and leaves De2_ unset; the intended
 * tensor (the 2D pattern extended) is used here. */
void icxo_heat_tensor(int ndim, const double *c6, double *D)
{
  if (ndim == 2) {
    D[0] = c6[0]; D[1] = c6[3];
    D[2] = c6[3]; D[3] = c6[1];
    return;
  }
  D[0] = c6[0]; D[1] = c6[3]; D[2] = c6[5];
  D[3] = c6[3]; D[4] = c6[1]; D[5] = c6[4];
  D[6] = c6[5]; D[7] = c6[4]; D[8] = c6[2];
}

/* LinearHeat::makeKeFint (src/problem/heat.cpp:79-109): Ke = sum_ip B^T De B jac with
 * B = dN/dX (ndim x ne), volume = sum_ip jac */
void icxo_heat_element(int ne, int ndim, const double *D, const double *X, double *Ke, double *volume)
{
  double N[MAXNE], dNdX[3 * MAXNE];
  memset(Ke, 0, (size_t)ne * ne * sizeof(double));
  double vol = 0.0;
  for (int ip = 0; ip < ne; ip++) {
    double jac = kinematics(ne, ndim, ip, X, N, dNdX);
    for (int i = 0; i < ne; i++)
      for (int j = 0; j < ne; j++) {
        double t = 0.0;
        for (int a = 0; a < ndim; a++)
          for (int b = 0; b < ndim; b++) t += dNdX[a * ne + i] * D[a * ndim + b] * dNdX[b * ne + j];
        Ke[i * ne + j] += t * jac;
      }
    vol += jac;
  }
  *volume = vol;
}

/* Helmholtz element: Ke = sum (N N^T + r_h^2 dN^T dN) jac, r_h = R/(2 sqrt 3);
 * Nint = sum N jac.  src/optimization/filter_helm.hpp:178-222 */
void icxo_helm_element(int ne, int ndim, double radius, const double *X, double *Ke,
                       double *Nint)
{
  double N[MAXNE], dNdX[3 * MAXNE];
  double rh2 = pow(radius / (2 * sqrt(3)), 2);
  memset(Ke, 0, ne * ne * sizeof(double));
  memset(Nint, 0, ne * sizeof(double));
  for (int ip = 0; ip < ne; ip++) {
    double jac = kinematics(ne, ndim, ip, X, N, dNdX);
    for (int a = 0; a < ne; a++) {
      for (int b = 0; b < ne; b++) {
        double g = 0.0;
        for (int d = 0; d < ndim; d++) g += dNdX[d * ne + a] * rh2 * dNdX[d * ne + b];
        Ke[a * ne + b] += (N[a] * N[b] + g) * jac;
      }
      Nint[a] += N[a] * jac;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Neumann tractions: src/io/boundary_condition_icarat.hpp:253-396,          */
/* src/io/boundary_condition_base.cpp:16-180 (judgeLine, generate),          */
/* 226-345 (side node lists), src/basis/nmatrix_surface.hpp:23-89            */
/* ------------------------------------------------------------------------ */
static int judge_side(int ne, int counter, const int *hit)
{
  if (ne == 4) {
    if (counter == 1) return 1;
    if (counter == 2) return 2;
    return (hit[1] - hit[0] == 1) ? 3 : 4;
  }
  if (counter == 3) return 1;
  if (counter == 5) return 3;
  if (counter == 6) return 4;
  if (hit[2] - hit[1] == 1) return 2;
  if (hit[1] - hit[0] == 1) return 5;
  return 6;
}

void icxo_neumann(int ndim, int dofnp, int ne, int nelx, const double *x,
                  const int *conn, const int *dof, int numeq,
                  const double *val3, const double *range6, double *fext)
{
  const double tol = 1.0e-7;
  double lo[3] = {range6[0] - tol, range6[2] - tol, range6[4] - tol};
  double hi[3] = {range6[1] + tol, range6[3] + tol, range6[5] + tol};
  static const int side4[5][2] = {{0, 0}, {0, 1}, {1, 2}, {2, 3}, {3, 0}};
  static const int side8[7][4] = {{0, 0, 0, 0}, {0, 1, 2, 3}, {3, 0, 4, 7}, {0, 1, 5, 4},
                                  {1, 2, 6, 5}, {2, 3, 7, 6}, {4, 5, 6, 7}};
  int numnode = (ne == 4) ? 2 : 4;
  const double g = 1.0 / sqrt(3.0);
  for (int e = 0; e < nelx; e++) {
    int hits[MAXNE], cnt = 0;
    for (int k = 0; k < ne; k++) {
      const double *p = x + 3 * conn[ne * e + k];
      int in = 1;
      for (int d = 0; d < ndim; d++) in = in && lo[d] <= p[d] && p[d] <= hi[d];
      if (!in) continue;
      hits[cnt++] = k;
      if (cnt != numnode) continue;
      int side = judge_side(ne, k, hits);
      double fe[4 * 3];
      memset(fe, 0, sizeof fe);
      if (ne == 4) {
        const int *sn = side4[side];
        for (int ip = 0; ip < 2; ip++) {
          double r = (ip == 0 ? -g : g);
          double Nl[2] = {0.5 * (1.0 - r), 0.5 * (1.0 + r)};
          double d0 = 0.0, d1 = 0.0;
          for (int a = 0; a < 2; a++) {
            double dn = (a == 0) ? -0.5 : 0.5;
            d0 += dn * x[3 * conn[ne * e + sn[a]] + 0];
            d1 += dn * x[3 * conn[ne * e + sn[a]] + 1];
          }
          double jac = sqrt(pow(d0, 2.0) + pow(d1, 2.0)) * 1.0;
          for (int a = 0; a < 2; a++)
            for (int c = 0; c < dofnp && c < 3; c++) fe[a * 3 + c] += Nl[a] * val3[c] * jac;
        }
        for (int a = 0; a < 2; a++)
          for (int c = 0; c < dofnp && c < 3; c++) {
            int id = dof[dofnp * conn[ne * e + sn[a]] + c];
            if (id < numeq) fext[id] += fe[a * 3 + c];
          }
      } else {
        const int *sn = side8[side];
        for (int ip = 0; ip < 4; ip++) {
          double r[2];
          gauss_point(4, ip, r);
          double Nq[4], dq[8];
          shape(4, r, Nq, dq);
          double t1[3] = {0, 0, 0}, t2[3] = {0, 0, 0};
          for (int a = 0; a < 4; a++)
            for (int d = 0; d < 3; d++) {
              t1[d] += dq[a] * x[3 * conn[ne * e + sn[a]] + d];
              t2[d] += dq[4 + a] * x[3 * conn[ne * e + sn[a]] + d];
            }
          double cx = t1[1] * t2[2] - t1[2] * t2[1];
          double cy = t1[2] * t2[0] - t1[0] * t2[2];
          double cz = t1[0] * t2[1] - t1[1] * t2[0];
          double jac = sqrt(cx * cx + cy * cy + cz * cz) * 1.0;
          for (int a = 0; a < 4; a++)
            for (int c = 0; c < dofnp && c < 3; c++) fe[a * 3 + c] += Nq[a] * val3[c] * jac;
        }
        for (int a = 0; a < 4; a++)
          for (int c = 0; c < dofnp && c < 3; c++) {
            int id = dof[dofnp * conn[ne * e + sn[a]] + c];
            if (id < numeq) fext[id] += fe[a * 3 + c];
          }
      }
    }
  }
}

/* ------------------------------------------------------------------------ */
/* node adjacency (shared by structural and Helmholtz CSR patterns)          */
/* ------------------------------------------------------------------------ */
typedef struct {
  int *eptr, *eid;  /* node -> elements */
  int *nptr, *nid;  /* node -> neighbour nodes (sorted, includes self) */
} adjacency;

static int cmp_int(const void *a, const void *b)
{
  int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}

static void build_adjacency(int ne, int numnp, int nelx, const int *conn, adjacency *A)
{
  A->eptr = calloc(numnp + 1, sizeof(int));
  for (int e = 0; e < nelx; e++)
    for (int k = 0; k < ne; k++) A->eptr[conn[ne * e + k] + 1]++;
  for (int n = 0; n < numnp; n++) A->eptr[n + 1] += A->eptr[n];
  A->eid = malloc((size_t)A->eptr[numnp] * sizeof(int));
  int *fill = calloc(numnp, sizeof(int));
  for (int e = 0; e < nelx; e++)
    for (int k = 0; k < ne; k++) {
      int n = conn[ne * e + k];
      A->eid[A->eptr[n] + fill[n]++] = e;
    }
  free(fill);
  A->nptr = calloc(numnp + 1, sizeof(int));
  int *tmpcnt = calloc(numnp, sizeof(int));
#pragma omp parallel for schedule(static)
  for (int n = 0; n < numnp; n++) {
    int buf[27 * 8 + 8], c = 0;
    for (int q = A->eptr[n]; q < A->eptr[n + 1]; q++)
      for (int k = 0; k < ne; k++) buf[c++] = conn[ne * A->eid[q] + k];
    qsort(buf, c, sizeof(int), cmp_int);
    int u = 0;
    for (int i = 0; i < c; i++)
      if (i == 0 || buf[i] != buf[i - 1]) u++;
    tmpcnt[n] = u;
  }
  for (int n = 0; n < numnp; n++) A->nptr[n + 1] = A->nptr[n] + tmpcnt[n];
  A->nid = malloc((size_t)A->nptr[numnp] * sizeof(int));
#pragma omp parallel for schedule(static)
  for (int n = 0; n < numnp; n++) {
    int buf[27 * 8 + 8], c = 0;
    for (int q = A->eptr[n]; q < A->eptr[n + 1]; q++)
      for (int k = 0; k < ne; k++) buf[c++] = conn[ne * A->eid[q] + k];
    qsort(buf, c, sizeof(int), cmp_int);
    int u = 0;
    for (int i = 0; i < c; i++)
      if (i == 0 || buf[i] != buf[i - 1]) A->nid[A->nptr[n] + u++] = buf[i];
  }
  free(tmpcnt);
}

static void free_adjacency(adjacency *A)
{
  free(A->eptr); free(A->eid); free(A->nptr); free(A->nid);
}

/* position of column c in the sorted row [lo, hi) */
static inline int find_col(const int *col, int lo, int hi, int c)
{
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (col[mid] < c) lo = mid + 1; else hi = mid;
  }
  return lo;
}

/* Greedy element colouring (elements sharing a node get different colours), so each
 * colour can be assembled in parallel without atomics and the summation order is
 * fixed: colour by colour, element id order inside a colour. */
static int colour_elements(int ne, int numnp, int nelx, const int *conn, int **colour_ptr,
                           int **order)
{
  adjacency A;
  build_adjacency(ne, numnp, nelx, conn, &A);
  int *col = malloc((size_t)nelx * sizeof(int));
  int ncol = 0;
  for (int e = 0; e < nelx; e++) {
    unsigned long long used = 0;
    for (int k = 0; k < ne; k++) {
      int n = conn[ne * e + k];
      for (int q = A.eptr[n]; q < A.eptr[n + 1]; q++) {
        int f = A.eid[q];
        if (f < e && col[f] < 64) used |= 1ULL << col[f];
      }
    }
    int c = 0;
    while (c < 63 && (used >> c & 1ULL)) c++;
    col[e] = c;
    if (c + 1 > ncol) ncol = c + 1;
  }
  free_adjacency(&A);
  int *ptr = calloc((size_t)ncol + 1, sizeof(int));
  for (int e = 0; e < nelx; e++) ptr[col[e] + 1]++;
  for (int c = 0; c < ncol; c++) ptr[c + 1] += ptr[c];
  int *ord = malloc((size_t)nelx * sizeof(int)), *fill = calloc((size_t)ncol, sizeof(int));
  for (int e = 0; e < nelx; e++) ord[ptr[col[e]] + fill[col[e]]++] = e;
  free(fill); free(col);
  *colour_ptr = ptr; *order = ord;
  return ncol;
}

/* ------------------------------------------------------------------------ */
/* structural system: the reference assembles K by triplets per element and  */
/* sums duplicates (src/analysis/linear.hpp:117-166, src/misc/assembling.hpp:17-38,
 * 77-84).  Same values here, built straight into CSR.                       */
/* ------------------------------------------------------------------------ */
struct icxo_sys {
  int ne, ndim, dofnp, state, numnp, nelx, numeq, nd;
  const double *x; const int *conn; const int *dof;
  double young1, young2, nu, p;
  double D1[9], D2[9];   /* heat: conductivity tensors of materials 1 and 2 */
  int *rowptr, *col; double *val; double *invdiag;
  double *volume; double *design; double *young;
  int ncol; int *cptr, *cord;
};

static void element_X(const icxo_sys *s, int e, double *X)
{
  for (int k = 0; k < s->ne; k++)
    for (int d = 0; d < s->ndim; d++) X[k * s->ndim + d] = s->x[3 * s->conn[s->ne * e + k] + d];
}

/* Ke of element e for the design of the last assembly: elastic E_e K0, or heat with
 * De = design^p De1 + (1 - design^p) De2 (Heat::makeSIMPDe, src/problem/heat.cpp:62-77) */
static void element_K(const icxo_sys *s, int e, const double *X, double *Ke, double *vol)
{
  if (s->state == ICXO_HEAT) {
    const double w = pow(s->design[e], s->p);
    double D[9];
    for (int i = 0; i < s->ndim * s->ndim; i++) D[i] = w * s->D1[i] + (1.0 - w) * s->D2[i];
    icxo_heat_element(s->ne, s->ndim, D, X, Ke, vol);
    return;
  }
  icxo_element_stiffness(s->ne, s->ndim, s->state, s->young[e], s->nu, X, Ke, vol);
}

void icxo_sys_set_conductivity(icxo_sys *s, const double *k1, const double *k2)
{
  icxo_heat_tensor(s->ndim, k1, s->D1);
  icxo_heat_tensor(s->ndim, k2, s->D2);
}

icxo_sys *icxo_sys_create(int ne, int ndim, int dofnp, int state, int numnp, int nelx,
                          const double *x, const int *conn, const int *dof, int numeq,
                          double young1, double young2, double nu, double p)
{
  icxo_sys *s = calloc(1, sizeof(*s));
  s->ne = ne; s->ndim = ndim; s->dofnp = dofnp; s->state = state;
  s->numnp = numnp; s->nelx = nelx; s->numeq = numeq; s->nd = ne * dofnp;
  s->x = x; s->conn = conn; s->dof = dof;
  s->young1 = young1; s->young2 = young2; s->nu = nu; s->p = p;
  adjacency A;
  build_adjacency(ne, numnp, nelx, conn, &A);
  s->rowptr = calloc((size_t)numeq + 1, sizeof(int));
  /* row lengths: free DOFs of neighbour nodes */
  for (int n = 0; n < numnp; n++) {
    int cnt = 0;
    for (int q = A.nptr[n]; q < A.nptr[n + 1]; q++)
      for (int c = 0; c < dofnp; c++) cnt += dof[dofnp * A.nid[q] + c] < numeq;
    for (int c = 0; c < dofnp; c++) {
      int r = dof[dofnp * n + c];
      if (r < numeq) s->rowptr[r + 1] = cnt;
    }
  }
  for (int r = 0; r < numeq; r++) s->rowptr[r + 1] += s->rowptr[r];
  s->col = malloc((size_t)s->rowptr[numeq] * sizeof(int));
  s->val = calloc((size_t)s->rowptr[numeq], sizeof(double));
#pragma omp parallel for schedule(static)
  for (int n = 0; n < numnp; n++)
    for (int c = 0; c < dofnp; c++) {
      int r = dof[dofnp * n + c];
      if (r >= numeq) continue;
      int w = s->rowptr[r];
      /* free DOFs are numbered in node order, so this walk is sorted */
      for (int q = A.nptr[n]; q < A.nptr[n + 1]; q++)
        for (int b = 0; b < dofnp; b++) {
          int cc = dof[dofnp * A.nid[q] + b];
          if (cc < numeq) s->col[w++] = cc;
        }
    }
  free_adjacency(&A);
  s->ncol = colour_elements(ne, numnp, nelx, conn, &s->cptr, &s->cord);
  s->invdiag = malloc((size_t)numeq * sizeof(double));
  s->volume = malloc((size_t)nelx * sizeof(double));
  s->design = malloc((size_t)nelx * sizeof(double));
  s->young = malloc((size_t)nelx * sizeof(double));
#pragma omp parallel for schedule(static)
  for (int e = 0; e < nelx; e++) {
    double X[MAXNE * 3], Ke[MAXDOF * MAXDOF];
    element_X(s, e, X);
    if (state == ICXO_HEAT) {
      const double I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
      double D[9];
      for (int i = 0; i < ndim * ndim; i++) D[i] = I[(i / ndim) * 3 + i % ndim];
      icxo_heat_element(ne, ndim, D, X, Ke, &s->volume[e]);
    } else {
      icxo_element_stiffness(ne, ndim, state, 1.0, nu, X, Ke, &s->volume[e]);
    }
  }
  return s;
}

void icxo_sys_destroy(icxo_sys *s)
{
  if (!s) return;
  free(s->rowptr); free(s->col); free(s->val); free(s->invdiag);
  free(s->volume); free(s->design); free(s->young); free(s->cptr); free(s->cord); free(s);
}

long long icxo_sys_nnz(const icxo_sys *s) { return s->rowptr[s->numeq]; }

/* Assembles K(design) over free DOFs and fint = K_full * u_dirichlet on the
 * free rows (the reference's fint with node.val = Dirichlet values,
 * src/analysis/linear.hpp:62-80).  node_val/node_flag/has may be NULL for
 * homogeneous constraints. */
void icxo_sys_assemble(icxo_sys *s, const double *design, const int *has,
                       const int *node_flag, const double *node_val, double *fint)
{
  const int ne = s->ne, nd = s->nd, dofnp = s->dofnp, numeq = s->numeq;
  memset(s->val, 0, (size_t)s->rowptr[numeq] * sizeof(double));
  if (fint) memset(fint, 0, (size_t)numeq * sizeof(double));
  for (int colour = 0; colour < s->ncol; colour++)
#pragma omp parallel for schedule(static)
  for (int t = s->cptr[colour]; t < s->cptr[colour + 1]; t++) {
    const int e = s->cord[t];
    double X[MAXNE * 3], Ke[MAXDOF * MAXDOF], vol, disp[MAXDOF];
    int idof[MAXDOF];
    element_X(s, e, X);
    s->design[e] = design[e];
    s->young[e] = s->state == ICXO_HEAT ? pow(design[e], s->p) : icxo_simp(s->young1, s->young2, design[e], s->p);
    element_K(s, e, X, Ke, &vol);
    int anyd = 0;
    for (int k = 0; k < ne; k++)
      for (int c = 0; c < dofnp; c++) {
        int n = s->conn[ne * e + k];
        idof[k * dofnp + c] = s->dof[dofnp * n + c];
        double v = 0.0;
        if (has && has[n] && node_flag[6 * n + c] == 1) v = node_val[6 * n + c];
        disp[k * dofnp + c] = v;
        anyd |= (v != 0.0);
      }
    for (int i = 0; i < nd; i++) {
      int r = idof[i];
      if (r >= numeq) continue;
      int lo = s->rowptr[r], hi = s->rowptr[r + 1];
      double fe = 0.0;
      for (int j = 0; j < nd; j++) {
        fe += Ke[i * nd + j] * disp[j];
        int c = idof[j];
        if (c >= numeq) continue;
        s->val[find_col(s->col, lo, hi, c)] += Ke[i * nd + j];
      }
      if (fint && anyd) fint[r] += fe;
    }
  }
  /* DiagonalPreconditioner: 1/diag, or 1 where the diagonal is zero */
#pragma omp parallel for schedule(static)
  for (int r = 0; r < numeq; r++) {
    int q = find_col(s->col, s->rowptr[r], s->rowptr[r + 1], r);
    double d = s->val[q];
    s->invdiag[r] = (d != 0.0) ? 1.0 / d : 1.0;
  }
}

void icxo_sys_csr(const icxo_sys *s, int *rowptr, int *col, double *val)
{
  memcpy(rowptr, s->rowptr, ((size_t)s->numeq + 1) * sizeof(int));
  memcpy(col, s->col, (size_t)s->rowptr[s->numeq] * sizeof(int));
  memcpy(val, s->val, (size_t)s->rowptr[s->numeq] * sizeof(double));
}

static void csr_spmv(int n, const int *rowptr, const int *col, const double *val,
                     const double *x, double *y)
{
#pragma omp parallel for schedule(static)
  for (int r = 0; r < n; r++) {
    double acc = 0.0;
    for (int q = rowptr[r]; q < rowptr[r + 1]; q++) acc += val[q] * x[col[q]];
    y[r] = acc;
  }
}

void icxo_sys_spmv(const icxo_sys *s, const double *x, double *y)
{
  csr_spmv(s->numeq, s->rowptr, s->col, s->val, x, y);
}

/* deterministic dot: fixed chunks summed in order, independent of thread timing */
static double dot(int n, const double *a, const double *b)
{
  enum { NCH = 256 };
  double part[NCH];
  int chunk = (n + NCH - 1) / NCH;
#pragma omp parallel for schedule(static)
  for (int c = 0; c < NCH; c++) {
    double s = 0.0;
    int lo = c * chunk, hi = lo + chunk < n ? lo + chunk : n;
    for (int i = lo; i < hi; i++) s += a[i] * b[i];
    part[c] = s;
  }
  double s = 0.0;
  for (int c = 0; c < NCH; c++) s += part[c];
  return s;
}

/* Eigen 3.4 ConjugateGradient<..., Lower|Upper, DiagonalPreconditioner> as
 * called by src/analysis/solver/solvers.cpp:64-92 (tol 1e-10, maxit 10000
 * by default, solvers.hpp:90-91): x0 = 0; threshold = max(tol^2 |b|^2,
 * DBL_MIN); stop on the recursive residual before preconditioning; the
 * returned iteration count excludes the converging pass; err = |r|/|b|. */
static int pcg_eigen(int n, const int *rowptr, const int *col, const double *val,
                     const double *invdiag, const double *b, double *x, double tol,
                     int maxit, double *err)
{
  memset(x, 0, (size_t)n * sizeof(double));
  double rhsNorm2 = dot(n, b, b);
  if (rhsNorm2 == 0.0) { *err = 0.0; return 0; }
  double *r = malloc((size_t)n * sizeof(double)), *p = malloc((size_t)n * sizeof(double));
  double *z = malloc((size_t)n * sizeof(double)), *t = malloc((size_t)n * sizeof(double));
  memcpy(r, b, (size_t)n * sizeof(double));
  const double considerAsZero = 2.2250738585072014e-308;
  double threshold = tol * tol * rhsNorm2;
  if (threshold < considerAsZero) threshold = considerAsZero;
  double rn2 = dot(n, r, r);
  int i = 0;
  if (rn2 < threshold) { *err = sqrt(rn2 / rhsNorm2); goto done; }
#pragma omp parallel for schedule(static)
  for (int k = 0; k < n; k++) p[k] = invdiag[k] * r[k];
  double absNew = dot(n, r, p);
  while (i < maxit) {
    csr_spmv(n, rowptr, col, val, p, t);
    double alpha = absNew / dot(n, p, t);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n; k++) { x[k] += alpha * p[k]; r[k] -= alpha * t[k]; }
    rn2 = dot(n, r, r);
    if (rn2 < threshold) break;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n; k++) z[k] = invdiag[k] * r[k];
    double absOld = absNew;
    absNew = dot(n, r, z);
    double beta = absNew / absOld;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n; k++) p[k] = z[k] + beta * p[k];
    i++;
  }
  *err = sqrt(rn2 / rhsNorm2);
done:
  free(r); free(p); free(z); free(t);
  return i;
}

int icxo_csr_cg(int n, const int *rowptr, const int *col, const double *val,
                const double *b, double *x, double tol, int maxit, double *err)
{
  double *invdiag = malloc((size_t)n * sizeof(double));
  for (int r = 0; r < n; r++) {
    double d = 0.0;
    for (int q = rowptr[r]; q < rowptr[r + 1]; q++)
      if (col[q] == r) d += val[q];
    invdiag[r] = (d != 0.0) ? 1.0 / d : 1.0;
  }
  int it = pcg_eigen(n, rowptr, col, val, invdiag, b, x, tol, maxit, err);
  free(invdiag);
  return it;
}

int icxo_sys_cg(const icxo_sys *s, const double *b, double *x, double tol, int maxit,
                double *err)
{
  return pcg_eigen(s->numeq, s->rowptr, s->col, s->val, s->invdiag, b, x, tol, maxit, err);
}

/* Direct solve (stands in for Eigen::SimplicialLDLT, solvers.cpp:178-200):
 * banded Cholesky of the SPD system in its natural ordering. */
static int banded_cholesky_solve(int n, const int *rowptr, const int *col,
                                 const double *val, const double *b, double *x)
{
  if (n <= 0) return -1;
  int bw = 0;
  for (int r = 0; r < n; r++)
    for (int q = rowptr[r]; q < rowptr[r + 1]; q++)
      if (r - col[q] > bw) bw = r - col[q];
  size_t W = (size_t)bw + 1;
  double *L = calloc((size_t)n * W, sizeof(double)); /* L[i*W + (i-j)] for j<=i */
  if (!L) return -1;
  for (int r = 0; r < n; r++)
    for (int q = rowptr[r]; q < rowptr[r + 1]; q++)
      if (col[q] <= r) L[(size_t)r * W + (r - col[q])] += val[q];
  for (int j = 0; j < n; j++) {
    double d = L[(size_t)j * W];
    int kmin = j - bw < 0 ? 0 : j - bw;
    for (int k = kmin; k < j; k++) { double l = L[(size_t)j * W + (j - k)]; d -= l * l; }
    if (d <= 0.0) { free(L); return -2; }
    d = sqrt(d);
    L[(size_t)j * W] = d;
    int imax = j + bw < n - 1 ? j + bw : n - 1;
#pragma omp parallel for schedule(static) if (bw > 256)
    for (int i = j + 1; i <= imax; i++) {
      double s = L[(size_t)i * W + (i - j)];
      int k0 = i - bw > kmin ? i - bw : kmin;
      for (int k = k0; k < j; k++) s -= L[(size_t)i * W + (i - k)] * L[(size_t)j * W + (j - k)];
      L[(size_t)i * W + (i - j)] = s / d;
    }
  }
  for (int i = 0; i < n; i++) {
    double s = b[i];
    int k0 = i - bw < 0 ? 0 : i - bw;
    for (int k = k0; k < i; k++) s -= L[(size_t)i * W + (i - k)] * x[k];
    x[i] = s / L[(size_t)i * W];
  }
  for (int i = n - 1; i >= 0; i--) {
    double s = x[i];
    int k1 = i + bw > n - 1 ? n - 1 : i + bw;
    for (int k = i + 1; k <= k1; k++) s -= L[(size_t)k * W + (k - i)] * x[k];
    x[i] = s / L[(size_t)i * W];
  }
  free(L);
  return 0;
}

int icxo_sys_direct(const icxo_sys *s, const double *b, double *x)
{
  return banded_cholesky_solve(s->numeq, s->rowptr, s->col, s->val, b, x);
}

/* divideValIntoNode with the |v| < 1e-15 flush: src/analysis/analysis_utility.hpp:8-31 */
void icxo_divide_val_into_node(int numnp, int dofnp, int numeq, const int *dof,
                               const double *u, const double *node_val_dirich,
                               double *nodeval)
{
  int dn = dofnp < 3 ? dofnp : 3;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < numnp; i++) {
    for (int j = 0; j < 3; j++) nodeval[3 * i + j] = 0.0;
    for (int j = 0; j < dn; j++) {
      int d = dof[dofnp * i + j];
      double v = d >= numeq ? (node_val_dirich ? node_val_dirich[6 * i + j] : 0.0) : u[d];
      if (fabs(v) < 1.0e-15) v = 0.0;
      nodeval[3 * i + j] = v;
    }
  }
}

/* u_e^T Ke u_e with Ke = E_e K0_e from the last assembly (the reference keeps
 * Ke_ from LinearAnalysis, opt_macro_oc.cpp:199-206). */
void icxo_sys_element_energy(const icxo_sys *s, const double *nodeval, double *out)
{
  const int ne = s->ne, nd = s->nd, dofnp = s->dofnp;
#pragma omp parallel for schedule(static)
  for (int e = 0; e < s->nelx; e++) {
    double X[MAXNE * 3], Ke[MAXDOF * MAXDOF], vol, d[MAXDOF];
    element_X(s, e, X);
    element_K(s, e, X, Ke, &vol);
    for (int k = 0; k < ne; k++)
      for (int c = 0; c < dofnp; c++) d[k * dofnp + c] = nodeval[3 * s->conn[ne * e + k] + c];
    double acc = 0.0;
    for (int i = 0; i < nd; i++) {
      double t = 0.0;
      for (int j = 0; j < nd; j++) t += Ke[i * nd + j] * d[j];
      acc += d[i] * t;
    }
    out[e] = acc;
  }
}

/* makeMisesStress: src/problem/elasticity.cpp:165-208.  Element-mean strain and
 * stress over the Gauss points (fac = 1/ipmax), then mises = sqrt(s^T V s) with
 * V = [1 -1/2 0; -1/2 1 0; 0 0 3] (2D: xx, yy, xy; plane strain keeps the reference's
 * 2D form without s_zz) or its 6x6 counterpart (3D: xx, yy, zz, xy, yz, zx).  De is
 * that of the last assembly (young_ from SIMP in makeKeFint); nodeval is node.val
 * (3 per node, after the |v| < 1e-15 flush). */
void icxo_sys_mises(const icxo_sys *s, const double *nodeval, double *mises)
{
  const int ne = s->ne, ndim = s->ndim, nd = s->nd;
#pragma omp parallel for schedule(static)
  for (int e = 0; e < s->nelx; e++) {
    double X[MAXNE * 3], N[MAXNE], dNdX[3 * MAXNE], B[6 * MAXDOF], D[36], d[MAXDOF];
    double stress[6] = {0, 0, 0, 0, 0, 0};
    int voigt;
    element_X(s, e, X);
    make_De(s->state, s->young[e], s->nu, D, &voigt);
    for (int k = 0; k < ne; k++)
      for (int c = 0; c < ndim; c++) d[k * ndim + c] = nodeval[3 * s->conn[ne * e + k] + c];
    const double fac = 1.0 / ne; /* ipmax = ne for quad4 / hexa8 */
    for (int ip = 0; ip < ne; ip++) {
      kinematics(ne, ndim, ip, X, N, dNdX);
      make_B(ne, ndim, dNdX, B);
      double strain[6];
      for (int v = 0; v < voigt; v++) {
        double t = 0.0;
        for (int j = 0; j < nd; j++) t += B[v * nd + j] * d[j];
        strain[v] = t;
      }
      for (int v = 0; v < voigt; v++) {
        double t = 0.0;
        for (int w = 0; w < voigt; w++) t += D[v * voigt + w] * strain[w];
        stress[v] += t * fac;
      }
    }
    double q;
    if (voigt == 3) {
      q = stress[0] * stress[0] - stress[0] * stress[1] + stress[1] * stress[1] + 3.0 * stress[2] * stress[2];
    } else {
      q = stress[0] * stress[0] + stress[1] * stress[1] + stress[2] * stress[2] - stress[0] * stress[1] -
          stress[1] * stress[2] - stress[0] * stress[2] +
          3.0 * (stress[3] * stress[3] + stress[4] * stress[4] + stress[5] * stress[5]);
    }
    mises[e] = sqrt(q);
  }
}

/* Stress constraint, 2D elastic jobs (constraint/opt-stress-constraint/sensitivity.cpp:8-155)
 * on the last assembly (design = rho_hat, young_e) and its solution u (nodeval: node.val
 * after the flush):
 *   s_e = rho_e^qp E2 vm_e / E_e (qp relaxation), g = (sum (s_e/slim)^P)^(1/P) - 1
 *   factor1_e = psum^(1/P-1) (s_e/slim)^(P-1) dvm/dsigma / slim, factor2_e = qp rho^(qp-1) E2 sigma_e / E_e
 *   K lambda = sum_e Fe, Fe = sum_ip (rho^qp E2 / E_e) factor1^T De Be / ipmax  (free DOFs)
 *   dg_e = factor1.factor2 - (p rho^(p-1) (E2 - E1) / E_e) (Ke u_e).lambda_e
 * lambda_e takes the Dirichlet value at constrained DOFs (numeqToElement, src/misc/
 * assembling.hpp:271-298).  The adjoint solve is direct or CG (cg_tol).  Returns -1 for
 * 3D jobs or E2 <= E1, as the reference exits there. */
int icxo_stress_constraint(const icxo_sys *s, const double *nodeval, const double *node_val_dirich,
                           double qp, double pnorm, double slim, int direct, double cg_tol,
                           double *g, double *dgds, double *sqp_out)
{
  const int ne = s->ne, ndim = s->ndim, nd = s->nd, nelx = s->nelx, numeq = s->numeq;
  if (ndim != 2 || !(s->young2 > s->young1)) return -1;
  double *sig = malloc((size_t)nelx * 3 * 8), *vm = malloc((size_t)nelx * 8), *sqp = malloc((size_t)nelx * 8);
  double psum = 0.0;
  for (int e = 0; e < nelx; e++) {
    double X[MAXNE * 3], N[MAXNE], dNdX[3 * MAXNE], B[6 * MAXDOF], D[36], d[MAXDOF], st[3] = {0, 0, 0};
    int voigt;
    element_X(s, e, X);
    make_De(s->state, s->young[e], s->nu, D, &voigt);
    for (int k = 0; k < ne; k++)
      for (int c = 0; c < 2; c++) d[k * 2 + c] = nodeval[3 * s->conn[ne * e + k] + c];
    for (int ip = 0; ip < ne; ip++) {
      kinematics(ne, ndim, ip, X, N, dNdX);
      make_B(ne, ndim, dNdX, B);
      double eps[3];
      for (int v = 0; v < 3; v++) { double t = 0.0; for (int j = 0; j < nd; j++) t += B[v * nd + j] * d[j]; eps[v] = t; }
      for (int v = 0; v < 3; v++) { double t = 0.0; for (int w = 0; w < 3; w++) t += D[v * 3 + w] * eps[w]; st[v] += t / ne; }
    }
    for (int v = 0; v < 3; v++) sig[3 * e + v] = st[v];
    vm[e] = sqrt(st[0] * st[0] - st[0] * st[1] + st[1] * st[1] + 3.0 * st[2] * st[2]);
    sqp[e] = pow(s->design[e], qp) * s->young2 * vm[e] / s->young[e];
    psum += pow(sqp[e] / slim, pnorm);
  }
  *g = pow(psum, 1.0 / pnorm) - 1.0;
  if (sqp_out) memcpy(sqp_out, sqp, (size_t)nelx * 8);
  const double dpndvm = pow(psum, (1.0 / pnorm) - 1.0);
  double *f1 = malloc((size_t)nelx * 3 * 8), *f12 = malloc((size_t)nelx * 8);
  double *load = calloc((size_t)numeq, 8), *lam = calloc((size_t)numeq, 8);
  for (int e = 0; e < nelx; e++) {
    const double *st = sig + 3 * e;
    const double dvm[3] = {(2.0 * st[0] - st[1]) / (2.0 * vm[e]), (2.0 * st[1] - st[0]) / (2.0 * vm[e]),
                           3.0 * st[2] / vm[e]};
    const double dvds = qp * pow(s->design[e], qp - 1.0) * s->young2;
    double acc = 0.0;
    for (int v = 0; v < 3; v++) {
      f1[3 * e + v] = dpndvm * pow(sqp[e] / slim, pnorm - 1) * dvm[v] / slim;
      acc += f1[3 * e + v] * (dvds * st[v] / s->young[e]);
    }
    f12[e] = acc;
    double X[MAXNE * 3], N[MAXNE], dNdX[3 * MAXNE], B[6 * MAXDOF], D[36], Fe[MAXDOF] = {0};
    int voigt;
    element_X(s, e, X);
    make_De(s->state, s->young[e], s->nu, D, &voigt);
    const double coef = pow(s->design[e], qp) * s->young2 / s->young[e];
    for (int ip = 0; ip < ne; ip++) {
      kinematics(ne, ndim, ip, X, N, dNdX);
      make_B(ne, ndim, dNdX, B);
      double t[3];
      for (int w = 0; w < 3; w++) { t[w] = 0.0; for (int v = 0; v < 3; v++) t[w] += f1[3 * e + v] * D[v * 3 + w]; }
      for (int j = 0; j < nd; j++) {
        double a = 0.0;
        for (int w = 0; w < 3; w++) a += t[w] * B[w * nd + j];
        Fe[j] += coef * a / ne;
      }
    }
    for (int k = 0; k < ne; k++)
      for (int c = 0; c < 2; c++) {
        const int dof = s->dof[2 * s->conn[ne * e + k] + c];
        if (dof < numeq) load[dof] += Fe[2 * k + c];
      }
  }
  int rc = 0;
  if (direct) rc = icxo_sys_direct(s, load, lam);
  else { double err; icxo_sys_cg(s, load, lam, cg_tol, 100000, &err); }
  if (rc == 0) {
    for (int e = 0; e < nelx; e++) {
      double X[MAXNE * 3], Ke[MAXDOF * MAXDOF], vol, d[MAXDOF], l[MAXDOF];
      element_X(s, e, X);
      element_K(s, e, X, Ke, &vol);
      for (int k = 0; k < ne; k++)
        for (int c = 0; c < 2; c++) {
          const int n = s->conn[ne * e + k], dof = s->dof[2 * n + c];
          d[2 * k + c] = nodeval[3 * n + c];
          l[2 * k + c] = dof < numeq ? lam[dof] : (node_val_dirich ? node_val_dirich[6 * n + c] : 0.0);
        }
      double f34 = 0.0;
      for (int i = 0; i < nd; i++) {
        double t = 0.0;
        for (int j = 0; j < nd; j++) t += Ke[i * nd + j] * d[j];
        f34 += t * l[i];
      }
      const double dvpds = s->p * pow(s->design[e], s->p - 1.0) * (s->young2 - s->young1);
      dgds[e] = f12[e] - (dvpds / s->young[e]) * f34;
    }
  }
  free(sig); free(vm); free(sqp); free(f1); free(f12); free(load); free(lam);
  return rc;
}

/* Reactions: LinearAnalysis::solve re-assembles after the solve
 * (src/analysis/linear.hpp:90-91, 117-166): fe = Ke u_e of every element summed
 * over all DOFs (assemblingReact, src/misc/assembling.hpp:77-84), then
 * divideReactionIntoNode (src/analysis/analysis_utility.hpp:33-51) with the
 * |r| < 1e-15 flush.  out: 3 per node (components beyond dofnp are 0). */
void icxo_sys_reactions(const icxo_sys *s, const double *nodeval, double *out)
{
  const int ne = s->ne, nd = s->nd, dofnp = s->dofnp;
  memset(out, 0, (size_t)s->numnp * 3 * sizeof(double));
  /* element colours: elements of one colour share no node */
  for (int col = 0; col < s->ncol; col++) {
#pragma omp parallel for schedule(static)
    for (int q = s->cptr[col]; q < s->cptr[col + 1]; q++) {
      const int e = s->cord[q];
      double X[MAXNE * 3], Ke[MAXDOF * MAXDOF], vol, d[MAXDOF];
      element_X(s, e, X);
      element_K(s, e, X, Ke, &vol);
      for (int k = 0; k < ne; k++)
        for (int c = 0; c < dofnp; c++) d[k * dofnp + c] = nodeval[3 * s->conn[ne * e + k] + c];
      for (int i = 0; i < nd; i++) {
        double t = 0.0;
        for (int j = 0; j < nd; j++) t += Ke[i * nd + j] * d[j];
        out[3 * s->conn[ne * e + i / dofnp] + i % dofnp] += t;
      }
    }
  }
  for (long long i = 0; i < (long long)s->numnp * 3; i++)
    if (fabs(out[i]) < 1.0e-15) out[i] = 0.0;
}

/* Heat sensitivity: opt-heat/opt_heat.cpp:158-194 without the projection factor:
 * dfds = -T_e^T (p s^(p-1) sum_ip B^T (De1 - De2) B jac) T_e, s = the element design of
 * the last assembly; nodeval is node.val (3 per node, component 0 used). */
void icxo_sys_heat_sensitivity(const icxo_sys *s, const double *nodeval, double *dfds)
{
  const int ne = s->ne;
#pragma omp parallel for schedule(static)
  for (int e = 0; e < s->nelx; e++) {
    double X[MAXNE * 3], Ke[MAXNE * MAXNE], vol, t[MAXNE], D[9];
    element_X(s, e, X);
    const double f = s->p * pow(s->design[e], s->p - 1.0);
    for (int i = 0; i < s->ndim * s->ndim; i++) D[i] = f * (s->D1[i] - s->D2[i]);
    icxo_heat_element(ne, s->ndim, D, X, Ke, &vol);
    for (int k = 0; k < ne; k++) t[k] = nodeval[3 * s->conn[ne * e + k]];
    double acc = 0.0;
    for (int i = 0; i < ne; i++) {
      double r = 0.0;
      for (int j = 0; j < ne; j++) r += Ke[i * ne + j] * t[j];
      acc += t[i] * r;
    }
    dfds[e] = -acc;
  }
}

/* Sensitivity: opt_macro_oc.cpp:182-212 (driver A, dgds = v) and
 * opt_mma_Helmholtz.cpp:173-205 (driver B, dgds = v/V0):
 * dfds = -(dE/E) * u^T Ke u.  `energy` is u^T Ke u with Ke = E K0. */
void icxo_sensitivity(int nelx, int driver_b, double y1, double y2, double p,
                      const double *design, const double *energy, const double *volume,
                      double V0, double *dfds, double *dgds)
{
#pragma omp parallel for schedule(static)
  for (int e = 0; e < nelx; e++) {
    double dens = design[e], young, dyoung;
    if (y2 >= y1) {
      young = (1.0 - pow(dens, p)) * y1 + pow(dens, p) * y2;
      dyoung = p * pow(dens, p - 1.0) * (y2 - y1);
    } else {
      young = pow(1.0 - dens, p) * y1 + (1.0 - pow(1.0 - dens, p)) * y2;
      dyoung = (y2 - y1) * p * pow(1.0 - dens, p - 1.0);
    }
    double E = dyoung / young;
    dfds[e] = -E * energy[e];
    dgds[e] = driver_b ? volume[e] / V0 : volume[e];
  }
}

/* ------------------------------------------------------------------------ */
/* cone density filter: src/optimization/filter.hpp:102-151, 191-234          */
/* ------------------------------------------------------------------------ */
struct icxo_filter {
  int nelx;
  double radius;
  int *ptr, *id; double *dist;
  double *aa; /* dfilters_ds: sum of weights cached by the last forward call */
};

static void element_centers(int ne, int ndim, int nelx, const double *x, const int *conn,
                            double *c)
{
#pragma omp parallel for schedule(static)
  for (int e = 0; e < nelx; e++)
    for (int d = 0; d < 3; d++) {
      double s = 0.0;
      if (d < ndim)
        for (int k = 0; k < ne; k++) s += x[3 * conn[ne * e + k] + d] / ne;
      c[3 * e + d] = s;
    }
}

static inline double center_dist(int ndim, const double *a, const double *b)
{
  double s = 0.0;
  for (int d = 0; d < ndim; d++) { double t = a[d] - b[d]; s += t * t; }
  return sqrt(s);
}

icxo_filter *icxo_filter_create(int ne, int ndim, int nelx, const double *x,
                                const int *conn, double radius, int brute)
{
  icxo_filter *f = calloc(1, sizeof(*f));
  f->nelx = nelx; f->radius = radius;
  double *c = malloc((size_t)nelx * 3 * sizeof(double));
  element_centers(ne, ndim, nelx, x, conn, c);
  int *cnt = calloc((size_t)nelx + 1, sizeof(int));
  /* cell grid of size radius for the candidate search; the test applied is
   * the reference's (distance <= radius), candidates are visited in
   * increasing element id so the neighbour order matches the O(n^2) scan */
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int e = 0; e < nelx; e++)
    for (int d = 0; d < 3; d++) {
      if (c[3 * e + d] < lo[d]) lo[d] = c[3 * e + d];
      if (c[3 * e + d] > hi[d]) hi[d] = c[3 * e + d];
    }
  double cell = radius > 0 ? radius : 1.0;
  int nc[3];
  for (int d = 0; d < 3; d++) { nc[d] = (int)((hi[d] - lo[d]) / cell) + 1; if (nc[d] < 1) nc[d] = 1; }
  long long ncell = (long long)nc[0] * nc[1] * nc[2];
  int *cptr = NULL, *cid = NULL, *ecell = NULL;
  if (!brute) {
    cptr = calloc((size_t)ncell + 1, sizeof(int));
    ecell = malloc((size_t)nelx * sizeof(int));
    for (int e = 0; e < nelx; e++) {
      int ix = (int)((c[3 * e] - lo[0]) / cell), iy = (int)((c[3 * e + 1] - lo[1]) / cell),
          iz = (int)((c[3 * e + 2] - lo[2]) / cell);
      if (ix >= nc[0]) ix = nc[0] - 1;
      if (iy >= nc[1]) iy = nc[1] - 1;
      if (iz >= nc[2]) iz = nc[2] - 1;
      ecell[e] = ix + nc[0] * (iy + nc[1] * iz);
      cptr[ecell[e] + 1]++;
    }
    for (long long q = 0; q < ncell; q++) cptr[q + 1] += cptr[q];
    cid = malloc((size_t)nelx * sizeof(int));
    int *fill = calloc((size_t)ncell, sizeof(int));
    for (int e = 0; e < nelx; e++) cid[cptr[ecell[e]] + fill[ecell[e]]++] = e;
    free(fill);
  }
  for (int pass = 0; pass < 2; pass++) {
    if (pass == 1) {
      for (int e = 0; e < nelx; e++) cnt[e + 1] += cnt[e];
      f->ptr = cnt;
      f->id = malloc((size_t)cnt[nelx] * sizeof(int));
      f->dist = malloc((size_t)cnt[nelx] * sizeof(double));
    }
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < nelx; i++) {
      int w = pass ? f->ptr[i] : 0;
      if (brute) {
        for (int j = 0; j < nelx; j++) {
          double d = center_dist(ndim, c + 3 * i, c + 3 * j);
          if (d <= radius) {
            if (pass) { f->id[w] = j; f->dist[w] = d; }
            w++;
          }
        }
      } else {
        int cand[4096], m = 0;
        int ci = ecell[i];
        int ix = ci % nc[0], iy = (ci / nc[0]) % nc[1], iz = ci / (nc[0] * nc[1]);
        for (int dz = -1; dz <= 1; dz++)
          for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
              int X = ix + dx, Y = iy + dy, Z = iz + dz;
              if (X < 0 || Y < 0 || Z < 0 || X >= nc[0] || Y >= nc[1] || Z >= nc[2]) continue;
              int q = X + nc[0] * (Y + nc[1] * Z);
              for (int t = cptr[q]; t < cptr[q + 1] && m < 4096; t++) cand[m++] = cid[t];
            }
        qsort(cand, m, sizeof(int), cmp_int);
        for (int t = 0; t < m; t++) {
          int j = cand[t];
          double d = center_dist(ndim, c + 3 * i, c + 3 * j);
          if (d <= radius) {
            if (pass) { f->id[w] = j; f->dist[w] = d; }
            w++;
          }
        }
      }
      if (!pass) cnt[i + 1] = w;
    }
  }
  free(cptr); free(cid); free(ecell); free(c);
  f->aa = calloc((size_t)nelx, sizeof(double));
  return f;
}

void icxo_filter_destroy(icxo_filter *f)
{
  if (!f) return;
  free(f->ptr); free(f->id); free(f->dist); free(f->aa); free(f);
}

long long icxo_filter_nnz(const icxo_filter *f) { return f->ptr[f->nelx]; }

void icxo_filter_neighbours(const icxo_filter *f, int *ptr, int *id, double *dist)
{
  memcpy(ptr, f->ptr, ((size_t)f->nelx + 1) * sizeof(int));
  memcpy(id, f->id, (size_t)f->ptr[f->nelx] * sizeof(int));
  memcpy(dist, f->dist, (size_t)f->ptr[f->nelx] * sizeof(double));
}

void icxo_filter_apply(icxo_filter *f, const double *in, double *out, int is_sens)
{
  const double r = f->radius;
  if (!is_sens) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < f->nelx; i++) {
      double aa = 0.0, bb = 0.0;
      for (int q = f->ptr[i]; q < f->ptr[i + 1]; q++) {
        double w = r - f->dist[q];
        aa += w;
        bb += w * in[f->id[q]];
      }
      out[i] = bb / aa;
      f->aa[i] = aa;
    }
  } else {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < f->nelx; i++) {
      double s = 0.0;
      for (int q = f->ptr[i]; q < f->ptr[i + 1]; q++) {
        double w = r - f->dist[q];
        s += w * in[f->id[q]] / f->aa[f->id[q]];
      }
      out[i] = s;
    }
  }
}

/* Filter::sensitivityFilter: src/optimization/filter.hpp:72-100 (aa summed afresh, the
 * neighbour terms as written there) */
void icxo_filter_sensitivity(const icxo_filter *f, const double *design, const double *dfds,
                             const double *volume, double *out)
{
  const double r = f->radius;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < f->nelx; i++) {
    double aa = 0.0, bb = 0.0;
    for (int q = f->ptr[i]; q < f->ptr[i + 1]; q++) {
      const int j = f->id[q];
      const double w = r - f->dist[q];
      aa += w;
      bb += (w * design[j] * dfds[j]) / (design[j] * volume[j]);
    }
    const double cc = (design[i] * volume[i]) / (design[i] * aa);
    out[i] = cc * bb;
  }
}

/* threshold / thresholdDerivative: src/optimization/filter.hpp:153-185 */
void icxo_threshold(int n, const double *in, double beta, double T, double *out)
{
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++)
    out[i] = 0.5 * (tanh(T * beta) + tanh(beta * (in[i] - T))) / tanh(T * beta);
}

void icxo_threshold_derivative(int n, const double *in, double beta, double T, double *out)
{
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++)
    out[i] = (0.5 * (beta * (1.0 - pow(tanh(beta * (in[i] - T)), 2.0))) / tanh(T * beta));
}

/* ------------------------------------------------------------------------ */
/* Helmholtz PDE filter: src/optimization/filter_helm.hpp:75-137, 178-256    */
/* ------------------------------------------------------------------------ */
struct icxo_helm {
  int ne, ndim, numnp, nelx;
  const int *conn;
  int *rowptr, *col; double *val; double *invdiag;
  double *Nint; /* nelx * ne */
  double *Nip;  /* ne x ne : N_n at Gauss point ip (row ip) */
};

icxo_helm *icxo_helm_create(int ne, int ndim, int numnp, int nelx, const double *x,
                            const int *conn, double radius)
{
  icxo_helm *h = calloc(1, sizeof(*h));
  h->ne = ne; h->ndim = ndim; h->numnp = numnp; h->nelx = nelx; h->conn = conn;
  adjacency A;
  build_adjacency(ne, numnp, nelx, conn, &A);
  h->rowptr = A.nptr; h->col = A.nid; A.nptr = NULL; A.nid = NULL;
  free(A.eptr); free(A.eid);
  h->val = calloc((size_t)h->rowptr[numnp], sizeof(double));
  h->Nint = malloc((size_t)nelx * ne * sizeof(double));
  int *cptr, *cord;
  int ncol = colour_elements(ne, numnp, nelx, conn, &cptr, &cord);
  for (int colour = 0; colour < ncol; colour++)
#pragma omp parallel for schedule(static)
  for (int t = cptr[colour]; t < cptr[colour + 1]; t++) {
    const int e = cord[t];
    double X[MAXNE * 3], Ke[MAXNE * MAXNE];
    for (int k = 0; k < ne; k++)
      for (int d = 0; d < ndim; d++) X[k * ndim + d] = x[3 * conn[ne * e + k] + d];
    icxo_helm_element(ne, ndim, radius, X, Ke, h->Nint + (size_t)ne * e);
    for (int a = 0; a < ne; a++) {
      int r = conn[ne * e + a];
      for (int b = 0; b < ne; b++) {
        int q = find_col(h->col, h->rowptr[r], h->rowptr[r + 1], conn[ne * e + b]);
        h->val[q] += Ke[a * ne + b];
      }
    }
  }
  free(cptr); free(cord);
  h->invdiag = malloc((size_t)numnp * sizeof(double));
  for (int r = 0; r < numnp; r++) {
    double d = h->val[find_col(h->col, h->rowptr[r], h->rowptr[r + 1], r)];
    h->invdiag[r] = d != 0.0 ? 1.0 / d : 1.0;
  }
  h->Nip = malloc((size_t)ne * ne * sizeof(double));
  for (int ip = 0; ip < ne; ip++) {
    double rr[3], N[MAXNE], dN[3 * MAXNE];
    gauss_point(ne, ip, rr);
    shape(ne, rr, N, dN);
    for (int n = 0; n < ne; n++) h->Nip[ip * ne + n] = N[n];
  }
  return h;
}

void icxo_helm_destroy(icxo_helm *h)
{
  if (!h) return;
  free(h->rowptr); free(h->col); free(h->val); free(h->invdiag); free(h->Nint);
  free(h->Nip); free(h);
}

/* forward (sens=0): f_n = sum_e rho_e int N_n; out_e = (1/ipmax) sum_ip N(ip).phi_e
 * adjoint (sens=1): f_n = sum_e s_e int N_n / ipmax; out_e = sum_ip N(ip).phi_e */
int icxo_helm_apply(icxo_helm *h, const double *in, double *out, int sens, double tol,
                    int maxit, int direct)
{
  const int ne = h->ne, n = h->numnp;
  double *f = calloc((size_t)n, sizeof(double)), *phi = calloc((size_t)n, sizeof(double));
  for (int e = 0; e < h->nelx; e++)
    for (int a = 0; a < ne; a++) {
      double v = in[e] * h->Nint[(size_t)ne * e + a];
      if (sens) v = v / ne;
      f[h->conn[ne * e + a]] += v;
    }
  int it = 0;
  double err;
  if (direct) {
    if (banded_cholesky_solve(n, h->rowptr, h->col, h->val, f, phi) != 0) it = -1;
  } else {
    it = pcg_eigen(n, h->rowptr, h->col, h->val, h->invdiag, f, phi, tol, maxit, &err);
  }
#pragma omp parallel for schedule(static)
  for (int e = 0; e < h->nelx; e++) {
    double s = 0.0;
    for (int ip = 0; ip < ne; ip++) {
      double t = 0.0;
      for (int a = 0; a < ne; a++) t += h->Nip[ip * ne + a] * phi[h->conn[ne * e + a]];
      s += t;
    }
    out[e] = sens ? s : s / ne;
  }
  free(f); free(phi);
  return it;
}

/* ------------------------------------------------------------------------ */
/* OC: src/optimization/opt-solver/OCSolver.hpp:20-68                        */
/* ------------------------------------------------------------------------ */
int icxo_oc_update(int n, double *x, const double *dfdx, double g, const double *dgdx,
                   const double *xmin, const double *xmax, double movelimit,
                   icxo_volume_fn gk, void *user)
{
  const double iota = 0.5, lmin = 0.0, lmax = 1.0e4, leps = 1.0e-5;
  const int maxiter = 100;
  double l0 = lmin, l1 = lmax;
  double *xn = malloc((size_t)n * sizeof(double));
  int counter = 0;
  while ((l1 - l0) / (l1 + l0) > leps && maxiter > counter) {
    double lambda = 0.5 * (l1 + l0);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
      double t = -dfdx[i] / (dgdx[i] * lambda);
      double v = pow(t > 0.0 ? t : 0.0, iota) * x[i];
      double lo = (1.0 - movelimit) * x[i];
      if (xmin[i] > lo) lo = xmin[i];
      double hi = (1.0 + movelimit) * x[i];
      if (xmax[i] < hi) hi = xmax[i];
      if (v < lo) v = lo;
      else if (v > hi) v = hi;
      xn[i] = v;
    }
    if (gk(xn, user) - g > 0.0) l0 = lambda; else l1 = lambda;
    counter++;
  }
  memcpy(x, xn, (size_t)n * sizeof(double));
  free(xn);
  return counter;
}

/* ------------------------------------------------------------------------ */
/* MMA (m constraints, dual interior point):                                  */
/* src/optimization/opt-solver/MMASolver.cpp:41-553                          */
/* ------------------------------------------------------------------------ */
struct icxo_mma {
  int n, m, iter;
  double xmamieps, epsimin, raa0, move, albefa, asyminit, asymdec, asyminc;
  double *a, *c, *d, *y, z, *lam, *mu, *s;
  double *low, *upp, *alpha, *beta, *p0, *q0, *pij, *qij, *b, *grad, *hess;
  double *xold1, *xold2;
  int last_outloop;
  /* GCMMA only (GCMMASolver.hpp:91-114) */
  int gcmma;
  double raa0eps, raaeps, r0, f0app;
  double *raa, *r, *fapp;
};

icxo_mma *icxo_mma_create(int n, int m, double ai, double ci, double di)
{
  icxo_mma *M = calloc(1, sizeof(*M));
  M->n = n; M->m = m; M->iter = 0;
  M->xmamieps = 1.0e-5; M->epsimin = sqrt((double)(n + m)) * 1e-9; M->raa0 = 0.00001;
  M->move = 0.15; M->albefa = 0.2; M->asyminit = 0.2; M->asymdec = 0.7; M->asyminc = 1.2;
  M->a = malloc(m * sizeof(double)); M->c = malloc(m * sizeof(double));
  M->d = malloc(m * sizeof(double)); M->y = calloc(m, sizeof(double));
  for (int j = 0; j < m; j++) { M->a[j] = ai; M->c[j] = ci; M->d[j] = di; }
  M->lam = calloc(m, sizeof(double)); M->mu = calloc(m, sizeof(double));
  M->s = calloc(2 * m, sizeof(double));
  size_t N = n;
  M->low = calloc(N, 8); M->upp = calloc(N, 8); M->alpha = calloc(N, 8); M->beta = calloc(N, 8);
  M->p0 = calloc(N, 8); M->q0 = calloc(N, 8); M->pij = calloc(N * m, 8); M->qij = calloc(N * m, 8);
  M->b = calloc(m, 8); M->grad = calloc(m, 8); M->hess = calloc((size_t)m * m, 8);
  M->xold1 = calloc(N, 8); M->xold2 = calloc(N, 8);
  return M;
}

void icxo_mma_destroy(icxo_mma *M)
{
  if (!M) return;
  free(M->a); free(M->c); free(M->d); free(M->y); free(M->lam); free(M->mu); free(M->s);
  free(M->low); free(M->upp); free(M->alpha); free(M->beta); free(M->p0); free(M->q0);
  free(M->pij); free(M->qij); free(M->b); free(M->grad); free(M->hess);
  free(M->xold1); free(M->xold2); free(M->raa); free(M->r); free(M->fapp); free(M);
}

/* asymptotes L, U from the x history (MMASolver.cpp:408-463, GCMMASolver.cpp:432-489) */
static void mma_asymptotes(icxo_mma *M, const double *xval, const double *xmin, const double *xmax)
{
  const int n = M->n;
  M->iter++;
  if (M->iter < 3) {
    for (int i = 0; i < n; i++) {
      M->low[i] = xval[i] - M->asyminit * (xmax[i] - xmin[i]);
      M->upp[i] = xval[i] + M->asyminit * (xmax[i] - xmin[i]);
    }
  } else {
    for (int i = 0; i < n; i++) {
      double zzz = (xval[i] - M->xold1[i]) * (M->xold1[i] - M->xold2[i]);
      double gamma = zzz < 0.0 ? M->asymdec : (zzz > 0.0 ? M->asyminc : 1.0);
      M->low[i] = xval[i] - gamma * (M->xold1[i] - M->low[i]);
      M->upp[i] = xval[i] + gamma * (M->upp[i] - M->xold1[i]);
      double xmami = fmax(M->xmamieps, xmax[i] - xmin[i]);
      M->low[i] = fmax(M->low[i], xval[i] - 100.0 * xmami);
      M->low[i] = fmin(M->low[i], xval[i] - 1.0e-5 * xmami);
      M->upp[i] = fmax(M->upp[i], xval[i] + 1.0e-5 * xmami);
      M->upp[i] = fmin(M->upp[i], xval[i] + 100.0 * xmami);
      double xmi = xmin[i] - 1.0e-6, xma = xmax[i] + 1.0e-6;
      if (xval[i] < xmi) {
        M->low[i] = xval[i] - (xma - xval[i]) / 0.9;
        M->upp[i] = xval[i] + (xma - xval[i]) / 0.9;
      }
      if (xval[i] > xma) {
        M->low[i] = xval[i] - (xval[i] - xmi) / 0.9;
        M->upp[i] = xval[i] + (xval[i] - xmi) / 0.9;
      }
    }
  }
}

static void mma_gensub(icxo_mma *M, const double *xval, const double *dfdx,
                       const double *gx, const double *dgdx, const double *xmin,
                       const double *xmax)
{
  const int n = M->n, m = M->m;
  mma_asymptotes(M, xval, xmin, xmax);
  for (int i = 0; i < n; i++) {
    double al = fmax(xmin[i], M->low[i] + M->albefa * (xval[i] - M->low[i]));
    al = fmax(al, xval[i] - M->move * (xmax[i] - xmin[i]));
    M->alpha[i] = fmin(al, xmax[i]);
    double be = fmin(xmax[i], M->upp[i] - M->albefa * (M->upp[i] - xval[i]));
    be = fmin(be, xval[i] + M->move * (xmax[i] - xmin[i]));
    M->beta[i] = fmax(be, xmin[i]);
    double dfdxp = fmax(0.0, dfdx[i]), dfdxm = fmax(0.0, -1.0 * dfdx[i]);
    double xmamiinv = 1.0 / fmax(M->xmamieps, xmax[i] - xmin[i]);
    double pq = 0.001 * fabs(dfdx[i]) + M->raa0 * xmamiinv;
    M->p0[i] = pow(M->upp[i] - xval[i], 2.0) * (dfdxp + pq);
    M->q0[i] = pow(xval[i] - M->low[i], 2.0) * (dfdxm + pq);
    for (int j = 0; j < m; j++) {
      double g = dgdx[i * m + j];
      double gp = fmax(0.0, g), gm = fmax(0.0, -1.0 * g);
      double pqj = 0.001 * fabs(g) + M->raa0 * xmamiinv;
      M->pij[i * m + j] = pow(M->upp[i] - xval[i], 2.0) * (gp + pqj);
      M->qij[i * m + j] = pow(xval[i] - M->low[i], 2.0) * (gm + pqj);
    }
  }
  for (int j = 0; j < m; j++) {
    M->b[j] = -gx[j];
    for (int i = 0; i < n; i++)
      M->b[j] += M->pij[i * m + j] / (M->upp[i] - xval[i]) +
                 M->qij[i * m + j] / (xval[i] - M->low[i]);
  }
}

static void mma_xyz(icxo_mma *M, double *x)
{
  const int n = M->n, m = M->m;
  double lamai = 0.0;
  for (int i = 0; i < m; i++) {
    if (M->lam[i] < 0.0) M->lam[i] = 0;
    M->y[i] = fmax(0.0, M->lam[i] - M->c[i]);
    lamai += M->lam[i] * M->a[i];
  }
  M->z = fmax(0.0, 10.0 * (lamai - 1.0));
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) {
    double pj = M->p0[i], qj = M->q0[i];
    for (int j = 0; j < m; j++) { pj += M->pij[i * m + j] * M->lam[j]; qj += M->qij[i * m + j] * M->lam[j]; }
    double v = (sqrt(pj) * M->low[i] + sqrt(qj) * M->upp[i]) / (sqrt(pj) + sqrt(qj));
    if (v < M->alpha[i]) v = M->alpha[i];
    if (v > M->beta[i]) v = M->beta[i];
    x[i] = v;
  }
}

static void mma_grad(icxo_mma *M, const double *x)
{
  for (int j = 0; j < M->m; j++) {
    double g = -M->b[j] - M->a[j] * M->z - M->y[j];
    for (int i = 0; i < M->n; i++)
      g += M->pij[i * M->m + j] / (M->upp[i] - x[i]) + M->qij[i * M->m + j] / (x[i] - M->low[i]);
    M->grad[j] = g;
  }
}

static void mma_hess(icxo_mma *M, const double *x)
{
  const int n = M->n, m = M->m;
  double *df2 = malloc((size_t)n * 8), *PQ = malloc((size_t)n * m * 8);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) {
    double pj = M->p0[i], qj = M->q0[i];
    for (int j = 0; j < m; j++) {
      pj += M->pij[i * m + j] * M->lam[j];
      qj += M->qij[i * m + j] * M->lam[j];
      PQ[i * m + j] = M->pij[i * m + j] / pow(M->upp[i] - x[i], 2.0) -
                      M->qij[i * m + j] / pow(x[i] - M->low[i], 2.0);
    }
    df2[i] = -1.0 / (2.0 * pj / pow(M->upp[i] - x[i], 3.0) + 2.0 * qj / pow(x[i] - M->low[i], 3.0));
    double xp = (sqrt(pj) * M->low[i] + sqrt(qj) * M->upp[i]) / (sqrt(pj) + sqrt(qj));
    if (xp < M->alpha[i]) df2[i] = 0.0;
    if (xp > M->beta[i]) df2[i] = 0.0;
  }
  for (int i = 0; i < m; i++)
    for (int j = 0; j < m; j++) {
      double h = 0.0;
      for (int k = 0; k < n; k++) h += (0.0 + PQ[k * m + i] * df2[k]) * PQ[k * m + j];
      M->hess[i * m + j] = h;
    }
  double lamai = 0.0;
  for (int j = 0; j < m; j++) {
    if (M->lam[j] < 0.0) M->lam[j] = 0.0;
    lamai += M->lam[j] * M->a[j];
    if (M->lam[j] > M->c[j]) M->hess[j * m + j] += -1.0;
    M->hess[j * m + j] += -M->mu[j] / M->lam[j];
  }
  if (lamai > 0.0)
    for (int j = 0; j < m; j++)
      for (int k = 0; k < m; k++) M->hess[j * m + k] += -10.0 * M->a[j] * M->a[k];
  double tr = 0.0;
  for (int i = 0; i < m; i++) tr += M->hess[i * m + i];
  double corr = 1e-4 * tr / m;
  if (-1.0 * corr < 1.0e-7) corr = -1.0e-7;
  for (int i = 0; i < m; i++) M->hess[i * m + i] += corr;
  free(df2); free(PQ);
}

static double mma_residual(icxo_mma *M, const double *x, double epsi)
{
  double nr = 0.0;
  for (int j = 0; j < M->m; j++) {
    double r = -M->b[j] - M->a[j] * M->z - M->y[j] + M->mu[j];
    for (int i = 0; i < M->n; i++)
      r += M->pij[i * M->m + j] / (M->upp[i] - x[i]) + M->qij[i * M->m + j] / (x[i] - M->low[i]);
    double r2 = M->mu[j] * M->lam[j] - epsi;
    if (nr < fabs(r)) nr = fabs(r);
    if (nr < fabs(r2)) nr = fabs(r2);
  }
  return nr;
}

static void mma_linesearch(icxo_mma *M)
{
  double theta = 1.005;
  for (int i = 0; i < M->m; i++) {
    if (theta < -1.01 * M->s[i] / M->lam[i]) theta = -1.01 * M->s[i] / M->lam[i];
    if (theta < -1.01 * M->s[i + M->m] / M->mu[i]) theta = -1.01 * M->s[i + M->m] / M->mu[i];
  }
  theta = 1.0 / theta;
  for (int i = 0; i < M->m; i++) {
    M->lam[i] = M->lam[i] + theta * M->s[i];
    M->mu[i] = M->mu[i] + theta * M->s[i + M->m];
  }
}

static void mma_factorize_solve(double *K, double *x, int n)
{
  for (int s = 0; s < n - 1; s++)
    for (int i = s + 1; i < n; i++) {
      K[i * n + s] = K[i * n + s] / K[s * n + s];
      for (int j = s + 1; j < n; j++) K[i * n + j] = K[i * n + j] - K[i * n + s] * K[s * n + j];
    }
  for (int i = 1; i < n; i++) {
    double a = 0.0;
    for (int j = 0; j < i; j++) a = a - K[i * n + j] * x[j];
    x[i] = x[i] + a;
  }
  x[n - 1] = x[n - 1] / K[(n - 1) * n + (n - 1)];
  for (int i = n - 2; i >= 0; i--) {
    double a = x[i];
    for (int j = i + 1; j < n; j++) a = a - K[i * n + j] * x[j];
    x[i] = a / K[i * n + i];
  }
}

static int mma_solve_dip(icxo_mma *M, double *x)
{
  const int m = M->m;
  for (int j = 0; j < m; j++) { M->lam[j] = M->c[j] / 2.0; M->mu[j] = 1.0; }
  const double tol = M->epsimin;
  double epsi = 1.0, err = 1.0;
  int outloop = 0, total = 0;
  while (epsi > tol && outloop < 100) {
    outloop++;
    int loop = 0;
    while (err > 0.9 * epsi && loop < 100) {
      loop++; total++;
      mma_xyz(M, x);
      mma_grad(M, x);
      for (int j = 0; j < m; j++) M->grad[j] = -1.0 * M->grad[j] - epsi / M->lam[j];
      mma_hess(M, x);
      if (m > 1) {
        mma_factorize_solve(M->hess, M->grad, m);
        for (int j = 0; j < m; j++) M->s[j] = M->grad[j];
      } else if (m > 0) {
        M->s[0] = M->grad[0] / M->hess[0];
      }
      for (int i = 0; i < m; i++)
        M->s[m + i] = -M->mu[i] + epsi / M->lam[i] - M->s[i] * M->mu[i] / M->lam[i];
      mma_linesearch(M);
      mma_xyz(M, x);
      err = mma_residual(M, x, epsi);
    }
    epsi = epsi * 0.1;
  }
  M->last_outloop = outloop;
  return total;
}

int icxo_mma_update(icxo_mma *M, double *x, const double *dfdx, const double *gx,
                    const double *dgdx, const double *xmin, const double *xmax)
{
  mma_gensub(M, x, dfdx, gx, dgdx, xmin, xmax);
  memcpy(M->xold2, M->xold1, (size_t)M->n * 8);
  memcpy(M->xold1, x, (size_t)M->n * 8);
  return mma_solve_dip(M, x);
}

static void copy_opt(double *dst, const double *src, size_t n)
{
  if (dst) memcpy(dst, src, n * sizeof(double));
}

void icxo_mma_subproblem(const icxo_mma *M, double *low, double *upp, double *alpha,
                         double *beta, double *p0, double *q0, double *pij, double *qij,
                         double *b, double *lam, double *y, double *z)
{
  size_t n = M->n, m = M->m;
  copy_opt(low, M->low, n); copy_opt(upp, M->upp, n);
  copy_opt(alpha, M->alpha, n); copy_opt(beta, M->beta, n);
  copy_opt(p0, M->p0, n); copy_opt(q0, M->q0, n);
  copy_opt(pij, M->pij, n * m); copy_opt(qij, M->qij, n * m);
  copy_opt(b, M->b, m); copy_opt(lam, M->lam, m); copy_opt(y, M->y, m);
  if (z) *z = M->z;
}

/* MMASolver::SetAsymptotes (src/optimization/opt-solver/MMASolver.hpp:91-97) */
void icxo_mma_set_asymptotes(icxo_mma *M, double init, double decrease, double increase)
{
  M->asyminit = init; M->asymdec = decrease; M->asyminc = increase;
}

/* ------------------------------------------------------------------------ */
/* GCMMA (m constraints): src/optimization/opt-solver/GCMMASolver.cpp:41-620. */
/* The subproblem and its dual interior point solve (SolveDIP and helpers,   */
/* :135-429) are those of MMA above; GCMMA adds the conservative raa0 / raa  */
/* terms, the constants r0 / r of the approximations and their values at the */
/* new point (ComputeApprox), which ConCheck compares with the true values.  */
/* ------------------------------------------------------------------------ */
icxo_mma *icxo_gcmma_create(int n, int m, double ai, double ci, double di)
{
  icxo_mma *M = icxo_mma_create(n, m, ai, ci, di); /* move 0.15, albefa 0.2, asym 0.2/0.7/1.2 */
  M->gcmma = 1;
  M->raa0eps = 1e-6; M->raaeps = 1e-6; M->raa0 = 1.0;
  M->raa = malloc(m * sizeof(double)); M->r = calloc(m, sizeof(double));
  M->fapp = calloc(m, sizeof(double));
  for (int j = 0; j < m; j++) M->raa[j] = 1.0;
  return M;
}

/* GenSub (GCMMASolver.cpp:546-596): alpha / beta without the MMA clamps, raa0 and raa[j]
 * in the p / q coefficients, r0 = f0x - sum(p0/(U-x) + q0/(x-L)), r = fx - ..., b = -r */
static void gcmma_gensub(icxo_mma *M, const double *xval, double f0x, const double *df0dx,
                         const double *fx, const double *dfdx, const double *xmin,
                         const double *xmax)
{
  const int n = M->n, m = M->m;
  M->r0 = 0.0;
  for (int j = 0; j < m; j++) M->r[j] = 0.0;
  for (int i = 0; i < n; i++) {
    M->alpha[i] = fmax(xmin[i], M->low[i] + M->albefa * (xval[i] - M->low[i]));
    M->alpha[i] = fmax(M->alpha[i], xval[i] - M->move * (xmax[i] - xmin[i]));
    M->beta[i] = fmin(xmax[i], M->upp[i] - M->albefa * (M->upp[i] - xval[i]));
    M->beta[i] = fmin(M->beta[i], xval[i] + M->move * (xmax[i] - xmin[i]));
    const double uxinv = 1.0 / (M->upp[i] - xval[i]), xlinv = 1.0 / (xval[i] - M->low[i]);
    const double xmamiinv = 1.0 / fmax(M->xmamieps, xmax[i] - xmin[i]);
    const double pq = 0.001 * fabs(df0dx[i]) + M->raa0 * xmamiinv;
    M->p0[i] = pow(M->upp[i] - xval[i], 2.0) * (fmax(0.0, df0dx[i]) + pq);
    M->q0[i] = pow(xval[i] - M->low[i], 2.0) * (fmax(0.0, -1.0 * df0dx[i]) + pq);
    M->r0 += M->p0[i] * uxinv + M->q0[i] * xlinv;
    for (int j = 0; j < m; j++) {
      const double g = dfdx[i * m + j];
      const double pqj = 0.001 * fabs(g) + M->raa[j] * xmamiinv;
      M->pij[i * m + j] = pow(M->upp[i] - xval[i], 2.0) * (fmax(0.0, g) + pqj);
      M->qij[i * m + j] = pow(xval[i] - M->low[i], 2.0) * (fmax(0.0, -1.0 * g) + pqj);
      M->r[j] += M->pij[i * m + j] * uxinv + M->qij[i * m + j] * xlinv;
    }
  }
  M->r0 = f0x - M->r0;
  for (int j = 0; j < m; j++) {
    M->r[j] = fx[j] - M->r[j];
    M->b[j] = -M->r[j];
  }
}

/* ComputeApprox (GCMMASolver.cpp:598-620) */
static void gcmma_approx(icxo_mma *M, const double *xmma)
{
  const int n = M->n, m = M->m;
  M->f0app = 0.0;
  for (int j = 0; j < m; j++) M->fapp[j] = 0.0;
  for (int i = 0; i < n; i++) {
    const double uxinv = 1.0 / (M->upp[i] - xmma[i]), xlinv = 1.0 / (xmma[i] - M->low[i]);
    M->f0app += M->p0[i] * uxinv + M->q0[i] * xlinv;
    for (int j = 0; j < m; j++) M->fapp[j] += M->pij[i * m + j] * uxinv + M->qij[i * m + j] * xlinv;
  }
  M->f0app += M->r0;
  for (int j = 0; j < m; j++) M->fapp[j] += M->r[j];
}

/* OuterUpdate (GCMMASolver.cpp:66-90): Asymp (asymptotes, then raa0 / raa decay to
 * max(eps, 0.1 raa), :432-507), GenSub, x history, dual solve from xval, ComputeApprox */
int icxo_gcmma_outer(icxo_mma *M, double *xmma, const double *xval, double f0x,
                     const double *df0dx, const double *fx, const double *dfdx,
                     const double *xmin, const double *xmax)
{
  mma_asymptotes(M, xval, xmin, xmax);
  M->raa0 = fmax(M->raa0eps, 0.1 * M->raa0);
  for (int j = 0; j < M->m; j++) M->raa[j] = fmax(M->raaeps, 0.1 * M->raa[j]);
  gcmma_gensub(M, xval, f0x, df0dx, fx, dfdx, xmin, xmax);
  memcpy(M->xold2, M->xold1, (size_t)M->n * 8);
  memcpy(M->xold1, xval, (size_t)M->n * 8);
  memcpy(xmma, xval, (size_t)M->n * 8);
  const int it = mma_solve_dip(M, xmma);
  gcmma_approx(M, xmma);
  return it;
}

/* InnerUpdate (GCMMASolver.cpp:92-113) with RaaUpdate (:510-544): raa0 / raa grow where
 * the approximation at the last xmma fell below the true value; asymptotes and the x
 * history stay */
int icxo_gcmma_inner(icxo_mma *M, double *xmma, double f0xnew, const double *fxnew,
                     const double *xval, double f0x, const double *df0dx, const double *fx,
                     const double *dfdx, const double *xmin, const double *xmax)
{
  double raacof = 0.0;
  for (int i = 0; i < M->n; i++) {
    const double xmami = fmax(M->xmamieps, xmax[i] - xmin[i]);
    const double xxux = (xmma[i] - xval[i]) / (M->upp[i] - xmma[i]);
    const double xxxl = (xmma[i] - xval[i]) / (xmma[i] - M->low[i]);
    raacof += xxux * xxxl * ((M->upp[i] - M->low[i]) / xmami);
  }
  raacof = fmax(1e-12, raacof);
  if (f0xnew > M->f0app + 0.5 * M->epsimin) {
    const double d0 = (1.0 / raacof) * (f0xnew - M->f0app);
    M->raa0 = fmin(1.1 * (M->raa0 + d0), 10.0 * M->raa0);
  }
  for (int j = 0; j < M->m; j++)
    if (fxnew[j] > M->fapp[j] + 0.5 * M->epsimin) {
      const double dj = (1.0 / raacof) * (fxnew[j] - M->fapp[j]);
      M->raa[j] = fmin(1.1 * (M->raa[j] + dj), 10.0 * M->raa[j]);
    }
  gcmma_gensub(M, xval, f0x, df0dx, fx, dfdx, xmin, xmax);
  memcpy(xmma, xval, (size_t)M->n * 8);
  const int it = mma_solve_dip(M, xmma);
  gcmma_approx(M, xmma);
  return it;
}

/* ConCheck (GCMMASolver.cpp:115-129) */
int icxo_gcmma_concheck(const icxo_mma *M, double f0xnew, const double *fxnew)
{
  if (M->f0app + M->epsimin < f0xnew) return 0;
  for (int j = 0; j < M->m; j++)
    if (M->fapp[j] + M->epsimin < fxnew[j]) return 0;
  return 1;
}

/* the GCMMA approximations (the subproblem's functions) at any x (test access) */
void icxo_gcmma_approx_at(const icxo_mma *M, const double *x, double *f0, double *fj)
{
  double s0 = 0.0;
  for (int j = 0; j < M->m; j++) fj[j] = M->r[j];
  for (int i = 0; i < M->n; i++) {
    const double uxinv = 1.0 / (M->upp[i] - x[i]), xlinv = 1.0 / (x[i] - M->low[i]);
    s0 += M->p0[i] * uxinv + M->q0[i] * xlinv;
    for (int j = 0; j < M->m; j++) fj[j] += M->pij[i * M->m + j] * uxinv + M->qij[i * M->m + j] * xlinv;
  }
  *f0 = s0 + M->r0;
}

void icxo_gcmma_state(const icxo_mma *M, double *f0app, double *fapp, double *raa0, double *raa)
{
  if (f0app) *f0app = M->f0app;
  if (raa0) *raa0 = M->raa0;
  for (int j = 0; j < M->m; j++) {
    if (fapp) fapp[j] = M->fapp[j];
    if (raa) raa[j] = M->raa[j];
  }
}

/* ------------------------------------------------------------------------ */
/* driver loops                                                               */
/* ------------------------------------------------------------------------ */
typedef struct {
  icxo_filter *f;
  int n; double beta, T; int projection;
  const double *volume;
  double *tmp, *tmp2;
} oc_vol_ctx;

static double oc_volume(const double *x, void *user)
{
  oc_vol_ctx *c = user;
  icxo_filter_apply(c->f, x, c->tmp, 0);
  const double *h = c->tmp;
  if (c->projection) { icxo_threshold(c->n, c->tmp, c->beta, c->T, c->tmp2); h = c->tmp2; }
  double g = 0.0;
  for (int i = 0; i < c->n; i++) g += h[i] * c->volume[i];
  return g;
}

/* objective and constraint of a trial design of the GCMMA driver
 * (opt_macro_gcmma.cpp:158-175, 192-207): filter, threshold, solve, f = u.fext,
 * g = sum rho_hat v - V0 */
static int gcmma_trial(const icxo_driver_cfg *g, icxo_sys *s, icxo_filter *flt, const double *xnew,
                       double beta, double V0, double *fil, double *thr, double *fint, double *rhs,
                       double *u, double *f, double *gv)
{
  const int nelx = g->nelx, numeq = g->numeq;
  icxo_filter_apply(flt, xnew, fil, 0);
  if (g->projection) icxo_threshold(nelx, fil, beta, 0.5, thr);
  else memcpy(thr, fil, (size_t)nelx * 8);
  icxo_sys_assemble(s, thr, g->has, g->node_flag, g->node_val, fint);
  for (int i = 0; i < numeq; i++) rhs[i] = g->fext[i] - fint[i];
  if (g->direct) { if (icxo_sys_direct(s, rhs, u) != 0) return -1; }
  else { double err; icxo_sys_cg(s, rhs, u, g->cg_tol, g->cg_maxit, &err); }
  double o = 0.0, v = 0.0;
  for (int i = 0; i < numeq; i++) o += u[i] * g->fext[i];
  for (int i = 0; i < nelx; i++) v += thr[i] * s->volume[i];
  *f = o;
  *gv = v - V0;
  return 0;
}

int icxo_run_driver(const icxo_driver_cfg *g, double *design, double *obj_hist,
                    double *vol_hist, double *last_dfds, double *last_u, int *cg_hist)
{
  const int nelx = g->nelx, numeq = g->numeq;
  const double T = 0.5;
  icxo_sys *s = icxo_sys_create(g->ne, g->ndim, g->dofnp, g->state, g->numnp, nelx, g->x,
                                g->conn, g->dof, numeq, g->young1, g->young2, g->nu, g->p);
  icxo_filter *flt = NULL;
  icxo_helm *hlm = NULL;
  icxo_mma *mma = NULL;
  const int cone = g->driver != 1;
  if (g->driver == 3 && g->ndim != 2) return -1;
  if (cone) flt = icxo_filter_create(g->ne, g->ndim, nelx, g->x, g->conn, g->filter_radius, 0);
  else {
    hlm = icxo_helm_create(g->ne, g->ndim, g->numnp, nelx, g->x, g->conn, g->filter_radius);
    mma = icxo_mma_create(nelx, 1, 0.0, 100000.0, 0.0);
  }
  if (g->driver == 2) mma = icxo_gcmma_create(nelx, 1, 0.0, 1000.0, 0.0);
  if (g->driver == 3) { /* MMASolver(n, 1) with SetAsymptotes(0.1, 0.7, 1.2) */
    mma = icxo_mma_create(nelx, 1, 0.0, 100000.0, 0.0);
    mma->asyminit = 0.1;
  }
  size_t E = (size_t)nelx * sizeof(double);
  double *fil = malloc(E), *thr = malloc(E), *dthr = malloc(E), *dfds = malloc(E),
         *dgds = malloc(E), *energy = malloc(E), *tmp = malloc(E), *tmp2 = malloc(E),
         *xmin = malloc(E), *xmax = malloc(E);
  double *fint = malloc((size_t)numeq * 8), *rhs = malloc((size_t)numeq * 8),
         *u = malloc((size_t)numeq * 8), *nodeval = malloc((size_t)g->numnp * 3 * 8);
  double *xnew = malloc(E), *u2 = malloc((size_t)numeq * 8);
  for (int i = 0; i < nelx; i++) { xmin[i] = g->xmin; xmax[i] = g->xmax; }
  double V0 = 0.0, gcon = 0.0;
  int rc = 0;
  for (int k = 1; k <= g->iterations; k++) {
    if (cone) icxo_filter_apply(flt, design, fil, 0);
    else icxo_helm_apply(hlm, design, fil, 0, g->helm_tol, 10000, g->helm_direct);
    double beta = pow(2, (int)(k / 40));
    if (beta > 16.0) beta = 16.0;
    if (g->driver == 2) { /* opt_macro_gcmma.cpp:91-92 */
      const int bt = (int)fmax((k - 100) / 40.0, 1.0);
      beta = fmin((double)bt, 8.0);
    }
    if (g->driver == 3) beta = fmin(pow(2.0, (int)(k / 40)), 8.0); /* opt_stress_constraint.cpp:94 */
    if (g->projection) {
      icxo_threshold(nelx, fil, beta, T, thr);
      icxo_threshold_derivative(nelx, fil, beta, T, dthr);
    } else {
      memcpy(thr, fil, E);
      for (int i = 0; i < nelx; i++) dthr[i] = 1.0;
    }
    icxo_sys_assemble(s, thr, g->has, g->node_flag, g->node_val, fint);
    for (int i = 0; i < numeq; i++) rhs[i] = g->fext[i] - fint[i];
    int it = 0;
    if (g->direct) { if (icxo_sys_direct(s, rhs, u) != 0) { rc = -1; break; } }
    else { double err; it = icxo_sys_cg(s, rhs, u, g->cg_tol, g->cg_maxit, &err); }
    if (cg_hist) cg_hist[k - 1] = it;
    icxo_divide_val_into_node(g->numnp, g->dofnp, numeq, g->dof, u, g->node_val, nodeval);
    if (g->driver == 3) { /* opt_stress_constraint.cpp:103-150: f = volume, g = stress p-norm */
      double vol = 0.0;
      for (int i = 0; i < nelx; i++) vol += thr[i] * s->volume[i];
      if (icxo_stress_constraint(s, nodeval, g->node_val, g->qp, g->p_norm, g->stress_limit, g->direct,
                                 g->cg_tol, &gcon, dgds, NULL)) { rc = -1; break; }
      obj_hist[k - 1] = vol;
      if (vol_hist) vol_hist[k - 1] = gcon;
      for (int i = 0; i < nelx; i++) { dfds[i] = s->volume[i] * dthr[i]; dgds[i] *= dthr[i]; }
      icxo_filter_apply(flt, dfds, tmp, 1); memcpy(dfds, tmp, E);
      icxo_filter_apply(flt, dgds, tmp, 1); memcpy(dgds, tmp, E);
      if (k == g->iterations) {
        if (last_dfds) memcpy(last_dfds, dgds, E);
        if (last_u) memcpy(last_u, u, (size_t)numeq * 8);
      }
      double gx[1] = {gcon};
      icxo_mma_update(mma, design, dfds, gx, dgds, xmin, xmax);
      continue;
    }
    double obj = 0.0;
    for (int i = 0; i < numeq; i++) obj += u[i] * g->fext[i];
    if (k == 1) {
      V0 = 0.0;
      for (int i = 0; i < nelx; i++) V0 += thr[i] * s->volume[i];
      gcon = (g->driver == 0) ? V0 : 0.0;
    } else {
      gcon = 0.0;
      if (g->driver != 1) {
        for (int i = 0; i < nelx; i++) gcon += thr[i] * s->volume[i];
        if (g->driver == 2) gcon -= V0;
      } else {
        for (int i = 0; i < nelx; i++) gcon += thr[i] * s->volume[i] / V0;
        gcon -= 1;
      }
    }
    obj_hist[k - 1] = obj;
    if (vol_hist) vol_hist[k - 1] = gcon;
    icxo_sys_element_energy(s, nodeval, energy);
    icxo_sensitivity(nelx, g->driver == 1, g->young1, g->young2, g->p, thr, energy, s->volume,
                     V0, dfds, dgds);
    for (int i = 0; i < nelx; i++) { dfds[i] *= dthr[i]; dgds[i] *= dthr[i]; }
    if (cone) {
      icxo_filter_apply(flt, dfds, tmp, 1); memcpy(dfds, tmp, E);
      icxo_filter_apply(flt, dgds, tmp, 1); memcpy(dgds, tmp, E);
    } else {
      icxo_helm_apply(hlm, dfds, tmp, 1, g->helm_tol, 10000, g->helm_direct); memcpy(dfds, tmp, E);
      icxo_helm_apply(hlm, dgds, tmp, 1, g->helm_tol, 10000, g->helm_direct); memcpy(dgds, tmp, E);
    }
    if (k == g->iterations) {
      if (last_dfds) memcpy(last_dfds, dfds, E);
      if (last_u) memcpy(last_u, u, (size_t)numeq * 8);
    }
    if (g->driver == 0) {
      oc_vol_ctx c = {flt, nelx, beta, T, g->projection, s->volume, tmp, tmp2};
      icxo_oc_update(nelx, design, dfds, V0, dgds, xmin, xmax, g->move, oc_volume, &c);
    } else if (g->driver == 1) {
      double gx[1] = {gcon};
      icxo_mma_update(mma, design, dfds, gx, dgds, xmin, xmax);
    } else { /* opt_macro_gcmma.cpp:150-215 */
      double gx[1] = {gcon}, fnew, gnew[1];
      icxo_gcmma_outer(mma, xnew, design, obj, dfds, gx, dgds, xmin, xmax);
      if (gcmma_trial(g, s, flt, xnew, beta, V0, tmp, tmp2, fint, rhs, u2, &fnew, gnew)) { rc = -1; break; }
      int conserv = icxo_gcmma_concheck(mma, fnew, gnew), inner = 0;
      for (; !conserv && inner < 15; ++inner) {
        icxo_gcmma_inner(mma, xnew, fnew, gnew, design, obj, dfds, gx, dgds, xmin, xmax);
        if (gcmma_trial(g, s, flt, xnew, beta, V0, tmp, tmp2, fint, rhs, u2, &fnew, gnew)) { rc = -1; break; }
        conserv = icxo_gcmma_concheck(mma, fnew, gnew);
      }
      if (rc) break;
      if (g->inner_hist) g->inner_hist[k - 1] = inner;
      memcpy(design, xnew, E);
    }
  }
  free(fil); free(thr); free(dthr); free(dfds); free(dgds); free(energy); free(tmp);
  free(tmp2); free(xmin); free(xmax); free(fint); free(rhs); free(u); free(nodeval);
  free(xnew); free(u2);
  icxo_sys_destroy(s); icxo_filter_destroy(flt); icxo_helm_destroy(hlm); icxo_mma_destroy(mma);
  return rc;
}
