// HIP kernels (gfx950) of libicx: the matrix-free structured-grid operators,
// the PCG vector kernels, sensitivity, filters, projection, OC and MMA pieces.
//
// Reductions are deterministic: every block reduces in a fixed tree, writes one
// partial, and the last block to arrive (atomic ticket) sums the partials in
// index order.  Nothing depends on which block finishes first, so a solve gives
// the same bits alone and in a batch.
#include "icx_internal.h"

#include <hip/hip_runtime.h>
#include <math.h>

namespace icx {

#define ICX_CHECK_LAUNCH() (void)hipGetLastError()

// ---------------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Reduces NV doubles over the grid.  Returns true in every thread of the last
// block, with tot[] holding the grid totals.
template <int NV>
__device__ bool grid_reduce(const Reduce &red, double (&v)[NV], double (&tot)[NV]) {
  __shared__ double sh[NV][16];
  __shared__ int am_last;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double w = wave_sum(v[k]);
    if (lane == 0) sh[k][wid] = w;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = 0.0;
      for (int w = 0; w < nw; ++w) s += sh[k][w];
      __hip_atomic_store(&red.partial[k * RED_MAXBLOCKS + blockIdx.x], s, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    }
    __threadfence();
    unsigned prev = atomicAdd(red.counter, 1u);
    am_last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (!am_last) return false;
  __threadfence();
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    acc[k] = 0.0;
    for (unsigned b = threadIdx.x; b < gridDim.x; b += blockDim.x)
      acc[k] += __hip_atomic_load(&red.partial[k * RED_MAXBLOCKS + b], __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();  // sh reused
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double w = wave_sum(acc[k]);
    if (lane == 0) sh[k][wid] = w;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double s = 0.0;
    for (int w = 0; w < nw; ++w) s += sh[k][w];
    tot[k] = s;
  }
  if (threadIdx.x == 0) __hip_atomic_store(red.counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return true;
}

static inline int ew_blocks(long long n, int per_thread = 1) {
  long long b = (n + 256LL * per_thread - 1) / (256LL * per_thread);
  if (b < 1) b = 1;
  if (b > 2048) b = 2048;
  return (int)b;
}

// ---------------------------------------------------------------------------------
// element operators in the Haar basis
// ---------------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void haar_fwd8(T (&v)[8]) {
#pragma unroll
  for (int q = 0; q < 8; q += 2) { T s = v[q] + v[q + 1], d = v[q + 1] - v[q]; v[q] = s; v[q + 1] = d; }
#pragma unroll
  for (int q : {0, 1, 4, 5}) { T s = v[q] + v[q + 2], d = v[q + 2] - v[q]; v[q] = s; v[q + 2] = d; }
#pragma unroll
  for (int q = 0; q < 4; ++q) { T s = v[q] + v[q + 4], d = v[q + 4] - v[q]; v[q] = s; v[q + 4] = d; }
}
template <typename T>
__device__ __forceinline__ void haar_inv8(T (&v)[8]) {
#pragma unroll
  for (int q = 0; q < 4; ++q) { T a = v[q], b = v[q + 4]; v[q] = a - b; v[q + 4] = a + b; }
#pragma unroll
  for (int q : {0, 1, 4, 5}) { T a = v[q], b = v[q + 2]; v[q] = a - b; v[q + 2] = a + b; }
#pragma unroll
  for (int q = 0; q < 8; q += 2) { T a = v[q], b = v[q + 1]; v[q] = a - b; v[q + 1] = a + b; }
}
template <typename T>
__device__ __forceinline__ void haar_fwd4(T (&v)[4]) {
#pragma unroll
  for (int q = 0; q < 4; q += 2) { T s = v[q] + v[q + 1], d = v[q + 1] - v[q]; v[q] = s; v[q + 1] = d; }
#pragma unroll
  for (int q = 0; q < 2; ++q) { T s = v[q] + v[q + 2], d = v[q + 2] - v[q]; v[q] = s; v[q + 2] = d; }
}
template <typename T>
__device__ __forceinline__ void haar_inv4(T (&v)[4]) {
#pragma unroll
  for (int q = 0; q < 2; ++q) { T a = v[q], b = v[q + 2]; v[q] = a - b; v[q + 2] = a + b; }
#pragma unroll
  for (int q = 0; q < 4; q += 2) { T a = v[q], b = v[q + 1]; v[q] = a - b; v[q + 1] = a + b; }
}

// Structural non-zeros of the character blocks.  Character chi has bit d set
// when the mode is odd under the mirror of axis d.  Member c of block chi is the
// Haar coefficient of displacement component c with difference pattern
// chi ^ (1 << c).  Blocks with one odd axis d hold the rigid translation along
// d (row/column c = d vanishes); blocks with two odd axes hold a rigid rotation
// coupling those two components, and the remaining component is uncoupled.
__host__ __device__ constexpr int popc3(int v) { return (v & 1) + ((v >> 1) & 1) + ((v >> 2) & 1); }
__host__ __device__ constexpr bool blk_nz3(int chi, int c, int cc) {
  return (popc3(chi) == 0 || popc3(chi) == 3)
             ? true
             : (popc3(chi) == 1
                    ? (c != ((chi & 1) ? 0 : ((chi & 2) ? 1 : 2)) &&
                       cc != ((chi & 1) ? 0 : ((chi & 2) ? 1 : 2)))
                    : ((c == (!(chi & 1) ? 0 : (!(chi & 2) ? 1 : 2))) ==
                       (cc == (!(chi & 1) ? 0 : (!(chi & 2) ? 1 : 2)))));
}
__host__ __device__ constexpr bool blk_nz2(int chi, int c, int cc) {
  return (chi == 0 || chi == 3) ? true : (c != ((chi & 1) ? 0 : 1) && cc != ((chi & 1) ? 0 : 1));
}

template <typename T>
__device__ __forceinline__ void elastic_hex8(T (&v)[3][8], const ElasticOp<T> &op, T E) {
  haar_fwd8(v[0]); haar_fwd8(v[1]); haar_fwd8(v[2]);
  T o[3][8];
#pragma unroll
  for (int chi = 0; chi < 8; ++chi) {
    const T m[3] = {v[0][chi ^ 1], v[1][chi ^ 2], v[2][chi ^ 4]};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      T acc = T(0);
#pragma unroll
      for (int cc = 0; cc < 3; ++cc)
        if (blk_nz3(chi, c, cc)) acc += op.blk[chi][c][cc] * m[cc];
      o[c][chi ^ (1 << c)] = E * acc;
    }
  }
  haar_inv8(o[0]); haar_inv8(o[1]); haar_inv8(o[2]);
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int q = 0; q < 8; ++q) v[c][q] = o[c][q];
}

template <typename T>
__device__ __forceinline__ void elastic_quad4(T (&v)[2][4], const ElasticOp<T> &op, T E) {
  haar_fwd4(v[0]); haar_fwd4(v[1]);
  T o[2][4];
#pragma unroll
  for (int chi = 0; chi < 4; ++chi) {
    const T m[2] = {v[0][chi ^ 1], v[1][chi ^ 2]};
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      T acc = T(0);
#pragma unroll
      for (int cc = 0; cc < 2; ++cc)
        if (blk_nz2(chi, c, cc)) acc += op.blk[chi][c][cc] * m[cc];
      o[c][chi ^ (1 << c)] = E * acc;
    }
  }
  haar_inv4(o[0]); haar_inv4(o[1]);
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int q = 0; q < 4; ++q) v[c][q] = o[c][q];
}

// u^T K0 u in the Haar basis: sum over characters of m^T B m
template <typename T>
__device__ __forceinline__ T energy_hex8(T (&v)[3][8], const ElasticOp<T> &op) {
  haar_fwd8(v[0]); haar_fwd8(v[1]); haar_fwd8(v[2]);
  T e = T(0);
#pragma unroll
  for (int chi = 0; chi < 8; ++chi) {
    const T m[3] = {v[0][chi ^ 1], v[1][chi ^ 2], v[2][chi ^ 4]};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      T acc = T(0);
#pragma unroll
      for (int cc = 0; cc < 3; ++cc)
        if (blk_nz3(chi, c, cc)) acc += op.blk[chi][c][cc] * m[cc];
      e += m[c] * acc;
    }
  }
  return e;
}
template <typename T>
__device__ __forceinline__ T energy_quad4(T (&v)[2][4], const ElasticOp<T> &op) {
  haar_fwd4(v[0]); haar_fwd4(v[1]);
  T e = T(0);
#pragma unroll
  for (int chi = 0; chi < 4; ++chi) {
    const T m[2] = {v[0][chi ^ 1], v[1][chi ^ 2]};
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      T acc = T(0);
#pragma unroll
      for (int cc = 0; cc < 2; ++cc)
        if (blk_nz2(chi, c, cc)) acc += op.blk[chi][c][cc] * m[cc];
      e += m[c] * acc;
    }
  }
  return e;
}

// ---------------------------------------------------------------------------------
// XCD-aware tile order: consecutive hardware block ids go round-robin over the 8
// XCDs; remap so each XCD works on a contiguous run of tiles (shared halo lines
// stay in that XCD's L2).
// ---------------------------------------------------------------------------------
__device__ __forceinline__ int xcd_tile(int b, int nblocks) {
  const int per = nblocks >> 3, rem = nblocks & 7;
  const int xcd = b & 7, local = b >> 3;
  return xcd < rem ? xcd * (per + 1) + local : rem * (per + 1) + (xcd - rem) * per + local;
}

// ---------------------------------------------------------------------------------
// structured operator kernels
//
// hexa8: a block owns a TI x TJ x TK brick of nodes and marches along z.  Per
// element layer it stages the next node layer in LDS (halo of one node), every
// thread applies one element (the (TI+1)(TJ+1) elements touching the owned
// nodes), writes its 24 outputs to LDS, and every owned node sums the four
// in-layer contributions; the contributions to the layer above are carried in
// registers.  Summation order is fixed, so results are deterministic.
//
// MODE 0: y = K x            (in = x)
// MODE 1: p = invd r + beta pold, y = K p, xacc += alpha_x pold  (in = r)
// ---------------------------------------------------------------------------------
template <typename T> struct OpArgs {
  Grid g;
  int TK, tiles_x, tiles_y, tiles_z;
  ElasticOp<T> op;
  const T *E;
  const uint8_t *mask;
  const T *in, *invd, *pold;
  T *pnew, *y, *xacc;
  SolverState *st;
  Reduce red;
};

template <typename T, int TI, int TJ, int MODE>
__global__ void __launch_bounds__((TI + 1) * (TJ + 1)) k_hex8(OpArgs<T> a) {
  constexpr int NT = (TI + 1) * (TJ + 1);
  constexpr int LW = TI + 2, LN = (TI + 2) * (TJ + 2);
  __shared__ T s_in[2][3][LN];
  __shared__ T s_out[24][NT];
  const Grid &g = a.g;
  if (a.st && a.st->done) return;  // solve already finished (uniform per launch)
  const int nblocks = a.tiles_x * a.tiles_y * a.tiles_z;
  const int t = xcd_tile(blockIdx.x, nblocks);
  const int ti = t % a.tiles_x, tj = (t / a.tiles_x) % a.tiles_y, tk = t / (a.tiles_x * a.tiles_y);
  const int i0 = ti * TI, j0 = tj * TJ, k0 = tk * a.TK;
  const int k1 = min(k0 + a.TK, g.sz);
  const int tid = threadIdx.x;
  const int tx = tid % (TI + 1), ty = tid / (TI + 1);
  const int ex = i0 - 1 + tx, ey = j0 - 1 + ty;
  const bool ex_ok = ex >= 0 && ex < g.nx && ey >= 0 && ey < g.ny;
  const bool node_thread = tid < TI * TJ;
  const int lx = tid % TI, ly = tid / TI;
  const int ni = i0 + lx, nj = j0 + ly;
  const bool node_ok = node_thread && ni < g.sx && nj < g.sy;
  const long long nn = g.nn;
  T beta = T(0), alpha_x = T(0);
  if (MODE == 1) { beta = (T)a.st->beta; alpha_x = (T)a.st->alpha_x; }

  auto load_layer = [&](int k, int buf) {
    for (int idx = tid; idx < LN; idx += NT) {
      const int li = idx % LW, lj = idx / LW;
      const int gi = i0 - 1 + li, gj = j0 - 1 + lj;
      const bool ok = gi >= 0 && gi < g.sx && gj >= 0 && gj < g.sy && k >= 0 && k < g.sz;
      const long long n = ok ? gi + (long long)g.sx * (gj + (long long)g.sy * k) : 0;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        T v = T(0);
        if (ok) {
          if (MODE == 0) v = a.in[c * nn + n];
          else v = a.invd[c * nn + n] * a.in[c * nn + n] + beta * a.pold[c * nn + n];
        }
        s_in[buf][c][idx] = v;
      }
    }
  };

  double dot = 0.0;
  T acc_top[3] = {T(0), T(0), T(0)};
  load_layer(k0 - 1, 0);
  for (int ez = k0 - 1; ez < k1; ++ez) {
    const int rel = ez - (k0 - 1);
    const int b0 = rel & 1, b1 = (rel + 1) & 1;
    load_layer(ez + 1, b1);
    __syncthreads();
    {
      T v[3][8];
      const bool ok = ex_ok && ez >= 0 && ez < g.nz;
      if (ok) {
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            const int qa = q & 1, qb = (q >> 1) & 1, qc = q >> 2;
            v[c][q] = s_in[qc ? b1 : b0][c][(tx + qa) + LW * (ty + qb)];
          }
        const T E = a.E[ex + (long long)g.nx * (ey + (long long)g.ny * ez)];
        elastic_hex8<T>(v, a.op, E);
      } else {
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
          for (int q = 0; q < 8; ++q) v[c][q] = T(0);
      }
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int q = 0; q < 8; ++q) s_out[c * 8 + q][tid] = v[c][q];
    }
    __syncthreads();
    if (node_thread) {
      T yb[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) { yb[c] = acc_top[c]; acc_top[c] = T(0); }
#pragma unroll
      for (int db = 0; db < 2; ++db)
#pragma unroll
        for (int da = 0; da < 2; ++da) {
          const int t2 = (lx + da) + (TI + 1) * (ly + db);
          const int q = (1 - da) + 2 * (1 - db);
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            yb[c] += s_out[c * 8 + q][t2];
            acc_top[c] += s_out[c * 8 + q + 4][t2];
          }
        }
      if (ez >= k0 && node_ok) {
        const long long n = ni + (long long)g.sx * (nj + (long long)g.sy * ez);
        const int ctr = (lx + 1) + LW * (ly + 1);
        const uint8_t m = a.mask ? a.mask[n] : 0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const T yv = ((m >> c) & 1) ? T(0) : yb[c];
          const T xv = s_in[b0][c][ctr];
          a.y[c * nn + n] = yv;
          if (MODE == 1) {
            a.pnew[c * nn + n] = xv;
            a.xacc[c * nn + n] += alpha_x * a.pold[c * nn + n];
          }
          dot += (double)xv * (double)yv;
        }
      }
    }
    __syncthreads();
  }
  if (a.red.partial) {
    double v[1] = {dot}, tot[1];
    if (grid_reduce<1>(a.red, v, tot) && threadIdx.x == 0) {
      SolverState *st = a.st;
      st->pAp = tot[0];
      st->alpha = st->absNew / tot[0];
      if (MODE == 1) st->alpha_x = st->alpha;
    }
  }
}

// quad4: a block owns TI nodes along x and TJ rows, marching along y.
template <typename T, int TI, int MODE>
__global__ void __launch_bounds__(TI + 1) k_quad4(OpArgs<T> a) {
  constexpr int NT = TI + 1;
  constexpr int LW = TI + 2;
  __shared__ T s_in[2][2][LW];
  __shared__ T s_out[8][NT];
  const Grid &g = a.g;
  if (a.st && a.st->done) return;  // solve already finished (uniform per launch)
  const int nblocks = a.tiles_x * a.tiles_y;
  const int t = xcd_tile(blockIdx.x, nblocks);
  const int ti = t % a.tiles_x, tj = t / a.tiles_x;
  const int i0 = ti * TI, j0 = tj * a.TK;
  const int j1 = min(j0 + a.TK, g.sy);
  const int tid = threadIdx.x;
  const int ex = i0 - 1 + tid;
  const bool ex_ok = ex >= 0 && ex < g.nx;
  const bool node_thread = tid < TI;
  const int ni = i0 + tid;
  const bool node_ok = node_thread && ni < g.sx;
  const long long nn = g.nn;
  T beta = T(0), alpha_x = T(0);
  if (MODE == 1) { beta = (T)a.st->beta; alpha_x = (T)a.st->alpha_x; }
  auto load_row = [&](int j, int buf) {
    for (int idx = tid; idx < LW; idx += NT) {
      const int gi = i0 - 1 + idx;
      const bool ok = gi >= 0 && gi < g.sx && j >= 0 && j < g.sy;
      const long long n = ok ? gi + (long long)g.sx * j : 0;
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        T v = T(0);
        if (ok) {
          if (MODE == 0) v = a.in[c * nn + n];
          else v = a.invd[c * nn + n] * a.in[c * nn + n] + beta * a.pold[c * nn + n];
        }
        s_in[buf][c][idx] = v;
      }
    }
  };
  double dot = 0.0;
  T acc_top[2] = {T(0), T(0)};
  load_row(j0 - 1, 0);
  for (int ey = j0 - 1; ey < j1; ++ey) {
    const int rel = ey - (j0 - 1);
    const int b0 = rel & 1, b1 = (rel + 1) & 1;
    load_row(ey + 1, b1);
    __syncthreads();
    {
      T v[2][4];
      const bool ok = ex_ok && ey >= 0 && ey < g.ny;
      if (ok) {
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
          for (int q = 0; q < 4; ++q) v[c][q] = s_in[(q >> 1) ? b1 : b0][c][tid + (q & 1)];
        const T E = a.E[ex + (long long)g.nx * ey];
        elastic_quad4<T>(v, a.op, E);
      } else {
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
          for (int q = 0; q < 4; ++q) v[c][q] = T(0);
      }
#pragma unroll
      for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int q = 0; q < 4; ++q) s_out[c * 4 + q][tid] = v[c][q];
    }
    __syncthreads();
    if (node_thread) {
      T yb[2];
#pragma unroll
      for (int c = 0; c < 2; ++c) { yb[c] = acc_top[c]; acc_top[c] = T(0); }
#pragma unroll
      for (int da = 0; da < 2; ++da) {
        const int t2 = tid + da;
        const int q = 1 - da;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          yb[c] += s_out[c * 4 + q][t2];
          acc_top[c] += s_out[c * 4 + q + 2][t2];
        }
      }
      if (ey >= j0 && node_ok) {
        const long long n = ni + (long long)g.sx * ey;
        const uint8_t m = a.mask ? a.mask[n] : 0;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          const T yv = ((m >> c) & 1) ? T(0) : yb[c];
          const T xv = s_in[b0][c][tid + 1];
          a.y[c * nn + n] = yv;
          if (MODE == 1) {
            a.pnew[c * nn + n] = xv;
            a.xacc[c * nn + n] += alpha_x * a.pold[c * nn + n];
          }
          dot += (double)xv * (double)yv;
        }
      }
    }
    __syncthreads();
  }
  if (a.red.partial) {
    double v[1] = {dot}, tot[1];
    if (grid_reduce<1>(a.red, v, tot) && threadIdx.x == 0) {
      SolverState *st = a.st;
      st->pAp = tot[0];
      st->alpha = st->absNew / tot[0];
      if (MODE == 1) st->alpha_x = st->alpha;
    }
  }
}

constexpr int HEX_TI = 15, HEX_TJ = 15;
constexpr int QUAD_TI = 127;

template <typename T>
static void setup_tiles(const Grid &g, OpArgs<T> &a, int &nblocks, int &nthreads) {
  if (g.ndim == 3) {
    a.tiles_x = (g.sx + HEX_TI - 1) / HEX_TI;
    a.tiles_y = (g.sy + HEX_TJ - 1) / HEX_TJ;
    // choose the z depth so the grid has enough blocks to fill 256 CUs
    const long long cols = (long long)a.tiles_x * a.tiles_y;
    int tk = 16;
    while (tk > 2 && cols * ((g.sz + tk - 1) / tk) < 1024) tk >>= 1;
    while (cols * ((g.sz + tk - 1) / tk) > RED_MAXBLOCKS) tk <<= 1;  // one partial per block
    a.TK = tk;
    a.tiles_z = (g.sz + tk - 1) / tk;
    nblocks = a.tiles_x * a.tiles_y * a.tiles_z;
    nthreads = (HEX_TI + 1) * (HEX_TJ + 1);
  } else {
    a.tiles_x = (g.sx + QUAD_TI - 1) / QUAD_TI;
    int tk = 16;
    while (tk > 2 && (long long)a.tiles_x * ((g.sy + tk - 1) / tk) < 512) tk >>= 1;
    while ((long long)a.tiles_x * ((g.sy + tk - 1) / tk) > RED_MAXBLOCKS) tk <<= 1;
    a.TK = tk;
    a.tiles_y = (g.sy + tk - 1) / tk;
    a.tiles_z = 1;
    nblocks = a.tiles_x * a.tiles_y;
    nthreads = QUAD_TI + 1;
  }
}

template <typename T>
void launch_elastic_matvec(const Grid &g, const ElasticOp<T> &op, const T *E, const uint8_t *mask,
                           const T *x, T *y, SolverState *st, Reduce red, hipStream_t s) {
  OpArgs<T> a{};
  a.g = g; a.op = op; a.E = E; a.mask = mask; a.in = x; a.y = y; a.st = st; a.red = red;
  int nb, nt;
  setup_tiles(g, a, nb, nt);
  if (g.ndim == 3) hipLaunchKernelGGL((k_hex8<T, HEX_TI, HEX_TJ, 0>), dim3(nb), dim3(nt), 0, s, a);
  else hipLaunchKernelGGL((k_quad4<T, QUAD_TI, 0>), dim3(nb), dim3(nt), 0, s, a);
  ICX_CHECK_LAUNCH();
}

template <typename T>
void launch_elastic_fusedA(const Grid &g, const ElasticOp<T> &op, const T *E, const uint8_t *mask,
                           const T *r, const T *invd, const T *pold, T *pnew, T *y, T *xacc,
                           SolverState *st, Reduce red, hipStream_t s) {
  OpArgs<T> a{};
  a.g = g; a.op = op; a.E = E; a.mask = mask; a.in = r; a.invd = invd; a.pold = pold;
  a.pnew = pnew; a.y = y; a.xacc = xacc; a.st = st; a.red = red;
  int nb, nt;
  setup_tiles(g, a, nb, nt);
  if (g.ndim == 3) hipLaunchKernelGGL((k_hex8<T, HEX_TI, HEX_TJ, 1>), dim3(nb), dim3(nt), 0, s, a);
  else hipLaunchKernelGGL((k_quad4<T, QUAD_TI, 1>), dim3(nb), dim3(nt), 0, s, a);
  ICX_CHECK_LAUNCH();
}

template void launch_elastic_matvec<double>(const Grid &, const ElasticOp<double> &, const double *,
                                            const uint8_t *, const double *, double *, SolverState *,
                                            Reduce, hipStream_t);
template void launch_elastic_matvec<float>(const Grid &, const ElasticOp<float> &, const float *,
                                           const uint8_t *, const float *, float *, SolverState *,
                                           Reduce, hipStream_t);
template void launch_elastic_fusedA<double>(const Grid &, const ElasticOp<double> &, const double *,
                                            const uint8_t *, const double *, const double *,
                                            const double *, double *, double *, double *,
                                            SolverState *, Reduce, hipStream_t);
template void launch_elastic_fusedA<float>(const Grid &, const ElasticOp<float> &, const float *,
                                           const uint8_t *, const float *, const float *,
                                           const float *, float *, float *, float *, SolverState *,
                                           Reduce, hipStream_t);

// ---------------------------------------------------------------------------------
// Helmholtz scalar operator: node-parallel gather over the (up to) 8 elements
// around each node; the element matrix is diagonal in the Haar basis.
// ---------------------------------------------------------------------------------
__global__ void k_helm_hex8(Grid g, HelmOp<double> op, const double *x, double *y,
                            SolverState *st, Reduce red) {
  if (st && red.partial && st->done) return;
  double dot = 0.0;
  for (long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x; n < g.nn;
       n += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(n % g.sx), j = (int)((n / g.sx) % g.sy), k = (int)(n / ((long long)g.sx * g.sy));
    double acc = 0.0;
#pragma unroll
    for (int dz = 0; dz < 2; ++dz)
#pragma unroll
      for (int dy = 0; dy < 2; ++dy)
#pragma unroll
        for (int dx = 0; dx < 2; ++dx) {
          const int ex = i - 1 + dx, ey = j - 1 + dy, ez = k - 1 + dz;
          if (ex < 0 || ey < 0 || ez < 0 || ex >= g.nx || ey >= g.ny || ez >= g.nz) continue;
          double v[8];
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            const long long m = (ex + (q & 1)) + (long long)g.sx * ((ey + ((q >> 1) & 1)) + (long long)g.sy * (ez + (q >> 2)));
            v[q] = x[m];
          }
          haar_fwd8(v);
#pragma unroll
          for (int q = 0; q < 8; ++q) v[q] *= op.kf[q];
          haar_inv8(v);
          const int q = (1 - dx) + 2 * (1 - dy) + 4 * (1 - dz);
          acc += v[q];
        }
    y[n] = acc;
    dot += x[n] * acc;
  }
  if (red.partial) {
    double v[1] = {dot}, tot[1];
    if (grid_reduce<1>(red, v, tot) && threadIdx.x == 0) {
      st->pAp = tot[0];
      st->alpha = st->absNew / tot[0];
    }
  }
}

__global__ void k_helm_quad4(Grid g, HelmOp<double> op, const double *x, double *y,
                             SolverState *st, Reduce red) {
  if (st && red.partial && st->done) return;
  double dot = 0.0;
  for (long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x; n < g.nn;
       n += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(n % g.sx), j = (int)(n / g.sx);
    double acc = 0.0;
#pragma unroll
    for (int dy = 0; dy < 2; ++dy)
#pragma unroll
      for (int dx = 0; dx < 2; ++dx) {
        const int ex = i - 1 + dx, ey = j - 1 + dy;
        if (ex < 0 || ey < 0 || ex >= g.nx || ey >= g.ny) continue;
        double v[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) v[q] = x[(ex + (q & 1)) + (long long)g.sx * (ey + (q >> 1))];
        haar_fwd4(v);
#pragma unroll
        for (int q = 0; q < 4; ++q) v[q] *= op.kf[q];
        haar_inv4(v);
        acc += v[(1 - dx) + 2 * (1 - dy)];
      }
    y[n] = acc;
    dot += x[n] * acc;
  }
  if (red.partial) {
    double v[1] = {dot}, tot[1];
    if (grid_reduce<1>(red, v, tot) && threadIdx.x == 0) {
      st->pAp = tot[0];
      st->alpha = st->absNew / tot[0];
    }
  }
}

void launch_helm_matvec(const Grid &g, const HelmOp<double> &op, const double *x, double *y,
                        SolverState *st, Reduce red, hipStream_t s) {
  const int nb = ew_blocks(g.nn);
  if (g.ndim == 3) hipLaunchKernelGGL(k_helm_hex8, dim3(nb), dim3(256), 0, s, g, op, x, y, st, red);
  else hipLaunchKernelGGL(k_helm_quad4, dim3(nb), dim3(256), 0, s, g, op, x, y, st, red);
  ICX_CHECK_LAUNCH();
}

// ---------------------------------------------------------------------------------
// PCG vector kernels
// ---------------------------------------------------------------------------------
// Eigen loop, after tmp = A p and alpha: x += alpha p; r -= alpha tmp; rr = r.r;
// z = invd r; rz = r.z.  Last block: convergence test, beta, absNew, counter.
template <typename T>
__global__ void k_pcg_update(long long n, SolverState *st, T *x, T *r, const T *p, const T *t, T *z,
                             const T *invd, Reduce red) {
  if (st->done) return;
  const double alpha = st->alpha;
  double rr = 0.0, rz = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    T xi = x[i], ri = r[i];
    xi += (T)alpha * p[i];
    ri -= (T)alpha * t[i];
    x[i] = xi;
    r[i] = ri;
    const T zi = invd[i] * ri;
    z[i] = zi;
    rr += (double)ri * (double)ri;
    rz += (double)ri * (double)zi;
  }
  double v[2] = {rr, rz}, tot[2];
  if (grid_reduce<2>(red, v, tot) && threadIdx.x == 0) {
    st->rr = tot[0];
    st->rz_new = tot[1];
    if (tot[0] < st->threshold) {
      st->done = 1;
    } else {
      st->absOld = st->absNew;
      st->absNew = tot[1];
      st->beta = tot[1] / st->absOld;
      st->iter += 1;
      if (st->iter >= st->maxit) st->done = 2;
    }
  }
}

template <typename T>
__global__ void k_pcg_pupdate(long long n, SolverState *st, T *p, const T *z);

// ---- the Eigen ConjugateGradient loop, one kernel per vector expression ----------
// (Eigen 3.4 ConjugateGradient.h as called by src/analysis/solver/solvers.cpp:64-92)
// alpha = absNew / p.dot(tmp)
__global__ void k_eig_dot_pt(long long n, SolverState *st, const double *p, const double *t, Reduce red) {
  if (st->done) return;
  double s = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    s += p[i] * t[i];
  double v[1] = {s}, tot[1];
  if (grid_reduce<1>(red, v, tot) && threadIdx.x == 0) {
    st->pAp = tot[0];
    st->alpha = st->absNew / tot[0];
  }
}
// y += sign * alpha * v   (x += alpha p; residual -= alpha tmp)
__global__ void k_eig_axpy(long long n, SolverState *st, double sign, double *y, const double *v) {
  if (st->done) return;
  const double a = sign * st->alpha;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    y[i] += a * v[i];
}
// residualNorm2 = residual.squaredNorm(); break if below threshold
__global__ void k_eig_norm(long long n, SolverState *st, const double *r, Reduce red) {
  if (st->done) return;
  double s = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    s += r[i] * r[i];
  double v[1] = {s}, tot[1];
  if (grid_reduce<1>(red, v, tot) && threadIdx.x == 0) {
    st->rr = tot[0];
    if (tot[0] < st->threshold) st->done = 1;
  }
}
// z = precond.solve(residual)
__global__ void k_eig_precond(long long n, SolverState *st, double *z, const double *invd, const double *r) {
  if (st->done) return;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    z[i] = invd[i] * r[i];
}
// absOld = absNew; absNew = residual.dot(z); beta = absNew / absOld; i++
__global__ void k_eig_dot_rz(long long n, SolverState *st, const double *r, const double *z, Reduce red) {
  if (st->done) return;
  double s = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    s += r[i] * z[i];
  double v[1] = {s}, tot[1];
  if (grid_reduce<1>(red, v, tot) && threadIdx.x == 0) {
    st->absOld = st->absNew;
    st->absNew = tot[0];
    st->beta = tot[0] / st->absOld;
    st->iter += 1;
    if (st->iter >= st->maxit) st->done = 2;
  }
}

void launch_eigen_iteration_tail(long long n, SolverState *st, double *x, double *r, double *p,
                                 const double *t, double *z, const double *invd, Reduce red,
                                 hipStream_t s) {
  const int nb = ew_blocks(n, 4);
  hipLaunchKernelGGL(k_eig_dot_pt, dim3(nb), dim3(256), 0, s, n, st, p, t, red);
  hipLaunchKernelGGL(k_eig_axpy, dim3(nb), dim3(256), 0, s, n, st, 1.0, x, p);
  hipLaunchKernelGGL(k_eig_axpy, dim3(nb), dim3(256), 0, s, n, st, -1.0, r, t);
  hipLaunchKernelGGL(k_eig_norm, dim3(nb), dim3(256), 0, s, n, st, r, red);
  hipLaunchKernelGGL(k_eig_precond, dim3(nb), dim3(256), 0, s, n, st, z, invd, r);
  hipLaunchKernelGGL(k_eig_dot_rz, dim3(nb), dim3(256), 0, s, n, st, r, z, red);
  hipLaunchKernelGGL(k_pcg_pupdate<double>, dim3(nb), dim3(256), 0, s, n, st, p, z);
  ICX_CHECK_LAUNCH();
}

template <typename T>
__global__ void k_pcg_pupdate(long long n, SolverState *st, T *p, const T *z) {
  if (st->done) return;
  const T beta = (T)st->beta;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = z[i] + beta * p[i];
}

// fused B: r -= alpha Ap; rr, rz (z = invd r recomputed by the next kernel A)
template <typename T>
__global__ void k_pcg_fusedB(long long n, SolverState *st, T *r, const T *Ap, const T *invd,
                             Reduce red) {
  if (st->done) return;
  const T alpha = (T)st->alpha;
  double rr = 0.0, rz = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    T ri = r[i];
    ri -= alpha * Ap[i];
    r[i] = ri;
    rr += (double)ri * (double)ri;
    rz += (double)ri * (double)(invd[i] * ri);
  }
  double v[2] = {rr, rz}, tot[2];
  if (grid_reduce<2>(red, v, tot) && threadIdx.x == 0) {
    st->rr = tot[0];
    st->rz_new = tot[1];
    st->absOld = st->absNew;
    st->absNew = tot[1];
    st->beta = tot[1] / st->absOld;
    st->iter += 1;
    st->inner_since += 1;
    if (tot[0] < st->threshold) st->done = 3;
    else if (tot[0] < st->delta2 * st->rr_ref) st->done = 4;
    else if (st->replace_every > 0 && st->inner_since >= st->replace_every) st->done = 4;
    if (st->iter >= st->maxit && st->done == 0) st->done = 2;
  }
}

template <typename T>
__global__ void k_flush_x(long long n, SolverState *st, T *x, const T *p) {
  const T a = (T)st->alpha_x;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    x[i] += a * p[i];
}

__global__ void k_refine_x(long long n, SolverState *st, double *xhi, float *xlo, const float *p) {
  const float a = (float)st->alpha_x;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const float xl = xlo[i] + a * p[i];
    xhi[i] += (double)xl;
    xlo[i] = 0.0f;
  }
}

__global__ void k_refine_r(long long n, SolverState *st, const double *b, const double *Kx, float *r32,
                           const float *invd32, Reduce red) {
  double rr = 0.0, rz = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double ri = b[i] - Kx[i];
    const float rf = (float)ri;
    r32[i] = rf;
    rr += ri * ri;
    rz += (double)rf * (double)(invd32[i] * rf);
  }
  double v[2] = {rr, rz}, tot[2];
  if (grid_reduce<2>(red, v, tot) && threadIdx.x == 0) {
    st->true_rr = tot[0];
    st->rr = tot[0];
    st->rr_ref = tot[0];
    st->alpha_x = 0.0;
    st->inner_since = 0;
    // continue the recurrence with the replaced residual
    st->beta = tot[1] / st->absOld;
    st->absNew = tot[1];
    st->done = (tot[0] < st->threshold) ? 1 : 0;
  }
}

template <typename T>
__global__ void k_pcg_init(long long n, SolverState *st, const T *b, T *r, T *p, const T *invd,
                           Reduce red, double tol, int maxit, int set_p) {
  double bb = 0.0, rz = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const T bi = b[i];
    r[i] = bi;
    const T zi = invd[i] * bi;
    if (set_p) p[i] = zi;
    bb += (double)bi * (double)bi;
    rz += (double)bi * (double)zi;
  }
  double v[2] = {bb, rz}, tot[2];
  if (grid_reduce<2>(red, v, tot) && threadIdx.x == 0) {
    st->rhsNorm2 = tot[0];
    double thr = tol * tol * tot[0];
    if (thr < 2.2250738585072014e-308) thr = 2.2250738585072014e-308;
    st->threshold = thr;
    st->rr = tot[0];
    st->rr_ref = tot[0];
    st->absNew = tot[1];
    st->absOld = tot[1];
    st->beta = 0.0;
    st->alpha = 0.0;
    st->alpha_x = 0.0;
    st->iter = 0;
    st->inner_since = 0;
    st->maxit = maxit;
    st->true_rr = tot[0];
    // Eigen: a zero rhs returns x = 0 at once; so does a converged start
    st->done = (tot[0] == 0.0 || tot[0] < thr) ? 1 : 0;
  }
}

template <typename T>
void launch_pcg_update(long long n, SolverState *st, T *x, T *r, const T *p, const T *t, T *z,
                       const T *invd, Reduce red, hipStream_t s) {
  hipLaunchKernelGGL(k_pcg_update<T>, dim3(ew_blocks(n, 4)), dim3(256), 0, s, n, st, x, r, p, t, z,
                     invd, red);
  ICX_CHECK_LAUNCH();
}
template <typename T>
void launch_pcg_pupdate(long long n, SolverState *st, T *p, const T *z, hipStream_t s) {
  hipLaunchKernelGGL(k_pcg_pupdate<T>, dim3(ew_blocks(n, 4)), dim3(256), 0, s, n, st, p, z);
  ICX_CHECK_LAUNCH();
}
template <typename T>
void launch_pcg_fusedB(long long n, SolverState *st, T *r, const T *Ap, const T *invd, Reduce red,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_pcg_fusedB<T>, dim3(ew_blocks(n, 4)), dim3(256), 0, s, n, st, r, Ap, invd, red);
  ICX_CHECK_LAUNCH();
}
template <typename T>
void launch_flush_x(long long n, SolverState *st, T *x, const T *p, hipStream_t s) {
  hipLaunchKernelGGL(k_flush_x<T>, dim3(ew_blocks(n, 4)), dim3(256), 0, s, n, st, x, p);
  ICX_CHECK_LAUNCH();
}
void launch_refine_x(long long n, SolverState *st, double *xhi, float *xlo, const float *p,
                     hipStream_t s) {
  hipLaunchKernelGGL(k_refine_x, dim3(ew_blocks(n, 4)), dim3(256), 0, s, n, st, xhi, xlo, p);
  ICX_CHECK_LAUNCH();
}
void launch_refine_r(long long n, SolverState *st, const double *b, const double *Kx, float *r32,
                     const float *invd32, Reduce red, hipStream_t s) {
  hipLaunchKernelGGL(k_refine_r, dim3(ew_blocks(n, 4)), dim3(256), 0, s, n, st, b, Kx, r32, invd32, red);
  ICX_CHECK_LAUNCH();
}
template <typename T>
void launch_pcg_init(long long n, SolverState *st, const T *b, T *r, T *p, const T *invd,
                     Reduce red, double tol, int maxit, int set_p, hipStream_t s) {
  hipLaunchKernelGGL(k_pcg_init<T>, dim3(ew_blocks(n, 4)), dim3(256), 0, s, n, st, b, r, p, invd, red,
                     tol, maxit, set_p);
  ICX_CHECK_LAUNCH();
}

#define ICX_INST_VEC(T)                                                                          \
  template void launch_pcg_update<T>(long long, SolverState *, T *, T *, const T *, const T *, T *, \
                                     const T *, Reduce, hipStream_t);                           \
  template void launch_pcg_pupdate<T>(long long, SolverState *, T *, const T *, hipStream_t);     \
  template void launch_pcg_fusedB<T>(long long, SolverState *, T *, const T *, const T *, Reduce,  \
                                     hipStream_t);                                               \
  template void launch_flush_x<T>(long long, SolverState *, T *, const T *, hipStream_t);         \
  template void launch_pcg_init<T>(long long, SolverState *, const T *, T *, T *, const T *,       \
                                   Reduce, double, int, int, hipStream_t);
ICX_INST_VEC(double)
ICX_INST_VEC(float)

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

__global__ void k_dot(long long n, const double *a, const double *b, double *out, Reduce red) {
  double s = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    s += a[i] * b[i];
  double v[1] = {s}, tot[1];
  if (grid_reduce<1>(red, v, tot) && threadIdx.x == 0) out[0] = tot[0];
}
void launch_dot(long long n, const double *a, const double *b, double *out, Reduce red, hipStream_t s) {
  hipLaunchKernelGGL(k_dot, dim3(ew_blocks(n, 4)), dim3(256), 0, s, n, a, b, out, red);
  ICX_CHECK_LAUNCH();
}

// ---------------------------------------------------------------------------------
// sensitivity (opt_macro_oc.cpp:182-212, opt_mma_Helmholtz.cpp:173-205)
// ---------------------------------------------------------------------------------
__device__ __forceinline__ double flush15(double v) { return fabs(v) < 1.0e-15 ? 0.0 : v; }

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
__global__ void k_oc_candidate(long long n, double lambda, double move, double xmin, double xmax,
                               const double *x, const double *dfdx, const double *dgdx, double *xnew) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    xnew[i] = oc_x(lambda, move, xmin, xmax, x[i], dfdx[i], dgdx[i]);
}
void launch_oc_candidate(long long n, double lambda, double move, double xmin, double xmax,
                         const double *x, const double *dfdx, const double *dgdx, double *xnew,
                         hipStream_t s) {
  hipLaunchKernelGGL(k_oc_candidate, dim3(ew_blocks(n)), dim3(256), 0, s, n, lambda, move, xmin, xmax, x,
                     dfdx, dgdx, xnew);
  ICX_CHECK_LAUNCH();
}

__global__ void k_volume(long long n, int project, double beta, double T, const double *rho,
                         const double *volv, double vol, double *out, Reduce red) {
  double s = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double h = project ? proj(rho[i], beta, T) : rho[i];
    s += h * (volv ? volv[i] : vol);
  }
  double v[1] = {s}, tot[1];
  if (grid_reduce<1>(red, v, tot) && threadIdx.x == 0) out[0] = tot[0];
}
void launch_volume(long long n, int project, double beta, double T, const double *rho,
                   const double *vol_or_null, double vol, double *out, Reduce red, hipStream_t s) {
  hipLaunchKernelGGL(k_volume, dim3(ew_blocks(n, 4)), dim3(256), 0, s, n, project, beta, T, rho,
                     vol_or_null, vol, out, red);
  ICX_CHECK_LAUNCH();
}

// one bisection step of the OC volume lambda with the cone filter, fused:
// sum_i threshold( sum_j w_ij x_new_j(lambda) / aa_i ) v
__global__ void k_oc_volume_cone(Grid g, Stencil st, const double *aa, double lambda, double move,
                                 double xmin, double xmax, const double *x, const double *dfdx,
                                 const double *dgdx, int project, double beta, double T, double vol,
                                 double *out, Reduce red) {
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
  double v[1] = {s}, tot[1];
  if (grid_reduce<1>(red, v, tot) && threadIdx.x == 0) out[0] = tot[0];
}
void launch_oc_volume_cone(const Grid &g, Stencil st, const double *aa, double lambda, double move,
                           double xmin, double xmax, const double *x, const double *dfdx,
                           const double *dgdx, int project, double beta, double T, double vol,
                           double *out, Reduce red, hipStream_t s) {
  hipLaunchKernelGGL(k_oc_volume_cone, dim3(ew_blocks(g.ne)), dim3(256), 0, s, g, st, aa, lambda, move,
                     xmin, xmax, x, dfdx, dgdx, project, beta, T, vol, out, red);
  ICX_CHECK_LAUNCH();
}

// ---------------------------------------------------------------------------------
// MMA (m = 1), src/optimization/opt-solver/MMASolver.cpp:400-512 (GenSub) and
// 184-398 (XYZofLAMBDA, DualGrad, DualHess, DualResidual sums)
// ---------------------------------------------------------------------------------
__global__ void k_mma_gensub(long long n, int iter, MMAVecs m, const double *xval, const double *dfdx,
                             const double *dgdx, double xmin, double xmax, double move, double albefa,
                             double asyminit, double asymdec, double asyminc, double raa0,
                             double xmamieps, double *bsum, Reduce red) {
  double s = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double xv = xval[i];
    double low, upp;
    const double xo1 = m.xold1[i], xo2 = m.xold2[i];
    if (iter < 3) {
      low = xv - asyminit * (xmax - xmin);
      upp = xv + asyminit * (xmax - xmin);
    } else {
      low = m.low[i]; upp = m.upp[i];
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
    m.low[i] = low; m.upp[i] = upp;
    double al = fmax(xmin, low + albefa * (xv - low));
    al = fmax(al, xv - move * (xmax - xmin));
    m.alpha[i] = fmin(al, xmax);
    double be = fmin(xmax, upp - albefa * (upp - xv));
    be = fmin(be, xv + move * (xmax - xmin));
    m.beta[i] = fmax(be, xmin);
    const double xmamiinv = 1.0 / fmax(xmamieps, xmax - xmin);
    const double df = dfdx[i];
    const double pq = 0.001 * fabs(df) + raa0 * xmamiinv;
    const double u2 = pow(upp - xv, 2.0), l2 = pow(xv - low, 2.0);
    m.p0[i] = u2 * (fmax(0.0, df) + pq);
    m.q0[i] = l2 * (fmax(0.0, -1.0 * df) + pq);
    const double dg = dgdx[i];
    const double pqj = 0.001 * fabs(dg) + raa0 * xmamiinv;
    const double p1 = u2 * (fmax(0.0, dg) + pqj), q1 = l2 * (fmax(0.0, -1.0 * dg) + pqj);
    m.p1[i] = p1; m.q1[i] = q1;
    s += p1 / (upp - xv) + q1 / (xv - low);
    m.xold2[i] = xo1;
    m.xold1[i] = xv;
  }
  double v[1] = {s}, tot[1];
  if (grid_reduce<1>(red, v, tot) && threadIdx.x == 0) bsum[0] = tot[0];
}
void launch_mma_gensub(long long n, int iter, const MMAVecs &m, const double *xval,
                       const double *dfdx, const double *dgdx, double xmin, double xmax,
                       double move, double albefa, double asyminit, double asymdec, double asyminc,
                       double raa0, double xmamieps, double *bsum, Reduce red, hipStream_t s) {
  hipLaunchKernelGGL(k_mma_gensub, dim3(ew_blocks(n, 4)), dim3(256), 0, s, n, iter, m, xval, dfdx, dgdx,
                     xmin, xmax, move, albefa, asyminit, asymdec, asyminc, raa0, xmamieps, bsum, red);
  ICX_CHECK_LAUNCH();
}

__global__ void k_mma_dual(long long n, double lam, MMAVecs m, double *x, int want_hess, double *out,
                           Reduce red) {
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
    if (want_hess) {
      const double PQ = m.p1[i] / pow(upp - xi, 2.0) - m.q1[i] / pow(xi - low, 2.0);
      double df2 = -1.0 / (2.0 * pj / pow(upp - xi, 3.0) + 2.0 * qj / pow(xi - low, 3.0));
      if (xp < m.alpha[i]) df2 = 0.0;
      if (xp > m.beta[i]) df2 = 0.0;
      s2 += (0.0 + PQ * df2) * PQ;
    }
  }
  double v[2] = {s1, s2}, tot[2];
  if (grid_reduce<2>(red, v, tot) && threadIdx.x == 0) { out[0] = tot[0]; out[1] = tot[1]; }
}
void launch_mma_dual(long long n, double lam, const MMAVecs &m, double *x, int want_hess, double *out,
                     Reduce red, hipStream_t s) {
  hipLaunchKernelGGL(k_mma_dual, dim3(ew_blocks(n, 4)), dim3(256), 0, s, n, lam, m, x, want_hess, out, red);
  ICX_CHECK_LAUNCH();
}

}  // namespace icx
