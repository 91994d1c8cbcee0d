// Device helpers shared by the libicx kernels (gfx950, wave64).
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "icx_internal.h"

namespace icx {

#define ICX_CHECK_LAUNCH() (void)hipGetLastError()

// Launch with an optional event pair taking the kernel's own start / end timestamps.
template <typename F, typename... Args>
inline void launch_ev(F kernel, int blocks, int threads, hipStream_t s, const LaunchEvents *ev, Args... args) {
  if (ev) hipExtLaunchKernelGGL(kernel, dim3(blocks), dim3(threads), 0, s, ev->start, ev->stop, 0, args...);
  else hipLaunchKernelGGL(kernel, dim3(blocks), dim3(threads), 0, s, args...);
}

// s_waitcnt vmcnt(0): every vector-memory load and store of this wave has completed
// (gfx9-family encoding; the expcnt and lgkmcnt fields are at their maxima, i.e. not
// waited on).  Placed where a pipelined kernel's previous-step traffic is due anyway,
// it also tells the compiler's wait insertion that nothing older is pending.
__device__ __forceinline__ void wait_vmem() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// Explicit fused multiply-add.  Expressions with two products (a*b + c*d) are written
// with it, so that every instantiation of a kernel rounds them identically (with
// contraction on, the compiler may otherwise fuse either product).
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Value held by lane + 1 of the same 16-lane row (DPP row_shl:1); the last lane of
// each row reads 0.
__device__ __forceinline__ float row_next(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x101, 0xF, 0xF, true));
}
__device__ __forceinline__ double row_next(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xFFFFFFFF), 0x101, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x101, 0xF, 0xF, true);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}

// load / store at a 32-bit byte offset from a uniform base (global_load with an SGPR
// base and a 32-bit VGPR offset; node vectors stay below 4 GiB)
template <typename T> __device__ __forceinline__ T ld_off(const T *base, uint32_t off) {
  return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + off);
}
template <typename T> __device__ __forceinline__ void st_off(T *base, uint32_t off, T v) {
  *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + off) = v;
}

// |v| < 1e-15 flushed to 0, as divideValIntoNode does before the sensitivity reads
// node values (src/analysis/analysis_utility.hpp:27-28)
__device__ __forceinline__ double flush15(double v) { return fabs(v) < 1.0e-15 ? 0.0 : v; }

// Snapshot of a solver state at kernel entry (scalar loads: one round trip).  Every
// field an iteration kernel reads is written only by other kernels (kernel A records a
// stop in `pending`, kernel B publishes it as `done`), so all waves of all blocks of a
// launch read the same values.
__device__ __forceinline__ SolverState load_state(const SolverState *st) { return *st; }

// Job and job-local block of a batched launch (see Batch in icx_internal.h); the
// search is uniform per block.
__device__ __forceinline__ void batch_block(const int *__restrict__ off, int njobs, int &j, int &blk) {
  const int b = blockIdx.x;
  int lo = 0, hi = njobs;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= b) lo = mid;
    else hi = mid;
  }
  j = lo;
  blk = b - off[lo];
}

// ---------------------------------------------------------------------------------
// reductions
//
// Kernels never finish a reduction themselves: every block writes one partial per
// reduced quantity (plain store) and the kernel that needs the total -- the next
// one on the stream -- sums the partials in index order.  Kernel boundaries make
// the partials visible; no device-scope atomics or fences are involved (on an
// 8-XCD part a per-block fence + atomic ticket costs ~20 ns per block, serialised),
// and every block of the consumer computes the same total bit for bit.
// ---------------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Sum of one value per thread over the block; the total is returned to every
// thread.  scratch: LDS of >= BS/64 doubles.  Contains two barriers.
template <int BS>
__device__ __forceinline__ double block_sum(double v, double *scratch) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) scratch[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int w = 0; w < BS / 64; ++w) t += scratch[w];
  __syncthreads();
  return t;
}

// Total of p[0], p[s], ..., p[(n-1)s] in a fixed order, returned to every thread.
template <int BS>
__device__ __forceinline__ double block_total(const double *p, int n, int s, double *scratch) {
  double a = 0.0;
  if (n <= 16 * BS) {
    // every load issued before the first add (one memory latency instead of one per
    // round of a partially unrolled loop); the same order of additions
    double v[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int q = threadIdx.x + k * BS;
      v[k] = q < n ? p[(long long)q * s] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) a += v[k];
  } else {
#pragma unroll 4
    for (int q = threadIdx.x; q < n; q += BS) a += p[(long long)q * s];
  }
  return block_sum<BS>(a, scratch);
}

// block-count policy of the streaming kernels (each partial array holds <= RED_MAXBLOCKS)
static inline int ew_blocks(long long n, int per_thread = 1, int cap = 2048) {
  long long b = (n + 256LL * per_thread - 1) / (256LL * per_thread);
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (int)b;
}
constexpr int RED_BLOCKS = 512;   // grid of the reducing streaming kernels

// ---------------------------------------------------------------------------------
// element operators in the per-element Haar basis
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
// E * haar_inv8(v), the scale folded into the first stage: two products per pair
// become one product and two fused multiply-adds
template <typename T>
__device__ __forceinline__ void haar_inv8_scaled(T (&v)[8], T E) {
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const T t = E * v[q + 4];
    const T a = v[q];
    v[q] = fma_t(E, a, -t);
    v[q + 4] = fma_t(E, a, t);
  }
#pragma unroll
  for (int q : {0, 1, 4, 5}) { T a = v[q], b = v[q + 2]; v[q] = a - b; v[q + 2] = a + b; }
#pragma unroll
  for (int q = 0; q < 8; q += 2) { T a = v[q], b = v[q + 1]; v[q] = a - b; v[q + 1] = a + b; }
}
template <typename T>
__device__ __forceinline__ void haar_inv4_scaled(T (&v)[4], T E) {
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const T t = E * v[q + 2];
    const T a = v[q];
    v[q] = fma_t(E, a, -t);
    v[q + 2] = fma_t(E, a, t);
  }
#pragma unroll
  for (int q = 0; q < 4; q += 2) { T a = v[q], b = v[q + 1]; v[q] = a - b; v[q + 1] = a + b; }
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
__host__ __device__ constexpr int first_odd(int chi) { return (chi & 1) ? 0 : ((chi & 2) ? 1 : 2); }
__host__ __device__ constexpr int first_even(int chi) { return !(chi & 1) ? 0 : (!(chi & 2) ? 1 : 2); }
__host__ __device__ constexpr bool blk_nz3(int chi, int c, int cc) {
  return (popc3(chi) == 0 || popc3(chi) == 3)
             ? true
             : (popc3(chi) == 1 ? (c != first_odd(chi) && cc != first_odd(chi))
                                : ((c == first_even(chi)) == (cc == first_even(chi))));
}
__host__ __device__ constexpr bool blk_nz2(int chi, int c, int cc) {
  return (chi == 0 || chi == 3) ? true : (c != ((chi & 1) ? 0 : 1) && cc != ((chi & 1) ? 0 : 1));
}

// o <- the character blocks of K0 applied to the Haar coefficients v of a hexa8 element
template <typename T>
__device__ __forceinline__ void haar_blocks8(const T (&v)[3][8], const ElasticOp<T> &op, T (&o)[3][8]) {
#pragma unroll
  for (int chi = 0; chi < 8; ++chi) {
    const T m[3] = {v[0][chi ^ 1], v[1][chi ^ 2], v[2][chi ^ 4]};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      T acc = T(0);
      bool first = true;
#pragma unroll
      for (int cc = 0; cc < 3; ++cc)
        if (blk_nz3(chi, c, cc)) {
          const T k = op.blk[chi][c < cc ? c : cc][c < cc ? cc : c];
          acc = first ? k * m[cc] : fma_t(k, m[cc], acc);
          first = false;
        }
      o[c][chi ^ (1 << c)] = acc;
    }
  }
}

// v <- E * K0 v for one hexa8 element (v[c][q], corner q = a + 2b + 4c)
template <typename T>
__device__ __forceinline__ void elastic_hex8(T (&v)[3][8], const ElasticOp<T> &op, T E) {
  haar_fwd8(v[0]); haar_fwd8(v[1]); haar_fwd8(v[2]);
  T o[3][8];
  haar_blocks8(v, op, o);
  haar_inv8_scaled(o[0], E); haar_inv8_scaled(o[1], E); haar_inv8_scaled(o[2], E);
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
      bool first = true;
#pragma unroll
      for (int cc = 0; cc < 2; ++cc)
        if (blk_nz2(chi, c, cc)) {
          const T k = op.blk[chi][c < cc ? c : cc][c < cc ? cc : c];
          acc = first ? k * m[cc] : fma_t(k, m[cc], acc);
          first = false;
        }
      o[c][chi ^ (1 << c)] = acc;
    }
  }
  haar_inv4_scaled(o[0], E); haar_inv4_scaled(o[1], E);
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
        if (blk_nz3(chi, c, cc)) acc += op.blk[chi][c < cc ? c : cc][c < cc ? cc : c] * m[cc];
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
        if (blk_nz2(chi, c, cc)) acc += op.blk[chi][c < cc ? c : cc][c < cc ? cc : c] * m[cc];
      e += m[c] * acc;
    }
  }
  return e;
}

// u^T K0 w of one quad4 element (K0 symmetric, block diagonal in the Haar basis)
template <typename T>
__device__ __forceinline__ T bilinear_quad4(T (&u)[2][4], T (&w)[2][4], const ElasticOp<T> &op) {
  haar_fwd4(u[0]); haar_fwd4(u[1]);
  haar_fwd4(w[0]); haar_fwd4(w[1]);
  T e = T(0);
#pragma unroll
  for (int chi = 0; chi < 4; ++chi) {
    const T mu[2] = {u[0][chi ^ 1], u[1][chi ^ 2]}, mw[2] = {w[0][chi ^ 1], w[1][chi ^ 2]};
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      T acc = T(0);
#pragma unroll
      for (int cc = 0; cc < 2; ++cc)
        if (blk_nz2(chi, c, cc)) acc += op.blk[chi][c < cc ? c : cc][c < cc ? cc : c] * mw[cc];
      e += mu[c] * acc;
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

// Totals of the pairs (p[2q], p[2q+1]), q < n, in a fixed order, returned to every
// thread.  One round of 16-byte loads and a single barrier; scratch: 2 BS/64 doubles
// used by this call only.
template <int BS>
__device__ __forceinline__ void block_total2(const double *p, int n, double *scratch, double &s0, double &s1) {
  double a = 0.0, b = 0.0;
#pragma unroll 4
  for (int q = threadIdx.x; q < n; q += BS) {
    const double2 v = reinterpret_cast<const double2 *>(p)[q];
    a += v.x;
    b += v.y;
  }
  a = wave_sum(a);
  b = wave_sum(b);
  if ((threadIdx.x & 63) == 0) {
    scratch[threadIdx.x >> 6] = a;
    scratch[BS / 64 + (threadIdx.x >> 6)] = b;
  }
  __syncthreads();
  s0 = s1 = 0.0;
#pragma unroll
  for (int w = 0; w < BS / 64; ++w) {
    s0 += scratch[w];
    s1 += scratch[BS / 64 + w];
  }
}

// ---------------------------------------------------------------------------------
// scalars of the fused / mixed PCG iteration i.  Kernel A of iteration i sums the
// (|r_i|^2, r_i.z_i) partials of kernel B of iteration i-1 and stores r_i.z_i in
// SolverState::rz_hist[i & 1]; kernel B of iteration i stores alpha_i in
// alpha_hist[i & 1].  Each total is formed once, so beta_i and the deferred step
// alpha_{i-1} are bit-identical wherever they are used.
// ---------------------------------------------------------------------------------
struct IterScalars {
  double rr, rz, beta, alpha_prev;
  int stop;  // 0 continue, 2 maxit, 3 tolerance met by the recursive residual, 4 replace
};

// st: a snapshot of the solver state taken at kernel entry; rr, rz: the totals of the
// (|r_i|^2, r_i.z_i) partials (block_total2 over P.PB, issued before the state arrives).
__device__ __forceinline__ IterScalars iter_scalars(const SolverState &st, int i, double rr, double rz) {
  IterScalars s;
  s.rr = rr;
  s.rz = rz;
  if (i > 0) {
    const bool odd = (i - 1) & 1;
    s.beta = s.rz / (odd ? st.rz_hist[1] : st.rz_hist[0]);
    s.alpha_prev = odd ? st.alpha_hist[1] : st.alpha_hist[0];
  } else {
    s.beta = 0.0;
    s.alpha_prev = 0.0;
  }
  s.stop = 0;
  if (s.rr < st.threshold) s.stop = 3;
  else if (st.delta2 > 0.0 && s.rr < st.delta2 * st.rr_ref) s.stop = 4;
  else if (i >= st.maxit) s.stop = 2;
  return s;
}

}  // namespace icx
