// PCG vector kernels of libicx (gfx950).  Every reducing kernel writes one partial
// per block; the totals are formed by the kernels that consume them (see
// icx_device.h), indexed by the parity of the iteration that produced them.
#include "icx_device.h"


namespace icx {

constexpr int BS = 256;

// (|r|^2, r.z) partials of the fused / mixed forms: written by kernel B (or the set-up
// or a residual replacement), read by the next kernel A only, so one slot suffices
__host__ __device__ static inline double *pb_slot(const PcgPartials &P) { return P.PB; }

int pcg_stream_blocks(long long n) { return ew_blocks(n, 4, RED_BLOCKS); }

// ---- fused / mixed forms --------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(BS) k_pcg_setup(long long n, const T *b, T *r, const T *invd, PcgPartials P) {
  __shared__ double scr[BS / 64];
  double bb = 0.0, rz = 0.0;
  for (long long i = blockIdx.x * (long long)BS + threadIdx.x; i < n; i += (long long)gridDim.x * BS) {
    const T bi = b[i];
    r[i] = bi;
    bb += (double)bi * (double)bi;
    rz += (double)bi * (double)(invd[i] * bi);
  }
  bb = block_sum<BS>(bb, scr);
  rz = block_sum<BS>(rz, scr);
  if (threadIdx.x == 0) {
    double *slot = pb_slot(P);
    slot[2 * blockIdx.x] = bb;
    slot[2 * blockIdx.x + 1] = rz;
  }
}

__global__ void __launch_bounds__(BS) k_pcg_setup_finalize(SolverState *st, PcgPartials P, double tol, int maxit,
                                                           double delta2) {
  __shared__ double scr[BS / 64];
  const double bb = block_total<BS>(pb_slot(P), P.nB, 2, scr);
  if (threadIdx.x == 0) {
    st->rhsNorm2 = bb;
    double thr = tol * tol * bb;
    if (thr < 2.2250738585072014e-308) thr = 2.2250738585072014e-308;
    st->threshold = thr;
    st->rr_ref = bb;
    st->delta2 = delta2;
    st->rr = bb;
    st->true_rr = bb;
    st->alpha_owed = 0.0;
    st->maxit = maxit;
    st->iter_cur = 0;
    st->iter = 0;
    st->stop_iter = 0;
    st->flush_iter = -1;
    st->pending = 0;
    // Eigen: a zero right-hand side returns x = 0 at once; so does a converged start
    st->done = (bb == 0.0 || bb < thr) ? 1 : 0;
  }
}

template <typename T>
void launch_pcg_setup(long long n, SolverState *st, const T *b, T *r, const T *invd, PcgPartials P, double tol,
                      int maxit, double delta2, hipStream_t s) {
  hipLaunchKernelGGL(k_pcg_setup<T>, dim3(P.nB), dim3(BS), 0, s, n, b, r, invd, P);
  hipLaunchKernelGGL(k_pcg_setup_finalize, dim3(1), dim3(BS), 0, s, st, P, tol, maxit, delta2);
  ICX_CHECK_LAUNCH();
}

// 16-byte vector of T
template <typename T> struct Vec16;
template <> struct Vec16<float> { using type = float4; static constexpr int W = 4; };
template <> struct Vec16<double> { using type = double2; static constexpr int W = 2; };

// element w of a vector (w is a compile-time constant after unrolling)
__device__ __forceinline__ float lane(const float4 &v, int w) { return w == 0 ? v.x : w == 1 ? v.y : w == 2 ? v.z : v.w; }
__device__ __forceinline__ double lane(const double2 &v, int w) { return w == 0 ? v.x : v.y; }
__device__ __forceinline__ void set_lane(float4 &v, int w, float x) {
  if (w == 0) v.x = x; else if (w == 1) v.y = x; else if (w == 2) v.z = x; else v.w = x;
}
__device__ __forceinline__ void set_lane(double2 &v, int w, double x) {
  if (w == 0) v.x = x; else v.y = x;
}

// Kernel B of PCG iteration i: alpha_i = r_i.z_i / p_i.Ap_i, r -= alpha_i Ap,
// partials (|r|^2, r.invd r).  Each thread moves U 16-byte vectors of r, Ap and invd
// per round; the first round's loads are issued before alpha_i is reduced.  Block 0
// publishes alpha_i.  The vectors are 16-byte aligned (device allocations).
template <typename T, int U>
__device__ __forceinline__ void fusedB_block(const BArgs<T> &a, int blk) {
  using V = typename Vec16<T>::type;
  constexpr int W = Vec16<T>::W;
  __shared__ double scr[BS / 64];
  __shared__ double scr2[2 * (BS / 64)];
  const long long n = a.n;
  SolverState *st = a.st;
  T *r = a.r;
  const T *Ap = a.Ap;
  const T *invd = a.invd;
  const PcgPartials &P = a.P;
  const int nblk = P.nB;
  if (blk >= nblk) return;     // padding block of a batched launch
  // the state, the first round of vectors and the p.Ap partials go out together
  const SolverState ss = load_state(st);
  const long long nv = n / W;
  const long long stride = (long long)nblk * BS;
  const long long v0 = blk * (long long)BS + threadIdx.x;
  V *rv = reinterpret_cast<V *>(r);
  const V *av = reinterpret_cast<const V *>(Ap);
  const V *dv = reinterpret_cast<const V *>(invd);
  const V *a2v = reinterpret_cast<const V *>(a.Ap2);
  const bool seams = a.sg.kmax > 0;
  V R[U], A[U], D[U], A2[U];
  uint32_t sf = 0;  // seam flags: bit W u + w for lane w of chunk u
  auto load = [&](long long base) {
    sf = 0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long long v = base + u * stride;
      if (v < nv) {
        R[u] = rv[v];
        A[u] = av[v];
        D[u] = dv[v];
        if (seams) {
          const uint32_t f = seam_flags_aos((uint32_t)(v * W), W, a.sg);
          sf |= f << (W * u);
          if (f) A2[u] = a2v[v];
        }
      }
    }
  };
  load(v0);
  // both exits are uniform and leave before the partials reduction, which only a running
  // iteration uses
  if (ss.done) return;
  if (ss.pending) {  // kernel A of this iteration stopped: publish the stop
    if (blk == 0 && threadIdx.x == 0) st->done = ss.pending;
    return;
  }
  const double pAp = block_total<BS>(P.PA, P.nA, 1, scr);
  const int it = ss.iter_cur;
  const double alpha_d = ((it & 1) ? ss.rz_hist[1] : ss.rz_hist[0]) / pAp;
  if (blk == 0 && threadIdx.x == 0) st->alpha_hist[it & 1] = alpha_d;
  const T alpha = (T)alpha_d;
  double rr = 0.0, rzn = 0.0;
  for (long long base = v0; base < nv; base += U * stride) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long long v = base + u * stride;
      if (v < nv) {
        V out;
#pragma unroll
        for (int w = 0; w < W; ++w) {
          T apw = lane(A[u], w);
          if ((sf >> (W * u + w)) & 1) apw += lane(A2[u], w);
          const T ri = fma_t(-alpha, apw, lane(R[u], w));
          set_lane(out, w, ri);
          rr = fma_t((double)ri, (double)ri, rr);
          rzn = fma_t((double)ri, (double)(lane(D[u], w) * ri), rzn);
        }
        rv[v] = out;
      }
    }
    if (base + U * stride < nv) load(base + U * stride);
  }
  // scalar tail (n not a multiple of W)
  const long long tail = nv * W + v0;
  if (tail < n) {
    T apt = Ap[tail];
    if (seams && seam_flags_aos((uint32_t)tail, 1, a.sg)) apt += a.Ap2[tail];
    const T ri = fma_t(-alpha, apt, r[tail]);
    r[tail] = ri;
    rr = fma_t((double)ri, (double)ri, rr);
    rzn = fma_t((double)ri, (double)(invd[tail] * ri), rzn);
  }
  rr = wave_sum(rr);
  rzn = wave_sum(rzn);
  if ((threadIdx.x & 63) == 0) {
    scr2[threadIdx.x >> 6] = rr;
    scr2[BS / 64 + (threadIdx.x >> 6)] = rzn;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s0 = 0.0, s1 = 0.0;
#pragma unroll
    for (int w = 0; w < BS / 64; ++w) {
      s0 += scr2[w];
      s1 += scr2[BS / 64 + w];
    }
    double *slot = pb_slot(P);
    slot[2 * blk] = s0;
    slot[2 * blk + 1] = s1;
    if (blk == 0) st->iter = it + 1;
  }
}

// Kernel B on a hexa8 box job with the per-node Jacobi factor (node-interleaved r and Ap,
// entry 3 n + c): each thread takes chunks of 4 consecutive nodes, 12 consecutive
// entries (three 16-byte vectors each of r and Ap, one of invs read once for the three
// components), so the preconditioner costs one word per node instead of one per entry.
// Nodes past the last full chunk (nn mod 4) are taken by thread 0 of block 0.
template <int U>
__device__ __forceinline__ void fusedB_node_block(const BArgs<float> &a, int blk) {
  __shared__ double scr[BS / 64];
  __shared__ double scr2[2 * (BS / 64)];
  SolverState *st = a.st;
  const PcgPartials &P = a.P;
  const int nblk = P.nB;
  if (blk >= nblk) return;     // padding block of a batched launch
  const SolverState ss = load_state(st);
  const long long nn = a.nn, nq = nn / 4;
  const long long stride = (long long)nblk * BS;
  const long long q0 = blk * (long long)BS + threadIdx.x;
  const bool seams = a.sg.kmax > 0;
  float4 R[U][3], A[U][3], A2[U][3];
  float4 S[U];
  uint32_t sf[U];
  auto load = [&](long long base) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long long q = base + u * stride;
      sf[u] = 0;
      if (q < nq) {
        S[u] = *reinterpret_cast<const float4 *>(a.invs + 4 * q);
        if (seams) sf[u] = seam_flags((uint32_t)(4 * q), 4, a.sg);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          R[u][k] = *reinterpret_cast<const float4 *>(a.r + 12 * q + 4 * k);
          A[u][k] = *reinterpret_cast<const float4 *>(a.Ap + 12 * q + 4 * k);
          if (sf[u]) A2[u][k] = *reinterpret_cast<const float4 *>(a.Ap2 + 12 * q + 4 * k);
        }
      }
    }
  };
  load(q0);
  if (ss.done) return;
  if (ss.pending) {
    if (blk == 0 && threadIdx.x == 0) st->done = ss.pending;
    return;
  }
  const double pAp = block_total<BS>(P.PA, P.nA, 1, scr);
  const int it = ss.iter_cur;
  const double alpha_d = ((it & 1) ? ss.rz_hist[1] : ss.rz_hist[0]) / pAp;
  if (blk == 0 && threadIdx.x == 0) st->alpha_hist[it & 1] = alpha_d;
  const float alpha = (float)alpha_d;
  double rr = 0.0, rzn = 0.0;
  auto node = [&](float sv, float apw, float &rv, int c) {
    const float ri = fma_t(-alpha, apw, rv);
    rv = ri;
    rr = fma_t((double)ri, (double)ri, rr);
    rzn = fma_t((double)ri, (double)(jacobi_inv(sv, a.ikd[c]) * ri), rzn);
  };
  for (long long base = q0; base < nq; base += U * stride) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long long q = base + u * stride;
      if (q < nq) {
        const float sv[4] = {S[u].x, S[u].y, S[u].z, S[u].w};
        float rv[12], av[12], a2[12];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          rv[4 * k] = R[u][k].x; rv[4 * k + 1] = R[u][k].y; rv[4 * k + 2] = R[u][k].z; rv[4 * k + 3] = R[u][k].w;
          av[4 * k] = A[u][k].x; av[4 * k + 1] = A[u][k].y; av[4 * k + 2] = A[u][k].z; av[4 * k + 3] = A[u][k].w;
          a2[4 * k] = A2[u][k].x; a2[4 * k + 1] = A2[u][k].y; a2[4 * k + 2] = A2[u][k].z;
          a2[4 * k + 3] = A2[u][k].w;
        }
        // entry e = 3 w + c: component c of node w of the chunk
#pragma unroll
        for (int e = 0; e < 12; ++e) {
          const int w = e / 3, c = e % 3;
          float apw = av[e];
          if ((sf[u] >> w) & 1) apw += a2[e];
          node(sv[w], apw, rv[e], c);
        }
#pragma unroll
        for (int k = 0; k < 3; ++k)
          *reinterpret_cast<float4 *>(a.r + 12 * q + 4 * k) =
              make_float4(rv[4 * k], rv[4 * k + 1], rv[4 * k + 2], rv[4 * k + 3]);
      }
    }
    if (base + U * stride < nq) load(base + U * stride);
  }
  if (blk == 0 && threadIdx.x == 0) {
    for (long long n = 4 * nq; n < nn; ++n) {
      const bool sfl = seams && seam_flags((uint32_t)n, 1, a.sg);
      for (int c = 0; c < 3; ++c) {
        float apw = a.Ap[3 * n + c];
        if (sfl) apw += a.Ap2[3 * n + c];
        float rv = a.r[3 * n + c];
        node(a.invs[n], apw, rv, c);
        a.r[3 * n + c] = rv;
      }
    }
  }
  rr = wave_sum(rr);
  rzn = wave_sum(rzn);
  if ((threadIdx.x & 63) == 0) {
    scr2[threadIdx.x >> 6] = rr;
    scr2[BS / 64 + (threadIdx.x >> 6)] = rzn;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s0 = 0.0, s1 = 0.0;
#pragma unroll
    for (int w = 0; w < BS / 64; ++w) {
      s0 += scr2[w];
      s1 += scr2[BS / 64 + w];
    }
    double *slot = pb_slot(P);
    slot[2 * blk] = s0;
    slot[2 * blk + 1] = s1;
    if (blk == 0) st->iter = it + 1;
  }
}

template <typename T, int U>
__device__ __forceinline__ void fusedB_any(const BArgs<T> &a, int blk) {
  if constexpr (sizeof(T) == 4) {
    if (a.invs) {
      fusedB_node_block<2>(a, blk);
      return;
    }
  }
  fusedB_block<T, U>(a, blk);
}

template <typename T, int U>
__global__ void __launch_bounds__(BS) k_pcg_fusedB(BArgs<T> a) {
  fusedB_any<T, U>(a, blockIdx.x);
}
template <typename T, int U>
__global__ void __launch_bounds__(BS)
    k_pcg_fusedB_batch(const BArgs<T> *__restrict__ args, const int *__restrict__ off, int njobs) {
  int j, blk;
  batch_block(off, njobs, j, blk);
  fusedB_any<T, U>(args[j], blk);
}

template <typename T>
void launch_pcg_fusedB(long long n, SolverState *st, T *r, const T *Ap, const T *invd, PcgPartials P,
                       const T *Ap2, const SeamGeom &sg, hipStream_t s, const LaunchEvents *ev,
                       const T *invs, const T *ikd) {
  BArgs<T> a{n, st, r, Ap, invd, P, Ap2, sg, invs, {T(0), T(0), T(0)}, n / 3};
  if (ikd)
    for (int c = 0; c < 3; ++c) a.ikd[c] = ikd[c];
  launch_ev((k_pcg_fusedB<T, 4>), P.nB, BS, s, ev, a);
  ICX_CHECK_LAUNCH();
}

__global__ void k_pcg_flush(long long n, const SolverState *st, double *x, const double *P0, const double *P1) {
  const int it = st->stop_iter;
  if (it <= 0 || st->flush_iter == it) return;
  const double a = st->alpha_owed;
  const double *p = ((it - 1) & 1) ? P1 : P0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    x[i] += a * p[i];
}
void launch_pcg_flush(long long n, SolverState *st, double *x, const double *P0, const double *P1, hipStream_t s) {
  hipLaunchKernelGGL(k_pcg_flush, dim3(ew_blocks(n, 4)), dim3(256), 0, s, n, st, x, P0, P1);
  ICX_CHECK_LAUNCH();
}

// FP64 entry of FP32 entry i: the same index, or with a node-interleaved FP32 vector over
// nn_aos nodes of nc components (entry nc n + c) the component-plane entry c nn_aos + n
__device__ __forceinline__ long long plane_index(long long i, long long nn_aos, int nc) {
  if (!nn_aos) return i;
  const long long n = i / nc;
  return (i - nc * n) * nn_aos + n;
}

__global__ void k_refine_x(long long n, const SolverState *st, double *xhi, float *xlo, const float *P0,
                           const float *P1, long long nn_aos, int nc) {
  const int it = st->stop_iter;
  const float a = (it > 0 && st->flush_iter != it) ? (float)st->alpha_owed : 0.0f;
  const float *p = ((it - 1) & 1) ? P1 : P0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const float xl = xlo[i] + a * p[i];
    xhi[plane_index(i, nn_aos, nc)] += (double)xl;
    xlo[i] = 0.0f;
  }
}
void launch_refine_x(long long n, SolverState *st, double *xhi, float *xlo, const float *P0, const float *P1,
                     long long nn_aos, int nc, hipStream_t s) {
  hipLaunchKernelGGL(k_refine_x, dim3(ew_blocks(n, 4)), dim3(256), 0, s, n, st, xhi, xlo, P0, P1, nn_aos, nc);
  ICX_CHECK_LAUNCH();
}

__global__ void __launch_bounds__(BS) k_refine_r(long long n, const SolverState *st, const double *b,
                                                 const double *Kx, float *r32, const float *invd32, PcgPartials P,
                                                 long long nn_aos, int nc) {
  __shared__ double scr[BS / 64];
  double rr = 0.0, rz = 0.0;
  for (long long i = blockIdx.x * (long long)BS + threadIdx.x; i < n; i += (long long)gridDim.x * BS) {
    const long long i64 = plane_index(i, nn_aos, nc);
    const double ri = b[i64] - Kx[i64];
    const float rf = (float)ri;
    r32[i] = rf;
    rr += ri * ri;
    rz += (double)rf * (double)(invd32[i] * rf);
  }
  rr = block_sum<BS>(rr, scr);
  rz = block_sum<BS>(rz, scr);
  if (threadIdx.x == 0) {
    double *slot = pb_slot(P);  // replaces the residual of iteration stop_iter
    slot[2 * blockIdx.x] = rr;
    slot[2 * blockIdx.x + 1] = rz;
  }
}
__global__ void __launch_bounds__(BS) k_refine_finalize(SolverState *st, PcgPartials P) {
  __shared__ double scr[BS / 64];
  const double rr = block_total<BS>(pb_slot(P), P.nB, 2, scr);
  if (threadIdx.x == 0) {
    st->pending = 0;
    st->true_rr = rr;
    st->rr = rr;
    if (rr < st->threshold) {
      st->done = 1;
    } else {
      // resume at the stopped iteration with the replaced residual; its owed x step
      // is already in x_hi
      st->rr_ref = rr;
      st->flush_iter = st->stop_iter;
      st->iter = st->stop_iter;
      st->done = (st->stop_iter >= st->maxit) ? 2 : 0;
    }
  }
}
void launch_refine_r(long long n, SolverState *st, const double *b, const double *Kx, float *r32,
                     const float *invd32, PcgPartials P, long long nn_aos, int nc, hipStream_t s) {
  hipLaunchKernelGGL(k_refine_r, dim3(P.nB), dim3(BS), 0, s, n, st, b, Kx, r32, invd32, P, nn_aos, nc);
  hipLaunchKernelGGL(k_refine_finalize, dim3(1), dim3(BS), 0, s, st, P);
  ICX_CHECK_LAUNCH();
}

template <typename T>
void launch_pcg_fusedB_batch(const BArgs<T> *args, const int *off, int njobs, int total_blocks, hipStream_t s) {
  if (njobs <= 0 || total_blocks <= 0) return;
  hipLaunchKernelGGL((k_pcg_fusedB_batch<T, 4>), dim3(total_blocks), dim3(BS), 0, s, args, off, njobs);
  ICX_CHECK_LAUNCH();
}
template void launch_pcg_fusedB_batch<float>(const BArgs<float> *, const int *, int, int, hipStream_t);
template void launch_pcg_fusedB_batch<double>(const BArgs<double> *, const int *, int, int, hipStream_t);

#define ICX_INST_PCG(T)                                                                                  \
  template void launch_pcg_setup<T>(long long, SolverState *, const T *, T *, const T *, PcgPartials,    \
                                    double, int, double, hipStream_t);                                   \
  template void launch_pcg_fusedB<T>(long long, SolverState *, T *, const T *, const T *, PcgPartials, \
                                     const T *, const SeamGeom &, hipStream_t, const LaunchEvents *,  \
                                     const T *, const T *);
ICX_INST_PCG(double)
ICX_INST_PCG(float)

// ---- the Eigen ConjugateGradient loop (FP64), one kernel per vector expression -----------
// (Eigen 3.4 ConjugateGradient.h as called by src/analysis/solver/solvers.cpp:64-92.)
// Iteration i: absNew_i = r_i.z_i is the PZ total of iteration i-1 (of the set-up for i = 0).
__host__ __device__ static inline double *slot1(double *base, int parity) { return base + (size_t)(parity & 1) * RED_MAXBLOCKS; }

__global__ void __launch_bounds__(BS) k_eig_setup(long long n, const double *b, double *r, double *p,
                                                  const double *invd, PcgPartials P) {
  __shared__ double scr[BS / 64];
  double bb = 0.0, rz = 0.0;
  for (long long i = blockIdx.x * (long long)BS + threadIdx.x; i < n; i += (long long)gridDim.x * BS) {
    const double bi = b[i];
    r[i] = bi;
    const double zi = invd[i] * bi;
    p[i] = zi;
    bb += bi * bi;
    rz += bi * zi;
  }
  bb = block_sum<BS>(bb, scr);
  rz = block_sum<BS>(rz, scr);
  if (threadIdx.x == 0) {
    slot1(P.PN, 1)[blockIdx.x] = bb;
    slot1(P.PZ, 1)[blockIdx.x] = rz;
  }
}
__global__ void __launch_bounds__(BS) k_eig_setup_finalize(SolverState *st, PcgPartials P, double tol, int maxit) {
  __shared__ double scr[BS / 64];
  const double bb = block_total<BS>(slot1(P.PN, 1), P.nB, 1, scr);
  if (threadIdx.x == 0) {
    st->rhsNorm2 = bb;
    double thr = tol * tol * bb;
    if (thr < 2.2250738585072014e-308) thr = 2.2250738585072014e-308;
    st->threshold = thr;
    st->rr = bb;
    st->maxit = maxit;
    st->iter_cur = 0;
    st->iter = 0;
    st->stop_iter = 0;
    // fields the fused / mixed forms leave behind: the host maps a non-zero `pending`
    // to `done`, so a stale stop code would end this solve before its first iteration
    st->pending = 0;
    st->flush_iter = -1;
    st->alpha_owed = 0.0;
    st->delta2 = 0.0;
    st->done = (bb == 0.0 || bb < thr) ? 1 : 0;
  }
}
void launch_eigen_setup(long long n, SolverState *st, const double *b, double *r, double *p, const double *invd,
                        PcgPartials P, double tol, int maxit, hipStream_t s) {
  hipLaunchKernelGGL(k_eig_setup, dim3(P.nB), dim3(BS), 0, s, n, b, r, p, invd, P);
  hipLaunchKernelGGL(k_eig_setup_finalize, dim3(1), dim3(BS), 0, s, st, P, tol, maxit);
  ICX_CHECK_LAUNCH();
}

// p.dot(tmp)
__global__ void __launch_bounds__(BS) k_eig_dot_pt(long long n, SolverState *st, const double *p,
                                                   const double *t, PcgPartials P) {
  __shared__ double scr[BS / 64];
  if (st->done) return;
  const int it = st->iter;  // first kernel of the iteration: publish it
  if (blockIdx.x == 0 && threadIdx.x == 0) st->iter_cur = it;
  double s = 0.0;
  for (long long i = blockIdx.x * (long long)BS + threadIdx.x; i < n; i += (long long)gridDim.x * BS) s += p[i] * t[i];
  s = block_sum<BS>(s, scr);
  if (threadIdx.x == 0) slot1(P.PA, it)[blockIdx.x] = s;
}
// alpha = absNew / p.dot(tmp);  y += sign * alpha * v  (x += alpha p; residual -= alpha tmp)
__global__ void __launch_bounds__(BS) k_eig_axpy(long long n, const SolverState *st, double sign, double *y,
                                                 const double *v, PcgPartials P) {
  __shared__ double scr[BS / 64];
  if (st->done) return;
  const int it = st->iter_cur;
  const double absNew = block_total<BS>(slot1(P.PZ, it - 1), P.nB, 1, scr);
  const double pAp = block_total<BS>(slot1(P.PA, it), P.nB, 1, scr);
  const double a = sign * (absNew / pAp);
  for (long long i = blockIdx.x * (long long)BS + threadIdx.x; i < n; i += (long long)gridDim.x * BS) y[i] += a * v[i];
}
// residualNorm2 = residual.squaredNorm()
__global__ void __launch_bounds__(BS) k_eig_norm(long long n, const SolverState *st, const double *r, PcgPartials P) {
  __shared__ double scr[BS / 64];
  if (st->done) return;
  const int it = st->iter_cur;
  double s = 0.0;
  for (long long i = blockIdx.x * (long long)BS + threadIdx.x; i < n; i += (long long)gridDim.x * BS) s += r[i] * r[i];
  s = block_sum<BS>(s, scr);
  if (threadIdx.x == 0) slot1(P.PN, it)[blockIdx.x] = s;
}
// `done` as one value for the whole block.  The two kernels below set `done` from
// block 0 while blocks of the same launch may still be starting; a per-wave read
// could split a block across its barriers.
__device__ __forceinline__ bool block_done(const SolverState *st) {
  __shared__ int s_done;
  if (threadIdx.x == 0) s_done = __hip_atomic_load(&st->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  return s_done != 0;
}

// if (residualNorm2 < threshold) break;  z = precond.solve(residual)
__global__ void __launch_bounds__(BS) k_eig_precond(long long n, SolverState *st, double *z, const double *invd,
                                                    const double *r, PcgPartials P) {
  __shared__ double scr[BS / 64];
  if (block_done(st)) return;
  const int it = st->iter_cur;
  const double rr = block_total<BS>(slot1(P.PN, it), P.nB, 1, scr);
  if (rr < st->threshold) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->rr = rr;
      st->stop_iter = it;
      st->iter = it;
      st->done = 1;
    }
    return;
  }
  for (long long i = blockIdx.x * (long long)BS + threadIdx.x; i < n; i += (long long)gridDim.x * BS)
    z[i] = invd[i] * r[i];
}
// absNew = residual.dot(z)
__global__ void __launch_bounds__(BS) k_eig_dot_rz(long long n, const SolverState *st, const double *r,
                                                   const double *z, PcgPartials P) {
  __shared__ double scr[BS / 64];
  if (st->done) return;
  const int it = st->iter_cur;
  double s = 0.0;
  for (long long i = blockIdx.x * (long long)BS + threadIdx.x; i < n; i += (long long)gridDim.x * BS) s += r[i] * z[i];
  s = block_sum<BS>(s, scr);
  if (threadIdx.x == 0) slot1(P.PZ, it)[blockIdx.x] = s;
}
// beta = absNew / absOld;  p = z + beta * p;  i++
__global__ void __launch_bounds__(BS) k_eig_pupdate(long long n, SolverState *st, double *p, const double *z,
                                                    PcgPartials P) {
  __shared__ double scr[BS / 64];
  if (block_done(st)) return;
  const int it = st->iter_cur;
  const double absNew = block_total<BS>(slot1(P.PZ, it), P.nB, 1, scr);
  const double absOld = block_total<BS>(slot1(P.PZ, it - 1), P.nB, 1, scr);
  const double beta = absNew / absOld;
  for (long long i = blockIdx.x * (long long)BS + threadIdx.x; i < n; i += (long long)gridDim.x * BS)
    p[i] = z[i] + beta * p[i];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->iter = it + 1;
    if (it + 1 >= st->maxit) {
      st->stop_iter = it + 1;
      st->done = 2;
    }
  }
}

void launch_eigen_iteration_tail(long long n, SolverState *st, double *x, double *r, double *p, const double *t,
                                 double *z, const double *invd, PcgPartials P, hipStream_t s) {
  const dim3 nb(P.nB), bs(BS);
  hipLaunchKernelGGL(k_eig_dot_pt, nb, bs, 0, s, n, st, p, t, P);
  hipLaunchKernelGGL(k_eig_axpy, nb, bs, 0, s, n, st, 1.0, x, p, P);
  hipLaunchKernelGGL(k_eig_axpy, nb, bs, 0, s, n, st, -1.0, r, t, P);
  hipLaunchKernelGGL(k_eig_norm, nb, bs, 0, s, n, st, r, P);
  hipLaunchKernelGGL(k_eig_precond, nb, bs, 0, s, n, st, z, invd, r, P);
  hipLaunchKernelGGL(k_eig_dot_rz, nb, bs, 0, s, n, st, r, z, P);
  hipLaunchKernelGGL(k_eig_pupdate, nb, bs, 0, s, n, st, p, z, P);
  ICX_CHECK_LAUNCH();
}

// ---- batched solves: gather the states of n jobs into one array (one block per job) ----
__global__ void k_gather_states(SolverState *const *src, SolverState *dst, int n) {
  constexpr int W = sizeof(SolverState) / sizeof(int);
  static_assert(sizeof(SolverState) % sizeof(int) == 0, "state is a whole number of words");
  const int j = blockIdx.x;
  if (j >= n) return;
  const int *s = reinterpret_cast<const int *>(src[j]);
  int *d = reinterpret_cast<int *>(dst + j);
  for (int w = threadIdx.x; w < W; w += blockDim.x) d[w] = s[w];
}
void launch_gather_states(SolverState *const *src, SolverState *dst, int n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_gather_states, dim3(n), dim3(64), 0, s, src, dst, n);
  ICX_CHECK_LAUNCH();
}

// ---- generic reductions ----------------------------------------------------------------
__global__ void __launch_bounds__(BS) k_finalize(Reduce red, int nblocks, int nv, double *out) {
  __shared__ double scr[BS / 64];
  for (int q = 0; q < nv; ++q) {
    const double t = block_total<BS>(red.partial + (size_t)q * RED_MAXBLOCKS, nblocks, 1, scr);
    if (threadIdx.x == 0) out[q] = t;
  }
}
void launch_finalize(const Reduce &red, int nblocks, int nv, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(BS), 0, s, red, nblocks, nv, out);
  ICX_CHECK_LAUNCH();
}

__global__ void __launch_bounds__(BS) k_dot(long long n, const double *a, const double *b, Reduce red) {
  __shared__ double scr[BS / 64];
  double s = 0.0;
  for (long long i = blockIdx.x * (long long)BS + threadIdx.x; i < n; i += (long long)gridDim.x * BS) s += a[i] * b[i];
  s = block_sum<BS>(s, scr);
  if (threadIdx.x == 0) red.partial[blockIdx.x] = s;
}
void launch_dot(long long n, const double *a, const double *b, double *out, Reduce red, hipStream_t s) {
  const int nb = ew_blocks(n, 4, RED_BLOCKS);
  hipLaunchKernelGGL(k_dot, dim3(nb), dim3(BS), 0, s, n, a, b, red);
  launch_finalize(red, nb, 1, out, s);
}

}  // namespace icx
