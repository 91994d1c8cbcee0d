// Single-reduction PCG iteration of the FP32 hexa8 mixed form (gfx950): one kernel per
// iteration.
//
// The iterates are those of the two-kernel form (Hestenes-Stiefel PCG with Jacobi
// preconditioner z = invd r); only the scalar r_k.z_k that beta_k needs is formed before
// r_k exists, from the expansion
//   r_k.z_k = rz_{k-1} - 2 alpha mu_{k-1} + alpha^2 nu_{k-1},
//   mu = Ap.z,  nu = Ap.invd Ap,  alpha = alpha_{k-1} = rz_{k-1} / pi_{k-1},  pi = p.Ap
// whose terms the launch of iteration k-1 reduced together with rz and |r|^2.  Launch k:
//   stop test on |r_{k-1}|^2;  alpha_{k-1};  r_k = r_{k-1} - alpha Ap_{k-1};
//   x_k = x_{k-1} + alpha p_{k-1};  z_k = invd r_k;  beta_k = (r_k.z_k)~ / rz_{k-1};
//   p_k = z_k + beta_k p_{k-1};  Ap_k = K p_k;  partials (rz_k, |r_k|^2, pi_k, mu_k, nu_k)
// with rz_k formed directly from the new r_k (alpha_k uses it).  So one grid-wide
// reduction phase per iteration instead of two, and one pass over the vectors: the
// update of r, x and p is fused into the operator's loader.
//
// Tiling: a block of 16 x 16 threads owns a 15 x 15 column of nodes over node layers
// [k0, k1) and applies the element layers k0-1 .. k1-1 (the first one also applied by
// the block below), so every owned node's Ap is complete and the quadratic nu can be
// summed per node.  The loader forms r_k, z_k and p_k for every patch node (the halo
// too) from r_{k-1}, Ap_{k-1}, p_{k-1} and invd with the same operations as the owning
// block, so all blocks agree bit for bit; loader slot 0 of an owning thread is its own
// node, whose r_k, p_k and x_k are stored at the commit of the layer.  r, p and Ap
// alternate between two buffers by launch parity (the halo is read from neighbours'
// columns); x is updated in place (owned nodes only).
//
// Restart (set-up, or after a residual replacement of r_{k-1} by the stop of launch k):
// iterate k-1 is formed again from the new r: p = z + beta' p_{k-2}, beta' = rz' / rz_{k-2}
// (0 at the set-up), Ap = K p and its partials, as the two-kernel form continues the
// direction after a replacement.
#include "icx_device.h"

namespace icx {

namespace {

constexpr int TI = 15, TJ = 15;
constexpr int NT = (TI + 1) * (TJ + 1);               // 256 threads
constexpr int LW = TI + 2, LN = (TI + 2) * (TJ + 2);  // 17 x 17 node patch
constexpr int NIN = TI * TJ;                          // owned nodes per layer
constexpr int NH0 = NT - NIN;                         // halo nodes in loader slot 0 (31)
constexpr int NH1 = LN - NIN - NH0;                   // halo nodes in loader slot 1 (33)
constexpr int NP = 5;                                 // partials per block (rz, rr, pi, mu, nu)
constexpr int PREC = 6;                               // doubles per partial record
static_assert(TI + 1 == 16, "element rows are 16-lane DPP rows");
static_assert(NH1 > 0 && NH1 <= 64, "loader slot 1 is held by the first wave");

// patch position of halo node h: the bottom row, the top row, then the left and the
// right column without their corners
__device__ __forceinline__ void halo_pos(int h, int &li, int &lj) {
  if (h < LW) {
    li = h; lj = 0;
  } else if (h < 2 * LW) {
    li = h - LW; lj = TJ + 1;
  } else if (h < 2 * LW + TJ) {
    li = 0; lj = 1 + (h - 2 * LW);
  } else {
    li = TI + 1; lj = 1 + (h - 2 * LW - TJ);
  }
}

// totals of the NP partials of the records p[PREC q ..], q < n, in a fixed order,
// returned to every thread (one barrier; scratch: NP NT/64 doubles)
__device__ __forceinline__ void block_total5(const double *p, int n, double (*scratch)[NT / 64], double (&t)[NP]) {
  double a[PREC] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const double2 *q2 = reinterpret_cast<const double2 *>(p);
#pragma unroll 2
  for (int q = threadIdx.x; q < n; q += NT) {
#pragma unroll
    for (int h = 0; h < PREC / 2; ++h) {
      const double2 v = q2[(PREC / 2) * q + h];
      a[2 * h] += v.x;
      a[2 * h + 1] += v.y;
    }
  }
#pragma unroll
  for (int v = 0; v < NP; ++v) a[v] = wave_sum(a[v]);
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int v = 0; v < NP; ++v) scratch[v][threadIdx.x >> 6] = a[v];
  }
  __syncthreads();
#pragma unroll
  for (int v = 0; v < NP; ++v) {
    t[v] = 0.0;
#pragma unroll
    for (int w = 0; w < NT / 64; ++w) t[v] += scratch[v][w];
  }
}

// block sums of NP values, written by thread 0 to out[0 .. NP-1]
__device__ __forceinline__ void block_store5(double (&a)[NP], double (*scratch)[NT / 64], double *out) {
#pragma unroll
  for (int v = 0; v < NP; ++v) a[v] = wave_sum(a[v]);
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int v = 0; v < NP; ++v) scratch[v][threadIdx.x >> 6] = a[v];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int v = 0; v < NP; ++v) {
      double s = 0.0;
#pragma unroll
      for (int w = 0; w < NT / 64; ++w) s += scratch[v][w];
      out[v] = s;
    }
  }
}

// MODE 0: iteration of launch parity a.slot.  MODE 1: restart.
template <int MODE>
__device__ __forceinline__ void sr_block(const SrArgs &a, int blk) {
  __shared__ float s_in[3][3][LN];
  __shared__ float s_h[6][NT];
  __shared__ float s_zd[2][6][NT];  // z and invd of the thread's own node, by layer parity
  __shared__ double s_red[NP][NT / 64];
  constexpr uint32_t W = 4;
  const Grid &g = a.g;
  const SolverState ss = load_state(a.st);  // used once the first loads are out

  const int nblocks = a.tiles_x * a.tiles_y * a.tiles_z;
  if (blk >= nblocks) return;  // padding block of a batched launch
  const int t = xcd_tile(blk, nblocks);
  const int ti = t % a.tiles_x, tj = (t / a.tiles_x) % a.tiles_y, tk = t / (a.tiles_x * a.tiles_y);
  const int i0 = ti * TI, j0 = tj * TJ, k0 = tk * a.TK;
  const int k1 = min(k0 + a.TK, g.sz);  // owned node layers [k0, k1)
  const int kfirst = k0 - 1;            // first node layer and first element layer
  const int tid = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  // element of this thread (tx, ty) in the (TI+1) x (TJ+1) patch of elements around
  // the column; owned node (tx, ty) for tx < TI, ty < TJ (patch node (tx+1, ty+1))
  const int tx = tid % (TI + 1), ty = tid / (TI + 1);
  const int ex = i0 - 1 + tx, ey = j0 - 1 + ty;
  const bool ex_ok = ex >= 0 && ex < g.nx && ey >= 0 && ey < g.ny;
  const uint32_t e_off = ex_ok ? (uint32_t)(ex + g.nx * ey) * W : 0u;
  const uint32_t e_layer = (uint32_t)(g.nx * g.ny) * W;
  const bool node_thread = tx < TI && ty < TJ;
  const int ni = i0 + tx, nj = j0 + ty;
  const bool node_ok = node_thread && ni < g.sx && nj < g.sy;
  const uint32_t cstr = (uint32_t)g.nn * W;
  const uint32_t n_layer = (uint32_t)(g.sx * g.sy) * W;
  const uint32_t n_off = node_ok ? (uint32_t)(ni + g.sx * nj) * W : 0u;
  const int ctr = (tx + 1) + LW * (ty + 1);

  // Loader slots: slot 0 is the thread's own node (owning threads) or one of the first
  // NH0 halo nodes; slot 1 holds the other NH1 halo nodes (first wave only).  Nodes
  // outside the grid are read at offset 0: they are corners of elements outside the
  // grid only, whose modulus is zero, so the finite value read never reaches a result.
  uint32_t pos_off[2];
  int pos_lds[2];
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    int li = 0, lj = 0;
    bool has = true;
    if (p == 0) {
      if (node_thread) { li = tx + 1; lj = ty + 1; }
      else halo_pos(ty == TJ ? tx : TI + 1 + ty, li, lj);
    } else {
      has = tid < NH1;
      if (has) halo_pos(NH0 + tid, li, lj);
    }
    const int gi = i0 - 1 + li, gj = j0 - 1 + lj;
    const bool ok = has && gi >= 0 && gi < g.sx && gj >= 0 && gj < g.sy;
    pos_off[p] = ok ? (uint32_t)(gi + g.sx * gj) * W : 0u;
    pos_lds[p] = has ? li + LW * lj : -1;
  }
  auto pos_live = [&](int p) { return p == 0 || wave == 0; };
  auto owned_layer = [&](int k) { return k >= k0 && k < k1; };

  // loader registers of one node layer: r_{k-1}, Ap_{k-1}, p_{k-1}, invd; the owned
  // node's x
  float R[2][3], AP[2][3], P[2][3], D[2][3];
  float X[3];
  auto prefetch = [&](int k) {
    const uint32_t lo = (k >= 0 && k < g.sz) ? (uint32_t)k * n_layer : 0u;
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      if (!pos_live(p)) continue;
      if (p == 1 && pos_lds[1] < 0) continue;
      const uint32_t o = pos_off[p] + lo;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        R[p][c] = ld_off(a.r_in, o + c * cstr);
        D[p][c] = ld_off(a.invd, o + c * cstr);
        P[p][c] = ld_off(a.p_in, o + c * cstr);
        if (MODE == 0) AP[p][c] = ld_off(a.ap_in, o + c * cstr);
      }
    }
    if (MODE == 0 && node_ok && owned_layer(k)) {
      const uint32_t n = n_off + (uint32_t)k * n_layer;
#pragma unroll
      for (int c = 0; c < 3; ++c) X[c] = ld_off(a.x, n + c * cstr);
    }
  };
  float alpha = 0.0f, beta = 0.0f;
  double part[NP] = {0.0, 0.0, 0.0, 0.0, 0.0};  // rz, rr, pi, mu, nu
  // node layer k from the loader registers into LDS buffer buf (p_k); the owned node's
  // updates are stored at once, next to this step's loads
  auto commit = [&](int buf, int k) {
    const bool own = node_ok && owned_layer(k);
    const uint32_t n = n_off + (uint32_t)k * n_layer;
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      if (!pos_live(p)) continue;
      if (p == 1 && pos_lds[1] < 0) continue;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const float rk = MODE == 0 ? fma_t(-alpha, AP[p][c], R[p][c]) : R[p][c];
        const float z = D[p][c] * rk;
        const float pk = fma_t(beta, P[p][c], z);
        s_in[buf][c][pos_lds[p]] = pk;
        if (p == 0 && own) {
          if (MODE == 0) {
            st_off(a.r_out, n + c * cstr, rk);
            st_off(a.x, n + c * cstr, fma_t(alpha, P[0][c], X[c]));
          }
          st_off(a.p_out, n + c * cstr, pk);
          s_zd[k & 1][c][tid] = z;
          s_zd[k & 1][3 + c][tid] = D[0][c];
          part[0] = fma_t((double)rk, (double)z, part[0]);
          part[1] = fma_t((double)rk, (double)rk, part[1]);
        }
      }
    }
  };
  auto load_E = [&](int ez) -> float {
    const uint32_t lo = (ez >= 0 && ez < g.nz) ? (uint32_t)ez * e_layer : 0u;
    return ld_off(a.E, e_off + lo);
  };
  auto load_mask = [&](int k) -> uint32_t {
    const bool in = owned_layer(k);
    const uint32_t n = n_off + (in ? (uint32_t)k * n_layer : 0u);
    return a.mask ? (uint32_t)ld_off(a.mask, n / W) : 0u;
  };

  float acc_top[3] = {0.0f, 0.0f, 0.0f};
  // Ap of an owned node is stored one step late, next to that step's loads
  float yv[3] = {0.0f, 0.0f, 0.0f};
  bool have_out = false;
  auto store_out = [&](int ez, int buf) {
    if (have_out) {
      const uint32_t n = n_off + (uint32_t)ez * n_layer;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        st_off(a.ap_out, n + c * cstr, yv[c]);
        const double y = yv[c];
        part[2] = fma_t((double)s_in[buf][c][ctr], y, part[2]);
        part[3] = fma_t((double)s_zd[ez & 1][c][tid], y, part[3]);
        part[4] = fma_t((double)s_zd[ez & 1][3 + c][tid] * y, y, part[4]);
      }
    }
  };
  auto layer = [&](int ez, int tt, const float &Ec, float &En, uint32_t mc, uint32_t &mn) {
    const int b0 = tt % 3, b1 = (tt + 1) % 3;
    wait_vmem();                      // the previous step's loads and stores
    store_out(ez - 1, (tt + 2) % 3);  // Ap of node layer ez-1 (before commit reuses its s_zd)
    commit(b1, ez + 1);               // node layer ez+1
    if (ez + 2 <= k1) prefetch(ez + 2);
    En = load_E(ez + 1);
    mn = load_mask(ez + 1);
    __syncthreads();
    float hown[2][3];
    {
      float v[3][8];
      const bool ok = ex_ok && ez >= 0 && ez < g.nz;
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const int qa = q & 1, qb = (q >> 1) & 1, qc = q >> 2;
          v[c][q] = s_in[qc ? b1 : b0][c][(tx + qa) + LW * (ty + qb)];
        }
      elastic_hex8<float>(v, a.op, ok ? Ec : 0.0f);
      // pair sums along x across the 16-lane row: this element's qa = 1 corner plus the
      // qa = 0 corner of element tx + 1; qb = 0 sums belong to node row ty - 1 (LDS),
      // qb = 1 sums to node row ty (registers)
#pragma unroll
      for (int qc = 0; qc < 2; ++qc)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          s_h[3 * qc + c][tid] = v[c][1 + 4 * qc] + row_next(v[c][4 * qc]);
          hown[qc][c] = v[c][3 + 4 * qc] + row_next(v[c][2 + 4 * qc]);
        }
    }
    __syncthreads();
    have_out = false;
    if (node_thread) {
      float yb[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        yb[c] = (acc_top[c] + hown[0][c]) + s_h[c][tid + TI + 1];
        acc_top[c] = hown[1][c] + s_h[3 + c][tid + TI + 1];
      }
      if (ez >= k0 && node_ok) {
#pragma unroll
        for (int c = 0; c < 3; ++c) yv[c] = ((mc >> c) & 1) ? 0.0f : yb[c];
        have_out = true;
      }
    }
  };

  // the first node layer, its moduli and the partial sums go out before the state is
  // used
  prefetch(kfirst);
  float Ea = load_E(kfirst), Eb = 0.0f;
  uint32_t ma = 0u, mb = 0u;
  double tot[NP];
  if (MODE == 0) {
    block_total5(a.part_in, nblocks, s_red, tot);
  } else {
    double rr2, rz2;
    block_total2<NT>(a.pb, a.npb, &s_red[0][0], rr2, rz2);
    tot[0] = rz2;
    tot[1] = rr2;
  }
  if (ss.done || ss.pending) return;  // uniform: no launch writes these fields for its own blocks to act on
  if (MODE == 0) {
    const int k = ss.iter2[a.slot];
    const double rz = tot[0], rr = tot[1], pap = tot[2], mu = tot[3], nu = tot[4];
    int stop = 0;
    if (k != ss.resume_iter && rr < ss.threshold) stop = 3;
    else if (ss.delta2 > 0.0 && rr < ss.delta2 * ss.rr_ref) stop = 4;
    else if (k - 1 >= ss.maxit) stop = 2;
    const double al = rz / pap;
    const double rz_next = rz - 2.0 * al * mu + al * al * nu;
    if (blk == 0 && tid == 0) {
      SolverState *st = a.st;
      st->iter_cur = k;
      if (stop) {
        st->pending = stop;
        st->stop_iter = k - 1;
        st->rr = rr;
        st->alpha_owed = 0.0;
        st->stop_slot = a.slot;
      } else {
        st->rz_hist[(k - 1) & 1] = rz;
        st->alpha_hist[(k - 1) & 1] = al;
        st->iter2[a.slot ^ 1] = k + 1;
        st->iter = k;
      }
    }
    if (stop) return;
    alpha = (float)al;
    beta = (float)(rz_next / rz);
  } else {
    // iterate stop_iter formed again: beta' = rz' / rz_{stop_iter - 1}
    const int k = ss.stop_iter;
    beta = k > 0 ? (float)(tot[0] / ss.rz_hist[(k - 1) & 1]) : 0.0f;
    if (blk == 0 && tid == 0) {
      SolverState *st = a.st;
      st->iter2[0] = k + 1;
      st->resume_iter = k + 1;
      st->stop_slot = 0;
    }
  }
  wait_vmem();
  commit(0, kfirst);
  prefetch(kfirst + 1);
  int ez = kfirst, tt = 0;
  for (; ez + 1 < k1; ez += 2, tt += 2) {
    layer(ez, tt, Ea, Eb, ma, mb);
    layer(ez + 1, tt + 1, Eb, Ea, mb, ma);
  }
  if (ez < k1) {
    layer(ez, tt, Ea, Eb, ma, mb);
    store_out(ez, tt % 3);
  } else {
    store_out(ez - 1, (tt + 2) % 3);
  }
  block_store5(part, s_red, a.part_out + PREC * (size_t)blk);
}

template <int MODE>
__global__ void __launch_bounds__(NT) k_hex8_sr(SrArgs a) {
  sr_block<MODE>(a, blockIdx.x);
}
__global__ void __launch_bounds__(NT) k_hex8_sr_batch(const SrArgs *__restrict__ args, const int *__restrict__ off,
                                                      int njobs) {
  int j, blk;
  batch_block(off, njobs, j, blk);
  sr_block<0>(args[j], blk);
}

// Before a restart: p_{k-2} is wanted in p[1]; after a stop of a launch of parity 1 it
// is in p[0] and is copied.
__global__ void __launch_bounds__(256) k_sr_move_p(long long n, const SolverState *st, const float *p0, float *p1) {
  const SolverState ss = load_state(st);
  if (ss.done || ss.stop_slot == 0) return;
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) p1[i] = p0[i];
}

// tile depth: the deepest brick that still gives about 1.4 rounds of four blocks per
// CU, else at least 400 blocks, else depth 2
void sr_tiles(const Grid &g, int &tx, int &ty, int &tz, int &tk) {
  tx = (g.sx + TI - 1) / TI;
  ty = (g.sy + TJ - 1) / TJ;
  const long long cols = (long long)tx * ty;
  auto blocks = [&](int t) { return cols * ((g.sz + t - 1) / t); };
  tk = 2;
  if (g.tile_depth > 0) {
    tk = g.tile_depth;
  } else {
    const int cand[] = {8, 6, 4, 3};
    bool found = false;
    for (int t : cand)
      if (!found && blocks(t) >= 1434) { tk = t; found = true; }
    for (int t : cand)
      if (!found && blocks(t) >= 400) { tk = t; found = true; }
  }
  while (blocks(tk) > RED_MAXBLOCKS) tk <<= 1;
  tz = (g.sz + tk - 1) / tk;
}

}  // namespace

SrArgs sr_args(const Grid &g, const ElasticOp<float> &op, const SrVectors &v, SolverState *st, int slot) {
  SrArgs a{};
  a.g = g;
  sr_tiles(g, a.tiles_x, a.tiles_y, a.tiles_z, a.TK);
  a.op = op;
  a.E = v.E;
  a.mask = v.mask;
  a.invd = v.invd;
  const int o = slot ^ 1;
  a.r_in = v.r[slot]; a.ap_in = v.ap[slot]; a.p_in = v.p[slot];
  a.r_out = v.r[o]; a.ap_out = v.ap[o]; a.p_out = v.p[o];
  a.x = v.x;
  a.st = st;
  a.part_in = v.part[slot];
  a.part_out = v.part[o];
  a.pb = v.pb;
  a.npb = v.npb;
  a.slot = slot;
  return a;
}

int sr_blocks(const Grid &g) {
  int tx, ty, tz, tk;
  sr_tiles(g, tx, ty, tz, tk);
  return tx * ty * tz;
}

void launch_sr_iteration(const Grid &g, const ElasticOp<float> &op, const SrVectors &v, SolverState *st, int slot,
                         hipStream_t s, const LaunchEvents *ev) {
  const SrArgs a = sr_args(g, op, v, st, slot & 1);
  launch_ev(k_hex8_sr<0>, a.tiles_x * a.tiles_y * a.tiles_z, NT, s, ev, a);
  ICX_CHECK_LAUNCH();
}

void launch_sr_batch(const SrArgs *args, const int *off, int njobs, int total_blocks, hipStream_t s) {
  if (njobs <= 0 || total_blocks <= 0) return;
  hipLaunchKernelGGL(k_hex8_sr_batch, dim3(total_blocks), dim3(NT), 0, s, args, off, njobs);
  ICX_CHECK_LAUNCH();
}

void launch_sr_restart(const Grid &g, const ElasticOp<float> &op, const SrVectors &v, SolverState *st,
                       hipStream_t s) {
  const long long n = 3 * g.nn;
  hipLaunchKernelGGL(k_sr_move_p, dim3(ew_blocks(n, 4)), dim3(256), 0, s, n, st, v.p[0], v.p[1]);
  // reads r[0] and p[1] (p_{k-2}), writes p[0], Ap[0] and part[0]
  SrArgs a = sr_args(g, op, v, st, 1);
  a.r_in = v.r[0];
  a.ap_in = nullptr;
  a.part_in = nullptr;
  hipLaunchKernelGGL(k_hex8_sr<1>, dim3(a.tiles_x * a.tiles_y * a.tiles_z), dim3(NT), 0, s, a);
  ICX_CHECK_LAUNCH();
}

}  // namespace icx
