// Host side of the libicx C ABI (include/icx.h).
//
// Problem set-up follows the reference's sequence (MeshIcarat box mesh,
// BCIcarat Dirichlet ranges, countDOF, Neumann tractions into fext_org); the
// element operators are built once per job in FP64 and applied matrix-free on
// the device; the PCG loops drive the kernels of icx_kernels.hip and poll a
// small device-side state block every few iterations.
#include "../../include/icx.h"
#include "../../include/icx_profile.h"
#include "icx_internal.h"
#include "../host/toml_job.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

using namespace icx;

struct BatchRun;
static void batch_delete(BatchRun *b);
static void batch_remove(icx_ctx *c, icx_job *j);

// Scratch of icx_matvec_noalloc (the entry point for callers that cannot allocate
// device memory): one array per GPU, held until the process ends.  Host code never
// uses the symbol itself (on the host it names a stand-in, not device memory); each
// device's address comes from hipGetSymbolAddress (scratch_address, called by
// icx_ctx_create with that device current) and is kept per device.
constexpr size_t ICX_NOALLOC_SCRATCH_BYTES = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[ICX_NOALLOC_SCRATCH_BYTES];

struct icx_ctx {
  int device = 0;
  unsigned char *scratch = nullptr;  // this device's copy of `scratch` (device address)
  hipStream_t stream = nullptr;
  std::string err;
  SolverState *h_state = nullptr;  // pinned mirrors for polling, two slots
  hipEvent_t ev_state[2] = {nullptr, nullptr};  // state copy into slot k landed
  double *h_scalar = nullptr;      // pinned scalars
  int *h_int = nullptr;            // pinned ints (iteration base)
  bool use_graphs = true;          // replay captured iteration chunks (icx_set_option)
  bool single_reduction = false;   // hexa8 mixed solves: one kernel per iteration (opt-in)
  bool pipeline = true;            // keep one chunk queued ahead of the state read
  int tile_depth = 0;              // hexa8 operator z depth per block, 0 = automatic
  bool heat_dense = false;         // heat jobs: dense element rows even where Haar-diagonal
  bool mma_host_dual = false;      // MMA / GCMMA dual: Newton loop on the host
  bool mfma_matvec = false;        // FP32 hexa8 icx_matvec: dense K0 on the matrix cores
  // profile window (icx_profile.h)
  std::vector<icx_job *> jobs;  // live jobs, destroyed with the context
  int prof_kernel = 0;
  std::vector<hipEvent_t> ev_pool;
  int ev_used = 0;
  double prof_ms = 0.0;
  long long prof_launches = 0;
  BatchRun *batch = nullptr;    // submitted solves (icx_submit / icx_poll), created on first use
};

namespace {

// -------------------------------------------------------------------------------------
// errors
// -------------------------------------------------------------------------------------
icx_status fail(icx_ctx *ctx, icx_status code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (ctx) ctx->err = buf;
  return code;
}

// The device address of `scratch` on `device`, looked up once per device (the caller
// has made `device` current) and kept for every later context on it.
hipError_t scratch_address(int device, unsigned char **out) {
  static std::mutex mu;
  static std::vector<unsigned char *> addr;
  std::lock_guard<std::mutex> lock(mu);
  if ((int)addr.size() <= device) addr.resize(device + 1, nullptr);
  if (!addr[device]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    addr[device] = static_cast<unsigned char *>(p);
  }
  *out = addr[device];
  return hipSuccess;
}

#define HIPCHK(ctx, call)                                                                     \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      return fail((ctx), ICX_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                  __FILE__, __LINE__);                                                        \
  } while (0)

// device buffer with RAII
template <typename T> struct DBuf {
  T *p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  hipError_t alloc(size_t count) {
    if (count == n && p) return hipSuccess;
    release();
    n = count;
    return hipMalloc((void **)&p, std::max<size_t>(count, 1) * sizeof(T));
  }
  size_t bytes() const { return n * sizeof(T); }
};

// Pinned host buffer (hipHostMalloc; freed with the owner).
template <typename T> struct HBuf {
  T *p = nullptr;
  size_t n = 0;
  HBuf() = default;
  HBuf(const HBuf &) = delete;
  HBuf &operator=(const HBuf &) = delete;
  ~HBuf() { release(); }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    n = 0;
  }
  hipError_t alloc(size_t count) {
    if (count == n && p) return hipSuccess;
    release();
    n = count;
    return hipHostMalloc((void **)&p, std::max<size_t>(count, 1) * sizeof(T), hipHostMallocDefault);
  }
};

// Event pair for one launch of the profiled kernel, bound to the launch itself
// (LaunchEvents: the kernel's own start / end timestamps).
struct ProfScope {
  icx_ctx *c;
  bool on;
  LaunchEvents le{};
  ProfScope(icx_ctx *ctx, int kernel) : c(ctx), on(ctx->prof_kernel == kernel) {
    if (!on) return;
    if ((int)c->ev_pool.size() < c->ev_used + 2) {
      hipEvent_t a, b;
      if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) { on = false; return; }
      c->ev_pool.push_back(a);
      c->ev_pool.push_back(b);
    }
    le.start = c->ev_pool[c->ev_used];
    le.stop = c->ev_pool[c->ev_used + 1];
  }
  const LaunchEvents *ev() const { return on ? &le : nullptr; }
  ~ProfScope() {
    if (!on) return;
    c->ev_used += 2;
    c->prof_launches += 1;
  }
};

// accumulate recorded event pairs (the stream must be synchronised)
void prof_collect(icx_ctx *c) {
  for (int i = 0; i + 1 < c->ev_used; i += 2) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c->ev_pool[i], c->ev_pool[i + 1]) == hipSuccess) c->prof_ms += ms;
  }
  c->ev_used = 0;
}

// -------------------------------------------------------------------------------------
// element matrices of an axis-aligned box element (FP64, 2-point Gauss rules as in
// src/basis/gauss_integration.cpp:72-88, 268-289; B and De as in
// src/basis/bmatrix.hpp:28-162 and src/problem/elasticity.cpp:34-98).  Local node
// q = a + 2b + 4c is the corner (a, b, c) of the box.
// -------------------------------------------------------------------------------------
void material_matrix(int state, double E, double nu, double *D, int &voigt) {
  if (state == ICX_SOLID3D) {
    voigt = 6;
    std::fill(D, D + 36, 0.0);
    const double lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
    const double mu = E / (2 * (1.0 + nu));
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) D[i * 6 + j] = (i == j) ? lam + 2 * mu : lam;
    D[3 * 6 + 3] = D[4 * 6 + 4] = D[5 * 6 + 5] = mu;
    return;
  }
  voigt = 3;
  std::fill(D, D + 9, 0.0);
  double f;
  if (state == ICX_PLANE_STRESS) {
    D[0] = D[4] = 1.0; D[1] = D[3] = nu; D[8] = 0.5 * (1.0 - nu);
    f = E / (1.0 - nu * nu);
  } else {
    D[0] = D[4] = 1.0; D[1] = D[3] = nu / (1.0 - nu);
    D[8] = (1.0 - 2.0 * nu) / (2.0 * (1.0 - nu));
    f = E * ((1.0 - nu) / ((1.0 - 2.0 * nu) * (1.0 + nu)));
  }
  for (int i = 0; i < 9; ++i) D[i] *= f;
}

// shape-function derivatives w.r.t. physical coordinates at Gauss point gp (box element)
struct BoxElement {
  int ndim, nq;
  double h[3];
  // N[q] and dN/dX[d][q] at Gauss point ip, jac = det J * w
  void eval(int ip, double *N, double (*dNdX)[8], double &jac) const {
    const double g = 1.0 / std::sqrt(3.0);
    double r[3];
    for (int d = 0; d < ndim; ++d) r[d] = ((ip >> d) & 1) ? g : -g;
    jac = 1.0;
    for (int d = 0; d < ndim; ++d) jac *= 0.5 * h[d];
    const double sc = ndim == 3 ? 0.125 : 0.25;
    for (int q = 0; q < nq; ++q) {
      double f[3], df[3];
      for (int d = 0; d < ndim; ++d) {
        const double s = ((q >> d) & 1) ? 1.0 : -1.0;
        f[d] = 1.0 + s * r[d];
        df[d] = s;
      }
      double n = sc;
      for (int d = 0; d < ndim; ++d) n *= f[d];
      N[q] = n;
      for (int d = 0; d < ndim; ++d) {
        double v = sc * df[d];
        for (int e = 0; e < ndim; ++e)
          if (e != d) v *= f[e];
        dNdX[d][q] = v * 2.0 / h[d];  // dr/dX = 2/h on a box
      }
    }
  }
};

// K0 (unit modulus), ndof = nq*ndim, node-major (q*ndim + c); returns the volume
double elastic_K0(const BoxElement &el, int state, double nu, std::vector<double> &K) {
  const int nd = el.nq * el.ndim;
  K.assign(nd * nd, 0.0);
  double D[36];
  int voigt;
  material_matrix(state, 1.0, nu, D, voigt);
  double vol = 0.0;
  for (int ip = 0; ip < el.nq; ++ip) {
    double N[8], dN[3][8], jac;
    el.eval(ip, N, dN, jac);
    std::vector<double> B(voigt * nd, 0.0);
    for (int q = 0; q < el.nq; ++q) {
      if (el.ndim == 2) {
        B[0 * nd + 2 * q] = dN[0][q];
        B[1 * nd + 2 * q + 1] = dN[1][q];
        B[2 * nd + 2 * q] = dN[1][q];
        B[2 * nd + 2 * q + 1] = dN[0][q];
      } else {
        B[0 * nd + 3 * q] = dN[0][q];
        B[1 * nd + 3 * q + 1] = dN[1][q];
        B[2 * nd + 3 * q + 2] = dN[2][q];
        B[3 * nd + 3 * q] = dN[1][q]; B[3 * nd + 3 * q + 1] = dN[0][q];
        B[4 * nd + 3 * q + 1] = dN[2][q]; B[4 * nd + 3 * q + 2] = dN[1][q];
        B[5 * nd + 3 * q] = dN[2][q]; B[5 * nd + 3 * q + 2] = dN[0][q];
      }
    }
    for (int i = 0; i < nd; ++i)
      for (int j = 0; j < nd; ++j) {
        double s = 0.0;
        for (int v = 0; v < voigt; ++v)
          for (int w = 0; w < voigt; ++w) s += B[v * nd + i] * D[v * voigt + w] * B[w * nd + j];
        K[i * nd + j] += s * jac;
      }
    vol += jac;
  }
  return vol;
}

// Helmholtz element matrix  sum (N N^T + r_h^2 dN^T dN) jac, r_h = R/(2 sqrt 3)
// (src/optimization/filter_helm.hpp:178-222)
void helm_K0(const BoxElement &el, double R, std::vector<double> &K) {
  const int nq = el.nq;
  K.assign(nq * nq, 0.0);
  const double rh2 = std::pow(R / (2 * std::sqrt(3.0)), 2);
  for (int ip = 0; ip < nq; ++ip) {
    double N[8], dN[3][8], jac;
    el.eval(ip, N, dN, jac);
    for (int a = 0; a < nq; ++a)
      for (int b = 0; b < nq; ++b) {
        double g = 0.0;
        for (int d = 0; d < el.ndim; ++d) g += dN[d][a] * rh2 * dN[d][b];
        K[a * nq + b] += (N[a] * N[b] + g) * jac;
      }
  }
}

inline double haar_sign(int sigma, int q, int ndim) {
  double s = 1.0;
  for (int d = 0; d < ndim; ++d)
    if ((sigma >> d) & 1) s *= ((q >> d) & 1) ? 1.0 : -1.0;
  return s;
}

constexpr int popc(int v) { return (v & 1) + ((v >> 1) & 1) + ((v >> 2) & 1); }
bool host_nz3(int chi, int c, int cc) {
  const int pc = popc(chi);
  if (pc == 0 || pc == 3) return true;
  if (pc == 1) {
    const int d = (chi & 1) ? 0 : ((chi & 2) ? 1 : 2);
    return c != d && cc != d;
  }
  const int d3 = !(chi & 1) ? 0 : (!(chi & 2) ? 1 : 2);
  return (c == d3) == (cc == d3);
}
bool host_nz2(int chi, int c, int cc) {
  if (chi == 0 || chi == 3) return true;
  const int d = (chi & 1) ? 0 : 1;
  return c != d && cc != d;
}

// Haar-basis blocks of K0; fails if the element matrix is not block diagonal
bool build_elastic_op(const BoxElement &el, const std::vector<double> &K, ElasticOp<double> &op,
                      std::string &why) {
  const int nq = el.nq, nc = el.ndim, nd = nq * nc;
  std::vector<double> Kh(nd * nd, 0.0);  // index (c*nq + sigma)
  for (int c = 0; c < nc; ++c)
    for (int s = 0; s < nq; ++s)
      for (int c2 = 0; c2 < nc; ++c2)
        for (int s2 = 0; s2 < nq; ++s2) {
          double v = 0.0;
          for (int q = 0; q < nq; ++q)
            for (int q2 = 0; q2 < nq; ++q2)
              v += haar_sign(s, q, nc) * K[(q * nc + c) * nd + q2 * nc + c2] * haar_sign(s2, q2, nc);
          Kh[(c * nq + s) * nd + c2 * nq + s2] = v;
        }
  double mx = 0.0;
  for (double v : Kh) mx = std::max(mx, std::fabs(v));
  const double tol = 1e-12 * mx;
  std::memset(&op, 0, sizeof op);
  const double scale = 1.0 / double(nq * nq);
  for (int c = 0; c < nc; ++c)
    for (int s = 0; s < nq; ++s)
      for (int c2 = 0; c2 < nc; ++c2)
        for (int s2 = 0; s2 < nq; ++s2) {
          const int chi = s ^ (1 << c), chi2 = s2 ^ (1 << c2);
          const double v = Kh[(c * nq + s) * nd + c2 * nq + s2];
          const bool nz = chi == chi2 && (nc == 3 ? host_nz3(chi, c, c2) : host_nz2(chi, c, c2));
          if (!nz) {
            if (std::fabs(v) > tol) {
              why = "element stiffness is not block diagonal in the Haar basis";
              return false;
            }
            continue;
          }
          op.blk[chi][c][c2] = v * scale;
        }
  // the blocks are symmetric (K0 is); make them exactly so -- the kernels read the
  // upper triangles only
  for (int chi = 0; chi < nq; ++chi)
    for (int c = 0; c < nc; ++c)
      for (int c2 = c + 1; c2 < nc; ++c2) {
        const double a = op.blk[chi][c][c2], b = op.blk[chi][c2][c];
        if (std::fabs(a - b) > tol * scale) {
          why = "element stiffness is not symmetric";
          return false;
        }
        op.blk[chi][c][c2] = op.blk[chi][c2][c] = 0.5 * (a + b);
      }
  for (int c = 0; c < nc; ++c) {
    op.kd[c] = K[(0 * nc + c) * nd + 0 * nc + c];
    for (int q = 1; q < nq; ++q)
      if (std::fabs(K[(q * nc + c) * nd + q * nc + c] - op.kd[c]) > 1e-12 * mx) {
        why = "element diagonal differs between nodes";
        return false;
      }
    if (!(op.kd[c] > 0.0)) {
      why = "element diagonal is not positive";
      return false;
    }
    op.ikd[c] = 1.0 / op.kd[c];
  }
  return true;
}

bool build_helm_op(const BoxElement &el, const std::vector<double> &K, HelmOp<double> &op,
                   std::string &why) {
  const int nq = el.nq;
  double mx = 0.0;
  for (double v : K) mx = std::max(mx, std::fabs(v));
  std::memset(&op, 0, sizeof op);
  for (int s = 0; s < nq; ++s)
    for (int s2 = 0; s2 < nq; ++s2) {
      double v = 0.0;
      for (int q = 0; q < nq; ++q)
        for (int q2 = 0; q2 < nq; ++q2)
          v += haar_sign(s, q, el.ndim) * K[q * nq + q2] * haar_sign(s2, q2, el.ndim);
      if (s == s2) op.kf[s] = v / double(nq * nq);
      else if (std::fabs(v) > 1e-12 * mx * nq) {
        why = "Helmholtz element matrix is not diagonal in the Haar basis";
        return false;
      }
    }
  op.kdiag = K[0];
  return true;
}

template <typename T> ElasticOp<T> cast_op(const ElasticOp<double> &a) {
  ElasticOp<T> b;
  for (int i = 0; i < 8; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k) b.blk[i][j][k] = (T)a.blk[i][j][k];
  for (int j = 0; j < 3; ++j) {
    b.kd[j] = (T)a.kd[j];
    b.ikd[j] = (T)a.ikd[j];
  }
  return b;
}

}  // namespace

// =====================================================================================
// job
// =====================================================================================
// One captured chunk of PCG iterations.  The kernels take the iteration index from
// the device state, so one graph serves every chunk of a solve (and every later solve
// of the same job); buffers and operator constants of a job never move.
struct ChunkGraph {
  hipGraphExec_t exec = nullptr;
  ChunkGraph() = default;
  ChunkGraph(const ChunkGraph &) = delete;
  ChunkGraph &operator=(const ChunkGraph &) = delete;
  ~ChunkGraph() { reset(); }
  void reset() {
    if (exec) (void)hipGraphExecDestroy(exec);
    exec = nullptr;
  }
};

struct PcgWork64 {
  DBuf<double> r, p, t, z, invd;
};
struct MixedWork {
  DBuf<float> r, pa, pb, Ap, Ap2, xlo, invd, invs, E, b32;  // Ap2: seam halves of kernel A
  DBuf<double> Kx;
  // box jobs in the two-kernel iteration: the FP32 vectors are node-interleaved (entry
  // nc n + c), the layout of kernel A; otherwise component planes
  bool aos = false;
  // single-reduction form (hexa8 solves): r = {r, r1}, p = {pa, pb}, Ap = {Ap, ap1};
  // per-block partials by launch parity
  bool sr = false;
  DBuf<float> r1, ap1;
  DBuf<double> part;
};
struct FusedWork {
  DBuf<double> r, pa, pb, Ap, invd, invs;
  // box jobs: the iteration's vectors node-interleaved (entry nc n + c), with x and b in
  // that layout (xa, ba) during the solve; x goes back to the job's component-plane
  // displacement at the stop
  DBuf<double> xa, ba;
  bool aos = false;
};
struct HelmWork {
  bool ready = false;
  double radius = 0;
  HelmOp<double> op;   // box jobs
  double rh2 = 0;      // mesh jobs: (R / (2 sqrt 3))^2
  DBuf<double> invd, b, x, r, p, t, z;
};
struct ConeWork {
  bool ready = false;
  double radius = 0;
  int n = 0;
  DBuf<int> off;
  DBuf<double> w, aa;
  // mesh jobs: neighbour lists in CSR (ptr, id, weights r - d)
  DBuf<int> ptr, id;
  DBuf<double> wt;
};
struct MMAWork {
  bool ready = false;
  int iter = 0;
  double asyminit = 0.2, asymdec = 0.7, asyminc = 1.2;  // SetAsymptotes (MMASolver.hpp:71-76)
  DBuf<double> low, upp, alpha, beta, p0, q0, p1, q1, xold1, xold2;
};
// GCMMASolver state with m = 1 (GCMMASolver.hpp:91-114): the MMA vectors (iter is
// outeriter) plus the conservative factors raa0 / raa, the constants r0 / r of the
// approximations and their values f0app / fapp at the last returned point.
struct GcmmaWork {
  MMAWork v;
  double raa0 = 1.0, raa = 1.0, r0 = 0.0, r = 0.0, f0app = 0.0, fapp = 0.0;
};
// Output step (icx_output_begin / _write): device temporaries, pinned host copies of
// the displacement (node-major), density and mises, and the event that completes them.
struct OutputWork {
  DBuf<double> ufull, nodes, mises;
  HBuf<double> h_disp, h_dens, h_mises;
  std::vector<uint32_t> geo;  // big-endian points, cells and cell types (fixed per job)
  size_t geo_points = 0, geo_cells = 0;
  hipEvent_t done = nullptr;
  bool pending = false;
  ~OutputWork() {
    if (done) (void)hipEventDestroy(done);
  }
};

struct OptWork {
  bool ready = false;
  bool in_step = false;  // between icx_opt_step_begin and icx_opt_step_end
  icx_opt_params par;
  int step = 0;
  double V0 = 0.0;
  double beta = 1.0;     // projection sharpness of the current step
  std::chrono::steady_clock::time_point t0;
  DBuf<double> x, fil, thr, dthr, dfds, dgds, tmp, tmp2;
  DBuf<double> xnew;     // GCMMA driver: the trial design
};

// ---- volume of a general quad4 / hexa8 ----------------------------------------------
// sum_ip det J at the 2 x 2 (x 2) Gauss points +-1/sqrt(3), weight 1 (the volume
// LinearElastic::makeKeFint accumulates, src/problem/elasticity.cpp:249), local nodes in
// the reference's order (src/problem/shape_function.cpp:97-119, 403-447).  Returns false
// when det J <= 0 at a Gauss point (inverted or degenerate element).
static bool general_volume(int nq, int ndim, const double *X, double &vol) {
  static const int corner[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                   {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
  const double g = 1.0 / std::sqrt(3.0);
  vol = 0.0;
  for (int ip = 0; ip < nq; ++ip) {
    double r[3] = {0, 0, 0};
    for (int d = 0; d < ndim; ++d) r[d] = corner[ip][d] * g;
    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    const double sc = ndim == 3 ? 0.125 : 0.25;
    for (int a = 0; a < nq; ++a)
      for (int d = 0; d < ndim; ++d) {
        double v = sc * corner[a][d];
        for (int e = 0; e < ndim; ++e)
          if (e != d) v *= 1.0 + corner[a][e] * r[e];
        for (int k = 0; k < ndim; ++k) J[d][k] += v * X[3 * a + k];
      }
    const double det = ndim == 2 ? J[0][0] * J[1][1] - J[0][1] * J[1][0]
                                 : J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) +
                                       J[0][1] * (J[1][2] * J[2][0] - J[1][0] * J[2][2]) +
                                       J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    if (!(det > 0.0)) return false;
    vol += det;
  }
  return true;
}

struct icx_job {
  icx_ctx *ctx = nullptr;
  icx_problem prob{};
  Grid g{};
  int nc = 0;          // dofnp
  int numeq = -1;
  BoxElement el{};
  double volume = 0.0; // element volume
  ElasticOp<double> op64{};
  ElasticOp<float> op32{};
  std::vector<float> K0f;  // hexa8 elastic: dense unit-modulus K0 (24 x 24) for the MFMA matvec
  // host-side BC state (reference Node::dirich)
  std::vector<int> has, node_flag;
  std::vector<double> node_val;
  std::vector<int> dof;
  std::vector<double> fext;
  bool any_dirich_value = false;
  // device state
  DBuf<int> d_dof;
  DBuf<uint8_t> d_mask;
  DBuf<double> d_fext, d_dirich, d_b, d_u, d_design, d_E, d_vec, d_vec2;
  DBuf<double> d_partial;   // generic reductions [3][RED_MAXBLOCKS]
  DBuf<double> d_pcgpart;   // PCG partials: PA, PB (x2), PN, PZ
  DBuf<SolverState> d_state;
  DBuf<double> d_scalars;
  DBuf<OcBisect> d_oc;      // device-side OC bisection state
  DBuf<MmaDual> d_mdual;    // device-side MMA dual loop state
  bool design_set = false;
  bool u_valid = false;
  PcgWork64 pw;
  ChunkGraph gr_plain, gr_fused, gr_mixed, gr_helm;
  void reset_graphs() {
    gr_plain.reset();
    gr_fused.reset();
    gr_mixed.reset();
    gr_helm.reset();
  }
  MixedWork mw;
  // unstructured mesh (icx_job_create_mesh): coordinates (numnp x 3), connectivity
  // (nelx x ne, the reference's local node order), element volumes; device operator
  bool mesh = false;
  std::vector<double> mx, mvol;
  std::vector<int> mconn;
  DBuf<int> d_conn, d_adj_ptr, d_slot;
  DBuf<double> d_xyz, d_ye, d_vol;
  MeshOp mop{};
  // state of a submitted (batched) solve: mixed form, or FP64 fused form
  struct Submitted {
    bool active = false;     // in the context's batch
    int form = ICX_SOLVE_MIXED;
    int was = 0;             // stop code that triggered the pending replacement
    bool replacing = false;
    double tol = 0.0;
    int maxit = 0;
    int refinements = 0;
    std::chrono::steady_clock::time_point t0;
    bool has_info = false;   // info / rc describe the last finished submitted solve
    icx_solve_info info{};
    icx_status rc = ICX_OK;
  } sub;
  FusedWork fw;
  HelmWork hw;
  ConeWork cw;
  OutputWork ow;
  // heat conduction (prob.state == ICX_HEAT): element matrices of materials 1 and 2,
  // and K1 - K2 for the sensitivity; d_E holds the weights design^p
  bool heat() const { return prob.state == ICX_HEAT; }
  HeatOp heat_op{}, heat_dK{};
  bool heat_ready = false;
  bool heat_haar_ok = false;  // both conductivity element matrices Haar-diagonal
  MMAWork mma;
  GcmmaWork gc;
  OptWork opt;

  Reduce red() const { return Reduce{d_partial.p}; }
  PcgPartials parts(int nA, int nB) const {
    PcgPartials P;
    double *b = d_pcgpart.p;
    P.PA = b;
    P.PB = b + 2 * RED_MAXBLOCKS;
    P.PN = b + 6 * RED_MAXBLOCKS;
    P.PZ = b + 8 * RED_MAXBLOCKS;
    P.nA = nA;
    P.nB = nB;
    return P;
  }
  long long nvec() const { return (long long)nc * g.nn; }
};

namespace {

hipStream_t S(icx_job *j) { return j->ctx->stream; }

// A stop recorded in `pending` is final once the launches before the copy have ended
// (the single-reduction iteration records its stop there only; the next launch of the
// two-kernel forms publishes it as `done`).
inline void settle(SolverState &st) {
  if (!st.done && st.pending) st.done = st.pending;
}

icx_status read_state(icx_job *j, SolverState &st) {
  icx_ctx *c = j->ctx;
  HIPCHK(c, hipMemcpyAsync(c->h_state, j->d_state.p, sizeof(SolverState), hipMemcpyDeviceToHost, S(j)));
  HIPCHK(c, hipStreamSynchronize(S(j)));
  st = *c->h_state;
  settle(st);
  if (c->ev_used) prof_collect(c);
  return ICX_OK;
}

icx_status read_scalars(icx_job *j, int n, double *out) {
  icx_ctx *c = j->ctx;
  HIPCHK(c, hipMemcpyAsync(c->h_scalar, j->d_scalars.p, n * sizeof(double), hipMemcpyDeviceToHost, S(j)));
  HIPCHK(c, hipStreamSynchronize(S(j)));
  for (int i = 0; i < n; ++i) out[i] = c->h_scalar[i];
  return ICX_OK;
}

icx_status upload_fext(icx_job *j) {
  icx_ctx *c = j->ctx;
  DBuf<double> tmp;
  HIPCHK(c, tmp.alloc(std::max(j->numeq, 1)));
  HIPCHK(c, hipMemcpyAsync(tmp.p, j->fext.data(), j->numeq * sizeof(double), hipMemcpyHostToDevice, S(j)));
  launch_dof_to_grid(j->g, j->nc, j->d_dof.p, j->numeq, tmp.p, nullptr, j->d_fext.p, S(j));
  HIPCHK(c, hipStreamSynchronize(S(j)));
  return ICX_OK;
}

// rhs b = fext - K u_D on the free DOFs (src/analysis/linear.hpp:62-80)
icx_status build_rhs(icx_job *j) {
  const long long N = j->nvec();
  icx_ctx *c = j->ctx;
  HIPCHK(c, hipMemcpyAsync(j->d_b.p, j->d_fext.p, N * sizeof(double), hipMemcpyDeviceToDevice, S(j)));
  if (j->any_dirich_value) {
    if (j->heat()) launch_heat_matvec(j->g, j->heat_op, j->d_E.p, nullptr, j->d_dirich.p, j->d_vec.p, nullptr, S(j));
    else if (j->mesh) launch_mesh_matvec(j->mop, j->d_E.p, nullptr, j->d_dirich.p, j->d_vec.p, nullptr, S(j));
    else launch_elastic_matvec<double>(j->g, j->op64, j->d_E.p, nullptr, j->d_dirich.p, j->d_vec.p, nullptr, S(j));
    launch_axpby(N, -1.0, j->d_vec.p, 1.0, j->d_b.p, S(j));
    launch_mask_zero(j->g, j->nc, j->d_mask.p, j->d_b.p, S(j));
  }
  return ICX_OK;
}

// Calls that change what a running submitted solve reads (its design and moduli, its
// load, the solve buffers themselves) are refused until icx_poll has finished it.
#define ICX_REFUSE_IF_SUBMITTED(j)                                                            \
  do {                                                                                        \
    if ((j)->sub.active)                                                                      \
      return fail((j)->ctx, ICX_ERR_STATE, "%s: the job has a submitted solve running (poll "  \
                  "until it finishes)", __func__);                                            \
  } while (0)

inline int poll_every(long long N) { return N > 2000000 ? 16 : (N > 100000 ? 12 : 8); }

// Launches one chunk of iterations: replays the job's captured graph, capturing it on
// first use.  Eager launches while a profile window is open (events time single kernels)
// or when graphs are switched off.
template <class F>
icx_status run_chunk(icx_job *j, ChunkGraph &g, F launch_chunk) {
  icx_ctx *c = j->ctx;
  if (!c->use_graphs || c->prof_kernel) {
    launch_chunk();
    return ICX_OK;
  }
  if (!g.exec) {
    hipGraph_t graph = nullptr;
    HIPCHK(c, hipStreamBeginCapture(S(j), hipStreamCaptureModeThreadLocal));
    launch_chunk();
    HIPCHK(c, hipStreamEndCapture(S(j), &graph));
    const hipError_t e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
      g.exec = nullptr;
      c->use_graphs = false;
      launch_chunk();
      return ICX_OK;
    }
  }
  HIPCHK(c, hipGraphLaunch(g.exec, S(j)));
  return ICX_OK;
}

// Runs chunks of a solve until the device state reports a stop and leaves that state in
// st.  Pipelined (the default): every chunk is followed on the stream by a copy of the
// state into one of two pinned slots, and the chunk after the one whose state the host
// waits for is already queued, so the device does not idle while the host polls.  A
// chunk queued after the stop costs only its launches: every kernel of a chunk returns
// at once when it reads `done`.  Eager launches under a profile window poll chunk by
// chunk.
template <class F>
icx_status run_to_stop(icx_job *j, ChunkGraph &g, F launch_chunk, SolverState &st) {
  icx_ctx *c = j->ctx;
  icx_status rc = ICX_OK;
  if (st.done) return ICX_OK;
  if (!c->pipeline || c->prof_kernel) {
    while (!st.done) {
      if ((rc = run_chunk(j, g, launch_chunk))) return rc;
      if ((rc = read_state(j, st))) return rc;
    }
    return ICX_OK;
  }
  auto post = [&](int slot) -> icx_status {
    HIPCHK(c, hipMemcpyAsync(c->h_state + slot, j->d_state.p, sizeof(SolverState), hipMemcpyDeviceToHost, S(j)));
    HIPCHK(c, hipEventRecord(c->ev_state[slot], S(j)));
    return ICX_OK;
  };
  if ((rc = run_chunk(j, g, launch_chunk)) || (rc = post(0))) return rc;
  for (int slot = 0; !st.done; slot ^= 1) {
    if ((rc = run_chunk(j, g, launch_chunk)) || (rc = post(slot ^ 1))) return rc;
    HIPCHK(c, hipEventSynchronize(c->ev_state[slot]));
    st = c->h_state[slot];
    settle(st);
  }
  return ICX_OK;
}

// ---- FP64 plain PCG: the Eigen ConjugateGradient loop ------------------------------
// One kernel per Eigen vector expression (launch_eigen_iteration_tail) after the
// operator; used for the structural parity form and for the Helmholtz filter solves.
template <class MatvecFn>
icx_status pcg_plain(icx_job *j, long long N, const double *b, double *x, const double *invd,
                     PcgWork64 &w, double tol, int maxit, MatvecFn matvec, icx_solve_info &info,
                     ChunkGraph &graph) {
  icx_ctx *c = j->ctx;
  const PcgPartials P = j->parts(0, pcg_stream_blocks(N));
  HIPCHK(c, hipMemsetAsync(x, 0, N * sizeof(double), S(j)));
  launch_eigen_setup(N, j->d_state.p, b, w.r.p, w.p.p, invd, P, tol, maxit, S(j));
  SolverState st;
  icx_status rc = read_state(j, st);
  if (rc) return rc;
  const int every = poll_every(N);
  auto chunk = [&]() {
    for (int k = 0; k < every; ++k) {
      matvec(w.p.p, w.t.p);  // tmp = A p
      launch_eigen_iteration_tail(N, j->d_state.p, x, w.r.p, w.p.p, w.t.p, w.z.p, invd, P, S(j));
    }
  };
  if ((rc = run_to_stop(j, graph, chunk, st))) return rc;
  const bool conv = st.done == 1;
  info.iterations = conv ? st.stop_iter : st.iter;
  info.matvecs = (conv && st.rhsNorm2 > 0 && st.rr < st.rhsNorm2) ? st.stop_iter + 1 : st.iter;
  if (st.rhsNorm2 == 0 || (conv && st.stop_iter == 0 && st.rr == st.rhsNorm2)) info.matvecs = 0;
  info.fp64_matvecs = info.matvecs;
  info.relres = st.rhsNorm2 > 0 ? std::sqrt(st.rr / st.rhsNorm2) : 0.0;
  return conv ? ICX_OK : ICX_ERR_NOCONV;
}

}  // namespace

// =====================================================================================
// matvec staging (icx_matvec allocates it per call, icx_matvec_noalloc takes it from the
// device's scratch array)
// =====================================================================================
namespace {

struct MatvecStaging {
  double *dof = nullptr;                           // numeq doubles: x in, y out
  float *xf = nullptr, *yf = nullptr, *Ef = nullptr;  // FP32 operator: grid vectors, moduli
};

size_t noalloc_round(size_t b) { return (b + 255) & ~size_t(255); }

icx_status matvec_check(icx_job *j, int prec, const double *x, double *y) {
  if (!j || !x || !y) return ICX_ERR_ARG;
  if (j->numeq < 0 || !j->design_set)
    return fail(j->ctx, ICX_ERR_STATE, "icx_matvec needs icx_make_dof and icx_set_design");
  if (prec != 0 && prec != 1) return fail(j->ctx, ICX_ERR_ARG, "icx_matvec: precision 0 (FP64) or 1 (FP32)");
  if (prec != 0 && j->mesh) return fail(j->ctx, ICX_ERR_ARG, "mesh jobs: FP64 operator only");
  if (prec != 0 && j->heat()) return fail(j->ctx, ICX_ERR_ARG, "heat jobs: FP64 operator only");
  return ICX_OK;
}

// y = K x through the staging buffers: every staged word is written before it is read
// (x by the copy, the FP32 operands by the conversions, y by the operator), so nothing
// has to start at zero
icx_status matvec_staged(icx_job *j, int prec, const double *x, double *y, const MatvecStaging &sg) {
  icx_ctx *c = j->ctx;
  const long long N = j->nvec();
  HIPCHK(c, hipMemcpyAsync(sg.dof, x, j->numeq * sizeof(double), hipMemcpyHostToDevice, S(j)));
  launch_dof_to_grid(j->g, j->nc, j->d_dof.p, j->numeq, sg.dof, nullptr, j->d_vec.p, S(j));
  if (j->mesh) {
    launch_mesh_matvec(j->mop, j->d_E.p, j->d_mask.p, j->d_vec.p, j->d_vec2.p, nullptr, S(j));
  } else if (j->heat()) {
    launch_heat_matvec(j->g, j->heat_op, j->d_E.p, j->d_mask.p, j->d_vec.p, j->d_vec2.p, nullptr, S(j));
  } else if (prec == 0) {
    launch_elastic_matvec<double>(j->g, j->op64, j->d_E.p, j->d_mask.p, j->d_vec.p, j->d_vec2.p, nullptr, S(j));
  } else {
    launch_convert_f64_f32(N, j->d_vec.p, sg.xf, S(j));
    launch_convert_f64_f32(j->g.ne, j->d_E.p, sg.Ef, S(j));
    ProfScope ps(c, ICX_PROF_MATVEC32);
    if (c->mfma_matvec && j->g.ndim == 3)
      launch_elastic_matvec_mfma(j->g, j->K0f.data(), sg.Ef, j->d_mask.p, sg.xf, sg.yf, S(j), ps.ev());
    else
      launch_elastic_matvec<float>(j->g, j->op32, sg.Ef, j->d_mask.p, sg.xf, sg.yf, nullptr, S(j), ps.ev());
    launch_convert_f32_f64(N, sg.yf, j->d_vec2.p, S(j));
  }
  launch_grid_to_dof(j->g, j->nc, j->d_dof.p, j->numeq, j->d_vec2.p, sg.dof, S(j));
  HIPCHK(c, hipMemcpyAsync(y, sg.dof, j->numeq * sizeof(double), hipMemcpyDeviceToHost, S(j)));
  HIPCHK(c, hipStreamSynchronize(S(j)));
  return ICX_OK;
}

}  // namespace

// =====================================================================================
// C ABI
// =====================================================================================
extern "C" {

icx_status icx_ctx_create(int device, icx_ctx **out) {
  if (!out) return ICX_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return ICX_ERR_HIP;
  if (device < 0 || device >= ndev) return ICX_ERR_ARG;
  auto *c = new icx_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || scratch_address(device, &c->scratch) != hipSuccess ||
      hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipHostMalloc((void **)&c->h_state, 2 * sizeof(SolverState), hipHostMallocDefault) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_state[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_state[1], hipEventDisableTiming) != hipSuccess ||
      hipHostMalloc((void **)&c->h_scalar, 64 * sizeof(double), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void **)&c->h_int, 16 * sizeof(int), hipHostMallocDefault) != hipSuccess) {
    (void)icx_ctx_destroy(c);
    return ICX_ERR_HIP;
  }
  *out = c;
  return ICX_OK;
}

icx_status icx_profile_begin(icx_ctx *ctx, int kernel) {
  if (!ctx) return ICX_ERR_ARG;
  if (kernel < ICX_PROF_MATVEC64 || kernel > ICX_PROF_MATVEC32) return fail(ctx, ICX_ERR_ARG, "unknown kernel id");
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  prof_collect(ctx);
  ctx->prof_kernel = kernel;
  ctx->prof_ms = 0.0;
  ctx->prof_launches = 0;
  return ICX_OK;
}

icx_status icx_set_option(icx_ctx *ctx, int option, long long value) {
  if (!ctx) return ICX_ERR_ARG;
  if (option == ICX_OPT_GRAPHS) {
    ctx->use_graphs = value != 0;
    return ICX_OK;
  }
  if (option == ICX_OPT_SINGLE_REDUCTION) {
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->single_reduction = value != 0;
    for (icx_job *j : ctx->jobs) j->reset_graphs();  // the chunk's launches change
    return ICX_OK;
  }
  if (option == ICX_OPT_PIPELINE) {
    ctx->pipeline = value != 0;
    return ICX_OK;
  }
  if (option == ICX_OPT_MMA_HOST_DUAL) {
    ctx->mma_host_dual = value != 0;
    return ICX_OK;
  }
  if (option == ICX_OPT_MFMA_MATVEC) {
    ctx->mfma_matvec = value != 0;
    return ICX_OK;
  }
  if (option == ICX_OPT_HEAT_DENSE) {
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->heat_dense = value != 0;
    for (icx_job *j : ctx->jobs) {
      j->heat_op.haar = (j->heat_haar_ok && !ctx->heat_dense) ? 1 : 0;
      j->reset_graphs();
    }
    return ICX_OK;
  }
  if (option == ICX_OPT_TILE_DEPTH) {
    if (value < 0 || value > 64) return fail(ctx, ICX_ERR_ARG, "tile depth %lld outside 0..64", value);
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    for (icx_job *j : ctx->jobs) {
      j->g.tile_depth = (int)value;
      j->reset_graphs();  // captured launches carry the old grid
    }
    ctx->tile_depth = (int)value;
    return ICX_OK;
  }
  return fail(ctx, ICX_ERR_ARG, "unknown option %d", option);
}

icx_status icx_profile_end(icx_ctx *ctx, double *seconds, long long *launches) {
  if (!ctx) return ICX_ERR_ARG;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  prof_collect(ctx);
  if (seconds) *seconds = ctx->prof_ms * 1e-3;
  if (launches) *launches = ctx->prof_launches;
  ctx->prof_kernel = 0;
  return ICX_OK;
}

icx_status icx_ctx_destroy(icx_ctx *ctx) {
  if (!ctx) return ICX_ERR_ARG;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  while (!ctx->jobs.empty()) icx_job_destroy(ctx->jobs.back());
  batch_delete(ctx->batch);
  ctx->batch = nullptr;
  for (hipEvent_t e : ctx->ev_pool) (void)hipEventDestroy(e);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  if (ctx->h_state) (void)hipHostFree(ctx->h_state);
  for (hipEvent_t e : ctx->ev_state)
    if (e) (void)hipEventDestroy(e);
  if (ctx->h_scalar) (void)hipHostFree(ctx->h_scalar);
  if (ctx->h_int) (void)hipHostFree(ctx->h_int);
  delete ctx;
  return ICX_OK;
}

const char *icx_last_error(const icx_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

icx_status icx_ctx_synchronize(icx_ctx *ctx) {
  if (!ctx) return ICX_ERR_ARG;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return ICX_OK;
}

// host BC state and the device buffers every job holds; registers the job
static icx_status job_finish(icx_ctx *ctx, icx_job *j, icx_job **out) {
  const Grid &g = j->g;
  j->has.assign(g.nn, 0);
  j->node_flag.assign(g.nn * 6, 0);
  j->node_val.assign(g.nn * 6, 0.0);
  (void)hipSetDevice(ctx->device);
  const long long N = j->nvec();
  if (j->d_partial.alloc(3 * RED_MAXBLOCKS) || j->d_pcgpart.alloc(10 * RED_MAXBLOCKS) || j->d_state.alloc(1) ||
      j->d_scalars.alloc(64) || j->d_fext.alloc(N) || j->d_b.alloc(N) || j->d_u.alloc(N) ||
      j->d_design.alloc(g.ne) || j->d_E.alloc(g.ne) || j->d_vec.alloc(N) || j->d_vec2.alloc(N) ||
      j->d_mask.alloc(g.nn)) {
    delete j;
    return fail(ctx, ICX_ERR_NOMEM, "device allocation failed");
  }
  (void)hipMemsetAsync(j->d_pcgpart.p, 0, 10 * RED_MAXBLOCKS * sizeof(double), ctx->stream);
  (void)hipMemsetAsync(j->d_state.p, 0, sizeof(SolverState), ctx->stream);
  (void)hipMemsetAsync(j->d_fext.p, 0, N * sizeof(double), ctx->stream);
  (void)hipMemsetAsync(j->d_mask.p, 0, g.nn, ctx->stream);
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) { delete j; return fail(ctx, ICX_ERR_HIP, "stream sync failed"); }
  ctx->jobs.push_back(j);
  *out = j;
  return ICX_OK;
}

icx_status icx_job_create(icx_ctx *ctx, const icx_problem *p, icx_job **out) {
  if (!ctx || !p || !out) return ICX_ERR_ARG;
  *out = nullptr;
  if (p->eltype != ICX_QUAD4 && p->eltype != ICX_HEXA8)
    return fail(ctx, ICX_ERR_ARG, "eltype must be ICX_QUAD4 or ICX_HEXA8");
  const int ndim = p->eltype == ICX_QUAD4 ? 2 : 3;
  for (int d = 0; d < ndim; ++d)
    if (p->div[d] < 1 || !(p->length[d] > 0.0))
      return fail(ctx, ICX_ERR_ARG, "division and length must be positive");
  const bool heat = p->state == ICX_HEAT;
  if (!heat && (ndim == 3) != (p->state == ICX_SOLID3D))
    return fail(ctx, ICX_ERR_ARG, "state must be ICX_SOLID3D exactly for hexa8");
  if (!heat && !(p->poisson > -1.0 && p->poisson < 0.5)) return fail(ctx, ICX_ERR_ARG, "poisson out of range");
  auto *j = new icx_job();
  j->ctx = ctx;
  j->prob = *p;
  Grid &g = j->g;
  g.ndim = ndim;
  g.nx = p->div[0]; g.ny = p->div[1]; g.nz = ndim == 3 ? p->div[2] : 1;
  g.sx = g.nx + 1; g.sy = g.ny + 1; g.sz = ndim == 3 ? g.nz + 1 : 1;
  g.nn = (long long)g.sx * g.sy * g.sz;
  g.ne = (long long)g.nx * g.ny * g.nz;
  g.tile_depth = ctx->tile_depth;
  // node vectors are addressed with 32-bit byte offsets (3 components x 8 bytes)
  if (g.nn > (1LL << 27)) { delete j; return fail(ctx, ICX_ERR_ARG, "mesh too large (more than 2^27 nodes)"); }
  j->nc = heat ? 1 : ndim;
  j->el.ndim = ndim;
  j->el.nq = ndim == 3 ? 8 : 4;
  for (int d = 0; d < 3; ++d) j->el.h[d] = d < ndim ? p->length[d] / p->div[d] : 1.0;
  if (heat) {
    j->volume = 1.0;
    for (int d = 0; d < ndim; ++d) j->volume *= j->el.h[d];
  } else {
    std::vector<double> K;
    j->volume = elastic_K0(j->el, p->state, p->poisson, K);
    std::string why;
    if (!build_elastic_op(j->el, K, j->op64, why)) { delete j; return fail(ctx, ICX_ERR_NUMERIC, "%s", why.c_str()); }
    j->op32 = cast_op<float>(j->op64);
    if (ndim == 3) j->K0f.assign(K.begin(), K.end());
  }
  return job_finish(ctx, j, out);
}

// Element volumes of a mesh job, in parallel over host threads; false with the first
// element whose Jacobian determinant is not positive in `bad`.
static bool mesh_volumes(const icx_job *j, std::vector<double> &vol, long long &bad) {
  const int nq = j->el.nq, ndim = j->g.ndim;
  const long long ne = j->g.ne;
  vol.assign(ne, 0.0);
  const unsigned nt = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
  std::vector<long long> first_bad(nt, -1);
  auto work = [&](unsigned t) {
    double X[24];
    for (long long e = t; e < ne; e += nt) {
      for (int a = 0; a < nq; ++a)
        for (int d = 0; d < 3; ++d) X[3 * a + d] = j->mx[3 * (size_t)j->mconn[e * nq + a] + d];
      double v;
      if (!general_volume(nq, ndim, X, v)) {
        if (first_bad[t] < 0) first_bad[t] = e;
        continue;
      }
      vol[e] = v;
    }
  };
  std::vector<std::thread> th;
  for (unsigned t = 1; t < nt; ++t) th.emplace_back(work, t);
  work(0);
  for (auto &t : th) t.join();
  bad = -1;
  for (long long b : first_bad)
    if (b >= 0 && (bad < 0 || b < bad)) bad = b;
  return bad < 0;
}

icx_status icx_job_get_problem(const icx_job *j, icx_problem *prob) {
  if (!j || !prob) return ICX_ERR_ARG;
  *prob = j->prob;
  return ICX_OK;
}

icx_status icx_job_create_toml(icx_ctx *ctx, const char *text, icx_job **out, int *numeq, double *design0) {
  if (!ctx || !text || !out) return ICX_ERR_ARG;
  *out = nullptr;
  icarat::TomlJob tj;
  std::string why;
  icx_status rc;
  try {
    const icarat::toml_lite::Table cfg = icarat::toml_lite::parse_string(text);
    rc = icarat::box_job_from_toml(ctx, cfg, tj, why);
  } catch (const std::exception &e) {
    rc = ICX_ERR_ARG;
    why = e.what();
  }
  if (rc) return fail(ctx, rc, "icx_job_create_toml: %s", why.c_str());
  *out = tj.job;
  if (numeq) *numeq = tj.numeq;
  if (design0) *design0 = tj.design0;
  return ICX_OK;
}

icx_status icx_job_create_mesh(icx_ctx *ctx, const icx_mesh *m, icx_job **out) {
  if (!ctx || !m || !out || !m->x || !m->conn) return ICX_ERR_ARG;
  *out = nullptr;
  if (m->eltype != ICX_QUAD4 && m->eltype != ICX_HEXA8)
    return fail(ctx, ICX_ERR_ARG, "eltype must be ICX_QUAD4 or ICX_HEXA8");
  const int ndim = m->eltype == ICX_QUAD4 ? 2 : 3, nq = m->eltype;
  if (m->numnp < 1 || m->nelx < 1) return fail(ctx, ICX_ERR_ARG, "empty mesh");
  if ((long long)m->numnp > (1LL << 27)) return fail(ctx, ICX_ERR_ARG, "mesh too large (more than 2^27 nodes)");
  // element-corner rows (e * nq + a) and their outputs (row * ndim) are 32-bit indices
  if ((long long)m->nelx * nq * ndim >= (1LL << 31))
    return fail(ctx, ICX_ERR_ARG, "mesh too large (nelx * nodes per element * dimension >= 2^31)");
  if (m->state == ICX_HEAT) return fail(ctx, ICX_ERR_ARG, "mesh jobs: elastic states only");
  if ((ndim == 3) != (m->state == ICX_SOLID3D))
    return fail(ctx, ICX_ERR_ARG, "state must be ICX_SOLID3D exactly for hexa8");
  if (!(m->poisson > -1.0 && m->poisson < 0.5)) return fail(ctx, ICX_ERR_ARG, "poisson out of range");
  for (long long k = 0; k < (long long)m->nelx * nq; ++k)
    if (m->conn[k] < 0 || m->conn[k] >= m->numnp) return fail(ctx, ICX_ERR_ARG, "connectivity entry %lld out of range", k);
  auto *j = new icx_job();
  j->ctx = ctx;
  j->mesh = true;
  icx_problem &p = j->prob;
  p.eltype = m->eltype;
  p.state = m->state;
  p.young1 = m->young1;
  p.young2 = m->young2;
  p.poisson = m->poisson;
  p.penalty = m->penalty;
  Grid &g = j->g;
  g.ndim = ndim;
  g.nn = m->numnp;
  g.ne = m->nelx;
  g.nx = g.ny = g.nz = 1;
  g.sx = m->numnp; g.sy = g.sz = 1;
  g.tile_depth = ctx->tile_depth;
  j->nc = ndim;
  j->el.ndim = ndim;
  j->el.nq = nq;
  j->mx.assign(m->x, m->x + 3 * (size_t)m->numnp);
  j->mconn.assign(m->conn, m->conn + (size_t)m->nelx * nq);
  long long bad;
  if (!mesh_volumes(j, j->mvol, bad)) {
    delete j;
    return fail(ctx, ICX_ERR_NUMERIC, "element %lld has a non-positive Jacobian determinant", bad);
  }
  // connectivity node-of-element-major, node adjacency (e * nq + a, increasing e)
  const long long ne = g.ne, nn = g.nn;
  std::vector<int> conn_t((size_t)nq * ne), ptr(nn + 1, 0), adj((size_t)nq * ne), slot((size_t)nq * ne);
  for (long long e = 0; e < ne; ++e)
    for (int a = 0; a < nq; ++a) {
      conn_t[(size_t)a * ne + e] = j->mconn[e * nq + a];
      ++ptr[j->mconn[e * nq + a] + 1];
    }
  for (long long n = 0; n < nn; ++n) ptr[n + 1] += ptr[n];
  {
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (long long e = 0; e < ne; ++e)
      for (int a = 0; a < nq; ++a) adj[fill[j->mconn[e * nq + a]]++] = (int)(e * nq + a);
  }
  // the device takes the inverse: the row of (e, a) in node order
  for (size_t k = 0; k < adj.size(); ++k) slot[adj[k]] = (int)k;
  (void)hipSetDevice(ctx->device);
  const int nd = nq * ndim;
  hipStream_t s = ctx->stream;
  std::vector<double> xyz((size_t)ndim * nn);
  for (long long n = 0; n < nn; ++n)
    for (int d = 0; d < ndim; ++d) xyz[(size_t)d * nn + n] = j->mx[3 * n + d];
  if (j->d_conn.alloc(conn_t.size()) || j->d_adj_ptr.alloc(ptr.size()) || j->d_slot.alloc(slot.size()) ||
      j->d_xyz.alloc(xyz.size()) || j->d_ye.alloc((size_t)ne * nd) || j->d_vol.alloc(ne)) {
    delete j;
    return fail(ctx, ICX_ERR_NOMEM, "device allocation failed");
  }
  if (hipMemcpyAsync(j->d_conn.p, conn_t.data(), conn_t.size() * sizeof(int), hipMemcpyHostToDevice, s) ||
      hipMemcpyAsync(j->d_adj_ptr.p, ptr.data(), ptr.size() * sizeof(int), hipMemcpyHostToDevice, s) ||
      hipMemcpyAsync(j->d_slot.p, slot.data(), slot.size() * sizeof(int), hipMemcpyHostToDevice, s) ||
      hipMemcpyAsync(j->d_xyz.p, xyz.data(), xyz.size() * sizeof(double), hipMemcpyHostToDevice, s) ||
      hipMemcpyAsync(j->d_vol.p, j->mvol.data(), ne * sizeof(double), hipMemcpyHostToDevice, s) ||
      hipStreamSynchronize(s)) {
    delete j;
    return fail(ctx, ICX_ERR_HIP, "upload of the mesh failed");
  }
  double D[36];
  int voigt;
  material_matrix(p.state, 1.0, p.poisson, D, voigt);
  j->mop = MeshOp{nq, ndim, nn, ne, j->d_conn.p, j->d_xyz.p, D[0], D[1], D[(voigt - 1) * voigt + voigt - 1],
                  j->d_adj_ptr.p, j->d_slot.p, j->d_ye.p};
  return job_finish(ctx, j, out);
}

// MeshMDPA::setParameter / generate (src/io/mesh_mdpa.hpp:84-216): the nodes of the
// "Begin Nodes" block in file order (conn refers to them by id - 1, as the reference's
// node array does), the elements of every "Begin Elements" block whose name carries the
// element tag ("2D4" quad4, "3D8" hexa8) in file order, "id property n1 .. nne" per line.
icx_status icx_mdpa_read(const char *path, int eltype, int *numnp, int *nelx, double *x, int *conn) {
  if (!path || !numnp || !nelx) return ICX_ERR_ARG;
  if (eltype != ICX_QUAD4 && eltype != ICX_HEXA8) return ICX_ERR_ARG;
  std::FILE *f = std::fopen(path, "r");
  if (!f) return ICX_ERR_ARG;
  const char *tag = eltype == ICX_QUAD4 ? "2D4" : "3D8";
  const int nq = eltype;
  std::vector<double> xs;
  std::vector<int> cs;
  int nnodes = 0, block = 0;  // block: 0 none, 1 nodes, 2 elements of the type, 3 other elements
  bool ok = true;
  char buf[4096];
  while (ok && std::fgets(buf, sizeof buf, f)) {
    std::string line(buf);
    while (!line.empty() && (line.back() == '\n' || line.back() == '\r')) line.pop_back();
    if (block == 0) {
      if (line.rfind("Begin Nodes", 0) == 0) block = 1;
      else if (line.rfind("Begin Elements", 0) == 0) block = line.find(tag) != std::string::npos ? 2 : 3;
      continue;
    }
    if ((block == 1 && line == "End Nodes") || (block >= 2 && line == "End Elements")) {
      block = 0;
      continue;
    }
    if (block == 1) {
      long id;
      double v[3];
      if (std::sscanf(line.c_str(), "%ld %lf %lf %lf", &id, &v[0], &v[1], &v[2]) != 4) { ok = false; break; }
      ++nnodes;
      xs.insert(xs.end(), v, v + 3);
    } else if (block == 2) {
      std::istringstream is(line);
      long id, prop;
      if (!(is >> id >> prop)) { ok = false; break; }
      for (int a = 0; a < nq; ++a) {
        long n;
        if (!(is >> n)) { ok = false; break; }
        cs.push_back((int)(n - 1));
      }
    }
  }
  std::fclose(f);
  if (!ok || block != 0 || nnodes == 0 || cs.empty()) return ICX_ERR_ARG;
  for (int v : cs)
    if (v < 0 || v >= nnodes) return ICX_ERR_ARG;
  *numnp = nnodes;
  *nelx = (int)(cs.size() / nq);
  if (x) std::memcpy(x, xs.data(), xs.size() * sizeof(double));
  if (conn) std::memcpy(conn, cs.data(), cs.size() * sizeof(int));
  return ICX_OK;
}

icx_status icx_job_destroy(icx_job *j) {
  if (!j) return ICX_ERR_ARG;
  icx_ctx *c = j->ctx;
  auto it = std::find(c->jobs.begin(), c->jobs.end(), j);
  if (it == c->jobs.end()) return ICX_ERR_ARG;  // not a live job of its context
  c->jobs.erase(it);
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  batch_remove(c, j);
  delete j;
  return ICX_OK;
}

icx_status icx_job_sizes(const icx_job *j, int *numnp, int *nelx, int *neq) {
  if (!j) return ICX_ERR_ARG;
  if (numnp) *numnp = (int)j->g.nn;
  if (nelx) *nelx = (int)j->g.ne;
  if (neq) *neq = (int)(j->g.nn * j->nc);
  return ICX_OK;
}

// node coordinates x = (L/div) * index (src/io/mesh_icarat.hpp:325-335, 440-455)
static void node_coord(const icx_job *j, long long n, double *x) {
  if (j->mesh) {
    for (int d = 0; d < 3; ++d) x[d] = j->mx[3 * n + d];
    return;
  }
  const Grid &g = j->g;
  const long long i = n % g.sx, jj = (n / g.sx) % g.sy, k = n / ((long long)g.sx * g.sy);
  x[0] = (j->prob.length[0] / j->prob.div[0]) * (double)i;
  x[1] = (j->prob.length[1] / j->prob.div[1]) * (double)jj;
  x[2] = g.ndim == 3 ? (j->prob.length[2] / j->prob.div[2]) * (double)k : 0.0;
}

// local corner (a,b,c) of the reference's element node order
static const int kRefCorner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                     {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

static long long elem_node(const icx_job *j, long long e, int k) {
  if (j->mesh) return j->mconn[e * j->el.nq + k];
  const Grid &g = j->g;
  const long long ex = e % g.nx, ey = (e / g.nx) % g.ny, ez = e / ((long long)g.nx * g.ny);
  const int *cc = kRefCorner[k];
  return (ex + cc[0]) + (long long)g.sx * ((ey + cc[1]) + (long long)g.sy * (ez + cc[2]));
}

icx_status icx_get_mesh(const icx_job *j, double *x, int *conn) {
  if (!j) return ICX_ERR_ARG;
  if (x)
    for (long long n = 0; n < j->g.nn; ++n) node_coord(j, n, x + 3 * n);
  if (conn) {
    const int ne = j->el.nq;
    for (long long e = 0; e < j->g.ne; ++e)
      for (int k = 0; k < ne; ++k) conn[e * ne + k] = (int)elem_node(j, e, k);
  }
  return ICX_OK;
}

icx_status icx_add_dirichlet(icx_job *j, const int *flag6, const double *val6, const double *range6) {
  if (!j || !flag6 || !val6 || !range6) return ICX_ERR_ARG;
  ICX_REFUSE_IF_SUBMITTED(j);
  if (j->numeq >= 0) return fail(j->ctx, ICX_ERR_STATE, "Dirichlet ranges must precede icx_make_dof");
  const double tol = 1.0e-7;
  const double lo[3] = {range6[0] - tol, range6[2] - tol, range6[4] - tol};
  const double hi[3] = {range6[1] + tol, range6[3] + tol, range6[5] + tol};
  for (long long n = 0; n < j->g.nn; ++n) {
    double x[3];
    node_coord(j, n, x);
    bool in = true;
    for (int d = 0; d < j->g.ndim; ++d) in = in && x[d] >= lo[d] && x[d] <= hi[d];
    if (!in) continue;
    if (!j->has[n]) {
      j->has[n] = 1;
      for (int k = 0; k < 6; ++k) { j->node_flag[6 * n + k] = 0; j->node_val[6 * n + k] = 0.0; }
    }
    for (int k = 0; k < j->nc; ++k) {
      if (j->node_flag[6 * n + k] != 1) j->node_flag[6 * n + k] = flag6[k];
      j->node_val[6 * n + k] = val6[k];
    }
  }
  return ICX_OK;
}

icx_status icx_make_dof(icx_job *j, int *numeq) {
  if (!j) return ICX_ERR_ARG;
  ICX_REFUSE_IF_SUBMITTED(j);
  const int nc = j->nc;
  const long long nn = j->g.nn;
  j->dof.assign(nn * nc, 0);
  int counter = 0;
  for (long long n = 0; n < nn; ++n)
    for (int c = 0; c < nc; ++c)
      if (!j->has[n] || j->node_flag[6 * n + c] == 0) j->dof[n * nc + c] = counter++;
  const int neq_free = counter;
  for (long long n = 0; n < nn; ++n)
    if (j->has[n])
      for (int c = 0; c < nc; ++c)
        if (j->node_flag[6 * n + c] != 0) j->dof[n * nc + c] = counter++;
  j->numeq = neq_free;
  j->fext.assign(neq_free, 0.0);
  std::vector<uint8_t> mask(nn, 0);
  std::vector<double> dir(nn * 6, 0.0);
  j->any_dirich_value = false;
  for (long long n = 0; n < nn; ++n)
    for (int c = 0; c < nc; ++c)
      if (j->dof[n * nc + c] >= neq_free) {
        mask[n] |= (uint8_t)(1u << c);
        dir[6 * n + c] = j->node_val[6 * n + c];
        if (dir[6 * n + c] != 0.0) j->any_dirich_value = true;
      }
  icx_ctx *c = j->ctx;
  HIPCHK(c, j->d_dof.alloc(nn * nc));
  HIPCHK(c, hipMemcpyAsync(j->d_dof.p, j->dof.data(), nn * nc * sizeof(int), hipMemcpyHostToDevice, S(j)));
  HIPCHK(c, hipMemcpyAsync(j->d_mask.p, mask.data(), nn, hipMemcpyHostToDevice, S(j)));
  if (j->any_dirich_value) {
    // Dirichlet values as a grid vector (zero on free DOFs)
    std::vector<double> grid(nn * nc, 0.0);
    for (long long n = 0; n < nn; ++n)
      for (int cc = 0; cc < nc; ++cc)
        if (mask[n] >> cc & 1) grid[cc * nn + n] = dir[6 * n + cc];
    HIPCHK(c, j->d_dirich.alloc(nn * 6));
    HIPCHK(c, hipMemcpyAsync(j->d_dirich.p, grid.data(), nn * nc * sizeof(double), hipMemcpyHostToDevice, S(j)));
    // keep the per-node layout too for node-value output
    j->node_val.resize(nn * 6);
  }
  HIPCHK(c, hipStreamSynchronize(S(j)));
  icx_status rc = upload_fext(j);
  if (rc) return rc;
  if (numeq) *numeq = neq_free;
  return ICX_OK;
}

icx_status icx_get_dof_map(const icx_job *j, int *dof) {
  if (!j || !dof) return ICX_ERR_ARG;
  if (j->numeq < 0) return ICX_ERR_STATE;
  std::memcpy(dof, j->dof.data(), j->dof.size() * sizeof(int));
  return ICX_OK;
}

// Neumann tractions: generateLoadMain + SurfaceLoad::generate
// (src/io/boundary_condition_icarat.hpp:300-396, src/io/boundary_condition_base.cpp:16-180).
// For each element, at the numnode-th of its nodes (local order) inside the box, the
// side is chosen by judgeLine and the side's consistent load sum N^T t |J| w is added.
icx_status icx_add_neumann(icx_job *j, const double *val3, const double *range6) {
  if (!j || !val3 || !range6) return ICX_ERR_ARG;
  ICX_REFUSE_IF_SUBMITTED(j);
  if (j->numeq < 0) return fail(j->ctx, ICX_ERR_STATE, "icx_add_neumann needs icx_make_dof first");
  const double tol = 1.0e-7;
  const double lo[3] = {range6[0] - tol, range6[2] - tol, range6[4] - tol};
  const double hi[3] = {range6[1] + tol, range6[3] + tol, range6[5] + tol};
  const int ne = j->el.nq, nd = j->g.ndim, nc = j->nc;
  const int numnode = ne == 4 ? 2 : 4;
  static const int side4[5][2] = {{0, 0}, {0, 1}, {1, 2}, {2, 3}, {3, 0}};
  static const int side8[7][4] = {{0, 0, 0, 0}, {0, 1, 2, 3}, {3, 0, 4, 7}, {0, 1, 5, 4},
                                  {1, 2, 6, 5}, {2, 3, 7, 6}, {4, 5, 6, 7}};
  const double g = 1.0 / std::sqrt(3.0);
  for (long long e = 0; e < j->g.ne; ++e) {
    int hits[8], cnt = 0;
    for (int k = 0; k < ne; ++k) {
      double x[3];
      node_coord(j, elem_node(j, e, k), x);
      bool in = true;
      for (int d = 0; d < nd; ++d) in = in && lo[d] <= x[d] && x[d] <= hi[d];
      if (!in) continue;
      hits[cnt++] = k;
      if (cnt != numnode) continue;
      int side;
      if (ne == 4) side = (k == 1) ? 1 : (k == 2 ? 2 : (hits[1] - hits[0] == 1 ? 3 : 4));
      else side = (k == 3) ? 1 : (k == 5 ? 3 : (k == 6 ? 4 : (hits[2] - hits[1] == 1 ? 2 : (hits[1] - hits[0] == 1 ? 5 : 6))));
      const int *sn = ne == 4 ? side4[side] : side8[side];
      const int ns = ne == 4 ? 2 : 4;
      double X[4][3];
      for (int a = 0; a < ns; ++a) node_coord(j, elem_node(j, e, sn[a]), X[a]);
      double fe[4][3] = {};
      if (ne == 4) {
        for (int ip = 0; ip < 2; ++ip) {
          const double r = ip == 0 ? -g : g;
          const double N[2] = {0.5 * (1.0 - r), 0.5 * (1.0 + r)};
          const double d0 = 0.5 * (X[1][0] - X[0][0]), d1 = 0.5 * (X[1][1] - X[0][1]);
          const double jac = std::sqrt(d0 * d0 + d1 * d1);
          for (int a = 0; a < 2; ++a)
            for (int cc = 0; cc < nc; ++cc) fe[a][cc] += N[a] * val3[cc] * jac;
        }
      } else {
        static const int sg[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
        static const int gp[4][2] = {{-1, -1}, {1, -1}, {-1, 1}, {1, 1}};
        for (int ip = 0; ip < 4; ++ip) {
          const double r0 = gp[ip][0] * g, r1 = gp[ip][1] * g;
          double N[4], t1[3] = {0, 0, 0}, t2[3] = {0, 0, 0};
          for (int a = 0; a < 4; ++a) {
            N[a] = 0.25 * (1.0 + sg[a][0] * r0) * (1.0 + sg[a][1] * r1);
            const double dr = 0.25 * sg[a][0] * (1.0 + sg[a][1] * r1);
            const double ds = 0.25 * sg[a][1] * (1.0 + sg[a][0] * r0);
            for (int d = 0; d < 3; ++d) { t1[d] += dr * X[a][d]; t2[d] += ds * X[a][d]; }
          }
          const double cx = t1[1] * t2[2] - t1[2] * t2[1], cy = t1[2] * t2[0] - t1[0] * t2[2],
                       cz = t1[0] * t2[1] - t1[1] * t2[0];
          const double jac = std::sqrt(cx * cx + cy * cy + cz * cz);
          for (int a = 0; a < 4; ++a)
            for (int cc = 0; cc < 3; ++cc) fe[a][cc] += N[a] * val3[cc] * jac;
        }
      }
      for (int a = 0; a < ns; ++a) {
        const long long n = elem_node(j, e, sn[a]);
        for (int cc = 0; cc < nc; ++cc) {
          const int id = j->dof[n * nc + cc];
          if (id < j->numeq) j->fext[id] += fe[a][cc];
        }
      }
    }
  }
  return upload_fext(j);
}

icx_status icx_set_fext(icx_job *j, const double *fext) {
  if (!j || !fext) return ICX_ERR_ARG;
  ICX_REFUSE_IF_SUBMITTED(j);
  if (j->numeq < 0) return fail(j->ctx, ICX_ERR_STATE, "icx_set_fext needs icx_make_dof first");
  std::memcpy(j->fext.data(), fext, j->numeq * sizeof(double));
  return upload_fext(j);
}

icx_status icx_get_fext(const icx_job *j, double *fext) {
  if (!j || !fext) return ICX_ERR_ARG;
  if (j->numeq < 0) return ICX_ERR_STATE;
  std::memcpy(fext, j->fext.data(), j->numeq * sizeof(double));
  return ICX_OK;
}

icx_status icx_add_nodal_load(icx_job *j, int node, int comp, double value) {
  if (!j) return ICX_ERR_ARG;
  ICX_REFUSE_IF_SUBMITTED(j);
  if (j->numeq < 0) return fail(j->ctx, ICX_ERR_STATE, "icx_add_nodal_load needs icx_make_dof first");
  if (node < 0 || node >= j->g.nn || comp < 0 || comp >= j->nc) return fail(j->ctx, ICX_ERR_ARG, "node/comp out of range");
  const int id = j->dof[(long long)node * j->nc + comp];
  if (id < j->numeq) j->fext[id] += value;
  return upload_fext(j);
}

icx_status icx_set_design(icx_job *j, const double *design) {
  if (!j || !design) return ICX_ERR_ARG;
  ICX_REFUSE_IF_SUBMITTED(j);
  icx_ctx *c = j->ctx;
  if (j->heat() && !j->heat_ready) return fail(c, ICX_ERR_STATE, "icx_set_conductivity first");
  // densities live in [0, 1]; NaN, infinities and values outside are refused before any
  // modulus is formed from them
  for (long long e = 0; e < j->g.ne; ++e)
    if (!(design[e] >= 0.0 && design[e] <= 1.0))
      return fail(c, ICX_ERR_NUMERIC, "icx_set_design: design[%lld] = %g is not in [0, 1]", e, design[e]);
  HIPCHK(c, hipMemcpyAsync(j->d_design.p, design, j->g.ne * sizeof(double), hipMemcpyHostToDevice, S(j)));
  if (j->heat()) launch_heat_weight(j->g.ne, j->prob.penalty, j->d_design.p, j->d_E.p, S(j));
  else launch_simp(j->g.ne, j->prob.young1, j->prob.young2, j->prob.penalty, j->d_design.p, j->d_E.p, nullptr, S(j));
  HIPCHK(c, hipStreamSynchronize(S(j)));
  j->design_set = true;
  j->u_valid = false;
  return ICX_OK;
}

// conductivity tensor of a [material] row (xx, yy, zz, xy, yz, xz): Heat::setParameter
static void heat_tensor(int ndim, const double *c6, double *D) {
  if (ndim == 2) {
    const double t[4] = {c6[0], c6[3], c6[3], c6[1]};
    std::memcpy(D, t, sizeof(t));
  } else {
    const double t[9] = {c6[0], c6[3], c6[5], c6[3], c6[1], c6[4], c6[5], c6[4], c6[2]};
    std::memcpy(D, t, sizeof(t));
  }
}
// sum_ip dN^T D dN jac in the corner order of BoxElement (src/problem/heat.cpp:95-106)
static void heat_element(const BoxElement &el, const double *D, double *K) {
  const int nq = el.nq, nd = el.ndim;
  for (int i = 0; i < nq * nq; ++i) K[i] = 0.0;
  for (int ip = 0; ip < nq; ++ip) {
    double N[8], dN[3][8], jac;
    el.eval(ip, N, dN, jac);
    for (int a = 0; a < nq; ++a)
      for (int b = 0; b < nq; ++b) {
        double t = 0.0;
        for (int x = 0; x < nd; ++x)
          for (int y = 0; y < nd; ++y) t += dN[x][a] * D[x * nd + y] * dN[y][b];
        K[a * nq + b] += t * jac;
      }
  }
}

icx_status icx_set_conductivity(icx_job *j, const double *k1, const double *k2) {
  if (!j || !k1 || !k2) return ICX_ERR_ARG;
  ICX_REFUSE_IF_SUBMITTED(j);
  if (!j->heat()) return fail(j->ctx, ICX_ERR_STATE, "icx_set_conductivity: not a heat job (ICX_HEAT)");
  if (k1[0] < k2[0])
    return fail(j->ctx, ICX_ERR_ARG, "the conductivity of material 1 must be at least that of material 2");
  double D1[9], D2[9], Dd[9];
  const int nd = j->g.ndim;
  heat_tensor(nd, k1, D1);
  heat_tensor(nd, k2, D2);
  for (int i = 0; i < nd * nd; ++i) Dd[i] = D1[i] - D2[i];
  j->heat_op.nq = j->heat_dK.nq = j->el.nq;
  heat_element(j->el, D1, j->heat_op.K1);
  heat_element(j->el, D2, j->heat_op.K2);
  heat_element(j->el, Dd, j->heat_dK.K1);
  // diagonal tensors on a box element: both element matrices Haar-diagonal (fast operator)
  {
    const int nq = j->el.nq;
    HelmOp<double> h1, h2;
    std::string why;
    const bool d1 = build_helm_op(j->el, std::vector<double>(j->heat_op.K1, j->heat_op.K1 + nq * nq), h1, why);
    const bool d2 = build_helm_op(j->el, std::vector<double>(j->heat_op.K2, j->heat_op.K2 + nq * nq), h2, why);
    j->heat_haar_ok = d1 && d2;
    j->heat_op.haar = (j->heat_haar_ok && !j->ctx->heat_dense) ? 1 : 0;
    j->reset_graphs();
    for (int q = 0; q < 8; ++q) {
      j->heat_op.kf1[q] = q < nq ? h1.kf[q] : 0.0;
      j->heat_op.kf2[q] = q < nq ? h2.kf[q] : 0.0;
    }
  }
  j->heat_ready = true;
  j->u_valid = false;
  return ICX_OK;
}

icx_status icx_matvec(icx_job *j, int prec, const double *x, double *y) {
  icx_status st = matvec_check(j, prec, x, y);
  if (st != ICX_OK) return st;
  icx_ctx *c = j->ctx;
  MatvecStaging sg;
  DBuf<double> tmp;
  DBuf<float> xf, yf, Ef;
  HIPCHK(c, tmp.alloc(std::max(j->numeq, 1)));
  sg.dof = tmp.p;
  if (prec != 0) {
    HIPCHK(c, xf.alloc(j->nvec())); HIPCHK(c, yf.alloc(j->nvec())); HIPCHK(c, Ef.alloc(j->g.ne));
    sg.xf = xf.p; sg.yf = yf.p; sg.Ef = Ef.p;
  }
  return matvec_staged(j, prec, x, y, sg);
}

icx_status icx_matvec_noalloc(icx_job *j, int prec, const double *x, double *y) {
  icx_status st = matvec_check(j, prec, x, y);
  if (st != ICX_OK) return st;
  icx_ctx *c = j->ctx;
  if (!c->scratch) return fail(c, ICX_ERR_STATE, "icx_matvec_noalloc: no scratch address for device %d", c->device);
  // carve the staging buffers out of the device's scratch array, 256-byte aligned
  const size_t sz_dof = noalloc_round(sizeof(double) * (size_t)std::max(j->numeq, 1));
  const size_t sz_vec = noalloc_round(sizeof(float) * (size_t)j->nvec());
  const size_t sz_el = noalloc_round(sizeof(float) * (size_t)j->g.ne);
  const size_t need = sz_dof + (prec != 0 ? 2 * sz_vec + sz_el : 0);
  if (need > ICX_NOALLOC_SCRATCH_BYTES)
    return fail(c, ICX_ERR_ARG,
                "icx_matvec_noalloc: this job needs %zu bytes of scratch (precision %d), the maximum is %zu",
                need, prec, (size_t)ICX_NOALLOC_SCRATCH_BYTES);
  MatvecStaging sg;
  sg.dof = reinterpret_cast<double *>(c->scratch);
  if (prec != 0) {
    sg.xf = reinterpret_cast<float *>(c->scratch + sz_dof);
    sg.yf = reinterpret_cast<float *>(c->scratch + sz_dof + sz_vec);
    sg.Ef = reinterpret_cast<float *>(c->scratch + sz_dof + 2 * sz_vec);
  }
  return matvec_staged(j, prec, x, y, sg);
}

icx_status icx_diagonal(icx_job *j, double *diag) {
  if (!j || !diag) return ICX_ERR_ARG;
  if (j->numeq < 0 || !j->design_set) return ICX_ERR_STATE;
  icx_ctx *c = j->ctx;
  if (j->heat()) launch_heat_invdiag(j->g, j->heat_op, j->d_E.p, j->d_mask.p, j->d_vec.p, S(j));
  else if (j->mesh) launch_mesh_diag(j->mop, j->d_E.p, j->d_mask.p, j->d_vec.p, true, S(j));
  else launch_elastic_invdiag(j->g, j->op64.kd, j->d_E.p, j->d_mask.p, j->d_vec.p, nullptr, S(j));
  DBuf<double> tmp;
  HIPCHK(c, tmp.alloc(std::max(j->numeq, 1)));
  launch_grid_to_dof(j->g, j->nc, j->d_dof.p, j->numeq, j->d_vec.p, tmp.p, S(j));
  HIPCHK(c, hipMemcpyAsync(diag, tmp.p, j->numeq * sizeof(double), hipMemcpyDeviceToHost, S(j)));
  HIPCHK(c, hipStreamSynchronize(S(j)));
  for (int i = 0; i < j->numeq; ++i) diag[i] = 1.0 / diag[i];
  return ICX_OK;
}

// ---- structural solves ------------------------------------------------------------------
static icx_status solve_plain(icx_job *j, double tol, int maxit, icx_solve_info &info) {
  icx_ctx *c = j->ctx;
  const long long N = j->nvec();
  PcgWork64 &w = j->pw;
  HIPCHK(c, w.r.alloc(N)); HIPCHK(c, w.p.alloc(N)); HIPCHK(c, w.t.alloc(N));
  HIPCHK(c, w.z.alloc(N)); HIPCHK(c, w.invd.alloc(N));
  if (j->heat()) launch_heat_invdiag(j->g, j->heat_op, j->d_E.p, j->d_mask.p, w.invd.p, S(j));
  else if (j->mesh) launch_mesh_diag(j->mop, j->d_E.p, j->d_mask.p, w.invd.p, true, S(j));
  else launch_elastic_invdiag(j->g, j->op64.kd, j->d_E.p, j->d_mask.p, w.invd.p, nullptr, S(j));
  auto mv = [&](const double *p, double *t) {
    ProfScope ps(c, ICX_PROF_MATVEC64);
    if (j->heat()) launch_heat_matvec(j->g, j->heat_op, j->d_E.p, j->d_mask.p, p, t, j->d_state.p, S(j), ps.ev());
    else if (j->mesh) launch_mesh_matvec(j->mop, j->d_E.p, j->d_mask.p, p, t, j->d_state.p, S(j), ps.ev());
    else launch_elastic_matvec<double>(j->g, j->op64, j->d_E.p, j->d_mask.p, p, t, j->d_state.p, S(j), ps.ev());
  };
  return pcg_plain(j, N, j->d_b.p, j->d_u.p, w.invd.p, w, tol, maxit, mv, info, j->gr_plain);
}

// Fused forms: kernel A (direction, operator, deferred x update, p.Ap) and kernel B
// (residual update, |r|^2, r.z) per iteration; the iteration index lives on the device
// (SolverState::iter, advanced by kernel B), so every chunk is the same launch sequence
// and the host only launches chunks and polls.
// FP64 fused form:
// FP64 fused form, in pieces shared by icx_solve and the batched path.
static PcgPartials fused_parts(icx_job *j) {
  if (j->mesh) return j->parts(mesh_gather_blocks(j->mop), pcg_stream_blocks(j->nvec()));
  return j->parts(elastic_fusedA_blocks<double>(j->g), pcg_stream_blocks(j->nvec()));
}

// Set-up of an FP64 fused solve: the Jacobi factor, x = 0, r = b, the partials of
// iteration 0; st receives the state after the set-up (done = 1 at once for a zero
// right-hand side).
static icx_status fused_begin(icx_job *j, double tol, int maxit, SolverState &st) {
  icx_ctx *c = j->ctx;
  const long long N = j->nvec();
  FusedWork &w = j->fw;
  HIPCHK(c, w.r.alloc(N)); HIPCHK(c, w.pa.alloc(N)); HIPCHK(c, w.pb.alloc(N));
  HIPCHK(c, w.Ap.alloc(N)); HIPCHK(c, w.invd.alloc(N)); HIPCHK(c, w.invs.alloc(j->g.nn));
  w.aos = !j->mesh;
  if (w.aos) { HIPCHK(c, w.xa.alloc(N)); HIPCHK(c, w.ba.alloc(N)); }
  const PcgPartials P = fused_parts(j);
  // mesh jobs: the plain form's Jacobi inverse (1 on constrained components)
  if (j->mesh) launch_mesh_diag(j->mop, j->d_E.p, j->d_mask.p, w.invd.p, true, S(j));
  else launch_elastic_invdiag_node<double>(j->g, j->op64, j->d_E.p, j->d_mask.p, w.invs.p, w.invd.p, S(j), w.aos);
  HIPCHK(c, hipMemsetAsync(j->d_u.p, 0, N * sizeof(double), S(j)));
  if (w.aos) {
    HIPCHK(c, hipMemsetAsync(w.xa.p, 0, N * sizeof(double), S(j)));
    launch_soa_to_aos_f64(j->g.nn, j->nc, j->d_b.p, w.ba.p, S(j));
  }
  HIPCHK(c, hipMemsetAsync(w.pa.p, 0, N * sizeof(double), S(j)));
  HIPCHK(c, hipMemsetAsync(w.pb.p, 0, N * sizeof(double), S(j)));
  launch_pcg_setup<double>(N, j->d_state.p, w.aos ? w.ba.p : j->d_b.p, w.r.p, w.invd.p, P, tol, maxit, 0.0, S(j));
  return read_state(j, st);
}

// x += alpha p still owed at a stop of the fused form (none when the set-up converged)
// the solution vector of the fused iteration: the job's displacement, or its
// node-interleaved copy
static double *fused_x(icx_job *j) { return j->fw.aos ? j->fw.xa.p : j->d_u.p; }

static void fused_flush(icx_job *j, const SolverState &st) {
  // mesh jobs update p in place in pa: both parities read it
  if (st.done != 1)
    launch_pcg_flush(j->nvec(), j->d_state.p, fused_x(j), j->fw.pa.p, j->mesh ? j->fw.pa.p : j->fw.pb.p, S(j));
  if (j->fw.aos) launch_aos_to_soa_f64(j->g.nn, j->nc, j->fw.xa.p, j->d_u.p, S(j));
}

// info of a stopped fused solve (its state read after the flush); true if converged
static bool fused_info(const SolverState &st, icx_solve_info &info) {
  info.matvecs = st.done == 1 ? 0 : st.stop_iter;
  info.fp64_matvecs = info.matvecs;
  info.iterations = st.done == 3 ? st.stop_iter - 1 : (st.done == 1 ? 0 : st.stop_iter);
  info.refinements = 0;
  info.relres = st.rhsNorm2 > 0 ? std::sqrt(st.rr / st.rhsNorm2) : 0.0;
  return st.done == 1 || st.done == 3;
}

static icx_status solve_fused(icx_job *j, double tol, int maxit, icx_solve_info &info) {
  icx_ctx *c = j->ctx;
  const long long N = j->nvec();
  FusedWork &w = j->fw;
  SolverState st;
  icx_status rc = fused_begin(j, tol, maxit, st);
  if (rc) return rc;
  const PcgPartials P = fused_parts(j);
  const int every = poll_every(N);
  auto chunk = [&]() {
    for (int k = 0; k < every; ++k) {
      if (j->mesh) {
        ProfScope ps(c, ICX_PROF_FUSED_A);
        launch_mesh_pcg_iteration<double>(j->mop, j->d_state.p, j->d_E.p, j->d_mask.p, w.r.p, w.invd.p, w.pa.p,
                                          w.Ap.p, j->d_u.p, P, S(j), ps.ev());
      } else {
        ProfScope ps(c, ICX_PROF_FUSED_A);
        launch_elastic_fusedA<double>(j->g, j->op64, j->d_E.p, j->d_mask.p, w.r.p, w.invd.p, w.invs.p, w.pa.p,
                                      w.pb.p, w.Ap.p, nullptr, fused_x(j), j->d_state.p, P, S(j), ps.ev());
      }
      {
        ProfScope ps(c, ICX_PROF_FUSED_B);
        launch_pcg_fusedB<double>(N, j->d_state.p, w.r.p, w.Ap.p, w.invd.p, P, nullptr, SeamGeom{}, S(j),
                                  ps.ev());
      }
    }
  };
  if ((rc = run_to_stop(j, j->gr_fused, chunk, st))) return rc;
  fused_flush(j, st);
  rc = read_state(j, st);
  if (rc) return rc;
  return fused_info(st, info) ? ICX_OK : ICX_ERR_NOCONV;
}

// FP32 inner PCG on the FP32 operator, FP64 x accumulation and FP64 residual
// replacement (reliable updates): whenever the FP32 recursive residual has dropped
// by a factor 10 since the last replacement (or meets the tolerance), x_hi += x_lo
// and r = b - K x_hi is recomputed with the FP64 operator; the search direction is
// kept.  Convergence is declared on the FP64 residual only.
// Mixed form, in pieces shared by icx_solve and the batched path (icx_submit/icx_poll).
static PcgPartials mixed_parts(icx_job *j) {
  if (j->mesh) return j->parts(mesh_gather_blocks(j->mop), pcg_stream_blocks(j->nvec()));
  return j->parts(elastic_fusedA_blocks<float>(j->g), pcg_stream_blocks(j->nvec()));
}
// seam layers of the job's kernel A (its tiling, as fusedA_args sets it up); none for
// mesh jobs
static SeamGeom mixed_seams(icx_job *j) {
  if (j->mesh) return SeamGeom{};
  OpArgs<float> a;
  MixedWork &w = j->mw;
  fusedA_args<float>(j->g, j->op32, w.E.p, j->d_mask.p, w.r.p, w.invd.p, w.invs.p, w.pa.p, w.pb.p, w.Ap.p, w.Ap2.p,
                     w.xlo.p,
                     j->d_state.p, mixed_parts(j), a);
  return fusedA_seams(a);
}

// Set-up of a mixed solve: FP32 operands, x = 0, r = b, partials; st receives the state
// after the set-up (done = 1 at once for a zero right-hand side).
static SrVectors sr_vectors(icx_job *j) {
  MixedWork &w = j->mw;
  SrVectors v{};
  v.r[0] = w.r.p; v.r[1] = w.r1.p;
  v.p[0] = w.pa.p; v.p[1] = w.pb.p;
  v.ap[0] = w.Ap.p; v.ap[1] = w.ap1.p;
  v.x = w.xlo.p;
  v.invd = w.invd.p;
  v.E = w.E.p;
  v.mask = j->d_mask.p;
  v.part[0] = w.part.p;
  v.part[1] = w.part.p + 6 * RED_MAXBLOCKS;
  const PcgPartials P = mixed_parts(j);
  v.pb = P.PB;
  v.npb = P.nB;
  return v;
}

// hexa8 solves take the single-reduction iteration when it is switched on
static bool use_sr(const icx_job *j) {
  return j->g.ndim == 3 && !j->heat() && !j->mesh && j->ctx->single_reduction;
}

// Set-up of a mixed solve.  sr: the single-reduction iteration (icx_sr.hip) instead of
// kernels A and B.
static icx_status mixed_begin(icx_job *j, double tol, int maxit, SolverState &st, bool sr) {
  icx_ctx *c = j->ctx;
  const long long N = j->nvec();
  MixedWork &w = j->mw;
  w.sr = sr;
  w.aos = !sr && !j->mesh;
  HIPCHK(c, w.r.alloc(N)); HIPCHK(c, w.pa.alloc(N)); HIPCHK(c, w.pb.alloc(N)); HIPCHK(c, w.Ap.alloc(N));
  HIPCHK(c, w.xlo.alloc(N)); HIPCHK(c, w.invd.alloc(N)); HIPCHK(c, w.E.alloc(j->g.ne)); HIPCHK(c, w.Kx.alloc(N));
  HIPCHK(c, w.invs.alloc(j->g.nn));
  HIPCHK(c, w.b32.alloc(N));
  if (sr) {
    HIPCHK(c, w.r1.alloc(N)); HIPCHK(c, w.ap1.alloc(N)); HIPCHK(c, w.part.alloc(12 * RED_MAXBLOCKS));
  } else if (!j->mesh) {
    HIPCHK(c, w.Ap2.alloc(N));
  }
  HIPCHK(c, hipMemsetAsync(w.pb.p, 0, N * sizeof(float), S(j)));
  const PcgPartials P = mixed_parts(j);
  launch_convert_f64_f32(j->g.ne, j->d_E.p, w.E.p, S(j));
  if (j->mesh) {  // the FP64 Jacobi inverse (1 on constrained components), rounded
    launch_mesh_diag(j->mop, j->d_E.p, j->d_mask.p, w.Kx.p, true, S(j));
    launch_convert_f64_f32(N, w.Kx.p, w.invd.p, S(j));
  } else {
    launch_elastic_invdiag_node<float>(j->g, j->op32, j->d_E.p, j->d_mask.p, w.invs.p, w.invd.p, S(j), w.aos);
  }
  HIPCHK(c, hipMemsetAsync(j->d_u.p, 0, N * sizeof(double), S(j)));
  HIPCHK(c, hipMemsetAsync(w.xlo.p, 0, N * sizeof(float), S(j)));
  HIPCHK(c, hipMemsetAsync(w.pa.p, 0, N * sizeof(float), S(j)));
  // r32 = b; the tolerance is tested on the FP64 |b| (the replacement residual is FP64)
  if (w.aos) launch_convert_f64_f32_aos(j->g.nn, j->nc, j->d_b.p, w.b32.p, S(j));
  else launch_convert_f64_f32(N, j->d_b.p, w.b32.p, S(j));
  launch_pcg_setup<float>(N, j->d_state.p, w.b32.p, w.r.p, w.invd.p, P, tol, maxit, 0.01, S(j));
  launch_dot(N, j->d_b.p, j->d_b.p, j->d_scalars.p, j->red(), S(j));
  double bb;
  icx_status rc = read_scalars(j, 1, &bb);
  if (rc) return rc;
  rc = read_state(j, st);
  if (rc) return rc;
  st.rhsNorm2 = bb;
  st.threshold = std::max(tol * tol * bb, 2.2250738585072014e-308);
  if (bb == 0.0) st.done = 1;
  st.iter2[0] = st.iter2[1] = 0;
  st.resume_iter = 0;
  st.stop_slot = 0;
  HIPCHK(c, hipMemcpyAsync(j->d_state.p, &st, sizeof(SolverState), hipMemcpyHostToDevice, S(j)));
  // w_0 = K invd r_0 and the partials of iteration 0 (nothing when done)
  if (sr) launch_sr_restart(j->g, j->op32, sr_vectors(j), j->d_state.p, S(j));
  HIPCHK(c, hipStreamSynchronize(S(j)));
  return ICX_OK;
}

// one iteration of the mixed form for one job: kernel A, kernel B
static void mixed_iteration(icx_job *j) {
  icx_ctx *c = j->ctx;
  MixedWork &w = j->mw;
  const PcgPartials P = mixed_parts(j);
  if (j->mesh) {  // one buffer holds p (updated in place)
    {
      ProfScope ps(c, ICX_PROF_FUSED_A);
      launch_mesh_pcg_iteration<float>(j->mop, j->d_state.p, w.E.p, j->d_mask.p, w.r.p, w.invd.p, w.pa.p, w.Ap.p,
                                       w.xlo.p, P, S(j), ps.ev());
    }
    ProfScope ps(c, ICX_PROF_FUSED_B);
    launch_pcg_fusedB<float>(j->nvec(), j->d_state.p, w.r.p, w.Ap.p, w.invd.p, P, nullptr, SeamGeom{}, S(j),
                             ps.ev());
    return;
  }
  {
    ProfScope ps(c, ICX_PROF_FUSED_A);
    launch_elastic_fusedA<float>(j->g, j->op32, w.E.p, j->d_mask.p, w.r.p, w.invd.p, w.invs.p, w.pa.p, w.pb.p,
                                 w.Ap.p,
                                 w.Ap2.p, w.xlo.p, j->d_state.p, P, S(j), ps.ev());
  }
  {
    ProfScope ps(c, ICX_PROF_FUSED_B);
    launch_pcg_fusedB<float>(j->nvec(), j->d_state.p, w.r.p, w.Ap.p, w.invd.p, P, w.Ap2.p, mixed_seams(j), S(j),
                             ps.ev(), (j->g.ndim == 3 && j->g.nn >= B_NODE_MIN_NODES) ? w.invs.p : nullptr,
                             j->op32.ikd);
  }
}

// residual replacement after a stop of the inner loop: x_hi += x_lo + alpha p;
// r = b - K x_hi (FP64); the finalisation sets done = 1 (converged), 0 (resume at the
// stopped iteration) or 2 (iteration limit)
static void mixed_replace(icx_job *j) {
  const long long N = j->nvec();
  MixedWork &w = j->mw;
  const PcgPartials P = mixed_parts(j);
  // single-reduction form: x is complete at the stop (no step owed)
  if (j->mesh) {  // p in place: both parities read the one buffer
    launch_refine_x(N, j->d_state.p, j->d_u.p, w.xlo.p, w.pa.p, w.pa.p, 0, j->nc, S(j));
    launch_mesh_matvec(j->mop, j->d_E.p, j->d_mask.p, j->d_u.p, w.Kx.p, nullptr, S(j));
    launch_refine_r(N, j->d_state.p, j->d_b.p, w.Kx.p, w.r.p, w.invd.p, P, 0, j->nc, S(j));
    return;
  }
  const long long nn_aos = w.aos ? j->g.nn : 0;
  launch_refine_x(N, j->d_state.p, j->d_u.p, w.xlo.p, w.pa.p, w.pb.p, nn_aos, j->nc, S(j));
  launch_elastic_matvec<double>(j->g, j->op64, j->d_E.p, j->d_mask.p, j->d_u.p, w.Kx.p, nullptr, S(j));
  launch_refine_r(N, j->d_state.p, j->d_b.p, w.Kx.p, w.r.p, w.invd.p, P, nn_aos, j->nc, S(j));
  if (w.sr) launch_sr_restart(j->g, j->op32, sr_vectors(j), j->d_state.p, S(j));
}

static void mixed_info(const SolverState &st, int refinements, icx_solve_info &info) {
  info.iterations = st.stop_iter;
  info.matvecs = st.stop_iter + refinements;
  info.fp64_matvecs = refinements;
  info.refinements = refinements;
  info.relres = st.rhsNorm2 > 0 ? std::sqrt(st.true_rr / st.rhsNorm2) : 0.0;
}

static icx_status solve_mixed(icx_job *j, double tol, int maxit, icx_solve_info &info) {
  SolverState st;
  const bool sr = use_sr(j);
  icx_status rc = mixed_begin(j, tol, maxit, st, sr);
  if (rc) return rc;
  const int every = poll_every(j->nvec());  // even: launch parities repeat chunk by chunk
  int refinements = 0;
  const SrVectors sv = sr ? sr_vectors(j) : SrVectors{};
  icx_ctx *c = j->ctx;
  auto chunk = [&]() {
    for (int k = 0; k < every; ++k) {
      if (sr) {
        ProfScope ps(c, ICX_PROF_FUSED_A);
        launch_sr_iteration(j->g, j->op32, sv, j->d_state.p, k & 1, S(j), ps.ev());
      } else {
        mixed_iteration(j);
      }
    }
  };
  while (!st.done || st.done >= 2) {
    if ((rc = run_to_stop(j, j->gr_mixed, chunk, st))) return rc;
    const int was = st.done;
    mixed_replace(j);
    ++refinements;
    rc = read_state(j, st);
    if (rc) return rc;
    if (st.done == 1) break;
    if (was == 2 || st.done == 2) { st.done = 2; break; }
  }
  mixed_info(st, refinements, info);
  return st.done == 1 ? ICX_OK : ICX_ERR_NOCONV;
}

// -------------------------------------------------------------------------------------
// Batched solves.  Submitted solves of any jobs advance together: every chunk is `every`
// iterations of [kernel A for the hexa8 jobs, kernel A for the quad4 jobs, kernel B for
// all jobs], one launch each per precision (FP32 for mixed-form jobs, FP64 for
// fused-form jobs; Batch tables in device memory), captured once per job set as a HIP
// graph.  Each job's blocks compute exactly what its own launches would (same tiles,
// same partial sums), so a job's result is bit-identical alone and in any mix; a job
// whose solve has stopped exits at once in every launch.  After each chunk the states
// of all jobs are gathered with one copy; mixed-form jobs run their replacements (FP64
// residual), fused-form jobs their owed x step at the stop; finished jobs leave the set.
// -------------------------------------------------------------------------------------
struct BatchRun {
  std::vector<icx_job *> active;
  bool dirty = true;
  DBuf<OpArgs<float>> a3, a2;
  DBuf<BArgs<float>> ab;
  DBuf<OpArgs<double>> d3, d2;  // fused-form (FP64) jobs
  DBuf<BArgs<double>> db;
  DBuf<int> offd3, offd2, offdb;
  int nd3 = 0, nd2 = 0, ndb = 0, nbd3 = 0, nbd2 = 0, nbdb = 0;
  DBuf<SrArgs> as0, as1;  // single-reduction jobs, launch parity 0 / 1
  DBuf<int> off3, off2, offb, offs;
  DBuf<SolverState *> sptr;
  DBuf<SolverState> sdev;
  SolverState *h_states = nullptr;  // pinned
  int h_cap = 0;
  int n3 = 0, n2 = 0, nab = 0, ns = 0, nb3 = 0, nb2 = 0, nbb = 0, nbs = 0, every = 8;
  ChunkGraph graph;
  ~BatchRun() {
    if (h_states) (void)hipHostFree(h_states);
  }
};

static void batch_delete(BatchRun *b) { delete b; }

static int pad8(int n) { return (n + 7) & ~7; }

static icx_status batch_rebuild(icx_ctx *c, BatchRun &B) {
  std::vector<OpArgs<float>> a3, a2;
  std::vector<BArgs<float>> ab;
  std::vector<OpArgs<double>> d3, d2;
  std::vector<BArgs<double>> db;
  std::vector<SrArgs> as0, as1;
  std::vector<int> o3{0}, o2{0}, ob{0}, os{0}, od3{0}, od2{0}, odb{0};
  std::vector<SolverState *> sp;
  long long maxN = 0;
  for (icx_job *j : B.active) {
    MixedWork &w = j->mw;
    sp.push_back(j->d_state.p);
    maxN = std::max(maxN, j->nvec());
    if (j->sub.form == ICX_SOLVE_FUSED) {
      FusedWork &f = j->fw;
      const PcgPartials P = fused_parts(j);
      OpArgs<double> a;
      const int nb = fusedA_args<double>(j->g, j->op64, j->d_E.p, j->d_mask.p, f.r.p, f.invd.p, f.invs.p, f.pa.p,
                                         f.pb.p, f.Ap.p, nullptr, fused_x(j), j->d_state.p, P, a);
      if (j->g.ndim == 3) { d3.push_back(a); od3.push_back(od3.back() + pad8(nb)); }
      else { d2.push_back(a); od2.push_back(od2.back() + pad8(nb)); }
      db.push_back(BArgs<double>{j->nvec(), j->d_state.p, f.r.p, f.Ap.p, f.invd.p, P, nullptr, SeamGeom{}, nullptr,
                                 {0.0, 0.0, 0.0}, j->g.nn});
      odb.push_back(odb.back() + pad8(P.nB));
      continue;
    }
    if (w.sr) {
      const SrVectors v = sr_vectors(j);
      as0.push_back(sr_args(j->g, j->op32, v, j->d_state.p, 0));
      as1.push_back(sr_args(j->g, j->op32, v, j->d_state.p, 1));
      os.push_back(os.back() + pad8(sr_blocks(j->g)));
      continue;
    }
    const PcgPartials P = mixed_parts(j);
    OpArgs<float> a;
    const int nb = fusedA_args<float>(j->g, j->op32, w.E.p, j->d_mask.p, w.r.p, w.invd.p, w.invs.p, w.pa.p, w.pb.p,
                                      w.Ap.p,
                                      w.Ap2.p, w.xlo.p, j->d_state.p, P, a);
    if (j->g.ndim == 3) { a3.push_back(a); o3.push_back(o3.back() + pad8(nb)); }
    else { a2.push_back(a); o2.push_back(o2.back() + pad8(nb)); }
    BArgs<float> ba{j->nvec(), j->d_state.p, w.r.p, w.Ap.p, w.invd.p, P, w.Ap2.p, fusedA_seams(a), nullptr,
                    {0.f, 0.f, 0.f}, j->g.nn};
    if (j->g.ndim == 3 && j->g.nn >= B_NODE_MIN_NODES) {  // the per-node Jacobi factor, as kernel A forms it
      ba.invs = w.invs.p;
      for (int c = 0; c < 3; ++c) ba.ikd[c] = j->op32.ikd[c];
    }
    ab.push_back(ba);
    ob.push_back(ob.back() + pad8(P.nB));
  }
  const int n = (int)B.active.size();
  B.n3 = (int)a3.size(); B.n2 = (int)a2.size(); B.nab = (int)ab.size(); B.ns = (int)as0.size();
  B.nb3 = o3.back(); B.nb2 = o2.back(); B.nbb = ob.back(); B.nbs = os.back();
  B.nd3 = (int)d3.size(); B.nd2 = (int)d2.size(); B.ndb = (int)db.size();
  B.nbd3 = od3.back(); B.nbd2 = od2.back(); B.nbdb = odb.back();
  B.every = poll_every(maxN);
  auto up = [&](auto &dbuf, const auto &vec) -> icx_status {
    if (vec.empty()) return ICX_OK;
    HIPCHK(c, dbuf.alloc(vec.size()));
    HIPCHK(c, hipMemcpyAsync(dbuf.p, vec.data(), vec.size() * sizeof(vec[0]), hipMemcpyHostToDevice, c->stream));
    return ICX_OK;
  };
  icx_status rc;
  if ((rc = up(B.a3, a3)) || (rc = up(B.a2, a2)) || (rc = up(B.ab, ab)) || (rc = up(B.off3, o3)) ||
      (rc = up(B.off2, o2)) || (rc = up(B.offb, ob)) || (rc = up(B.sptr, sp)) || (rc = up(B.as0, as0)) ||
      (rc = up(B.as1, as1)) || (rc = up(B.offs, os)) || (rc = up(B.d3, d3)) || (rc = up(B.d2, d2)) ||
      (rc = up(B.db, db)) || (rc = up(B.offd3, od3)) || (rc = up(B.offd2, od2)) || (rc = up(B.offdb, odb)))
    return rc;
  HIPCHK(c, B.sdev.alloc(std::max(n, 1)));
  if (B.h_cap < n) {
    if (B.h_states) (void)hipHostFree(B.h_states);
    B.h_states = nullptr;
    B.h_cap = 0;
    HIPCHK(c, hipHostMalloc((void **)&B.h_states, n * sizeof(SolverState), hipHostMallocDefault));
    B.h_cap = n;
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  B.graph.reset();
  B.dirty = false;
  return ICX_OK;
}

// `every` is even: the launch parities of the single-reduction jobs repeat chunk by chunk
static void batch_chunk(icx_ctx *c, BatchRun &B) {
  for (int k = 0; k < B.every; ++k) {
    launch_sr_batch((k & 1) ? B.as1.p : B.as0.p, B.offs.p, B.ns, B.nbs, c->stream);
    launch_fusedA_batch<float>(3, B.a3.p, B.off3.p, B.n3, B.nb3, c->stream);
    launch_fusedA_batch<float>(2, B.a2.p, B.off2.p, B.n2, B.nb2, c->stream);
    launch_pcg_fusedB_batch<float>(B.ab.p, B.offb.p, B.nab, B.nbb, c->stream);
    launch_fusedA_batch<double>(3, B.d3.p, B.offd3.p, B.nd3, B.nbd3, c->stream);
    launch_fusedA_batch<double>(2, B.d2.p, B.offd2.p, B.nd2, B.nbd2, c->stream);
    launch_pcg_fusedB_batch<double>(B.db.p, B.offdb.p, B.ndb, B.nbdb, c->stream);
  }
}

static icx_status batch_states(icx_ctx *c, BatchRun &B) {
  const int n = (int)B.active.size();
  launch_gather_states(B.sptr.p, B.sdev.p, n, c->stream);
  HIPCHK(c, hipMemcpyAsync(B.h_states, B.sdev.p, n * sizeof(SolverState), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int q = 0; q < n; ++q) settle(B.h_states[q]);
  return ICX_OK;
}

static void batch_finish(icx_job *j, const SolverState &st, icx_status rc) {
  auto &b = j->sub;
  b.info = icx_solve_info{};
  if (b.form == ICX_SOLVE_FUSED) fused_info(st, b.info);
  else mixed_info(st, b.refinements, b.info);
  b.info.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - b.t0).count();
  b.rc = rc;
  b.has_info = true;
  b.active = false;
  b.replacing = false;
  j->u_valid = true;
}

static void batch_remove(icx_ctx *c, icx_job *j) {
  if (!c->batch) return;
  auto &v = c->batch->active;
  auto it = std::find(v.begin(), v.end(), j);
  if (it != v.end()) {
    v.erase(it);
    c->batch->dirty = true;
  }
  j->sub.active = false;
}

icx_status icx_solve(icx_job *j, int form, double tol, int maxit, double *u, icx_solve_info *info) {
  if (!j) return ICX_ERR_ARG;
  ICX_REFUSE_IF_SUBMITTED(j);
  if (j->numeq < 0 || !j->design_set)
    return fail(j->ctx, ICX_ERR_STATE, "icx_solve needs icx_make_dof and icx_set_design");
  if (!(tol > 0.0) || maxit < 1) return fail(j->ctx, ICX_ERR_ARG, "tol must be > 0 and maxit >= 1");
  if (j->heat() && form != ICX_SOLVE_PLAIN) return fail(j->ctx, ICX_ERR_ARG, "heat jobs: ICX_SOLVE_PLAIN only");
  icx_ctx *c = j->ctx;
  (void)hipSetDevice(c->device);
  icx_solve_info inf{};
  hipEvent_t e0, e1;
  HIPCHK(c, hipEventCreate(&e0));
  HIPCHK(c, hipEventCreate(&e1));
  HIPCHK(c, hipEventRecord(e0, S(j)));
  icx_status rc = build_rhs(j);
  if (!rc) {
    if (form == ICX_SOLVE_PLAIN) rc = solve_plain(j, tol, maxit, inf);
    else if (form == ICX_SOLVE_FUSED) rc = solve_fused(j, tol, maxit, inf);
    else if (form == ICX_SOLVE_MIXED) rc = solve_mixed(j, tol, maxit, inf);
    else rc = fail(c, ICX_ERR_ARG, "unknown solve form %d", form);
  }
  (void)hipEventRecord(e1, S(j));
  (void)hipEventSynchronize(e1);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  inf.seconds = ms * 1e-3;
  if (info) *info = inf;
  if (rc && rc != ICX_ERR_NOCONV) return rc;
  j->u_valid = true;
  if (u) {
    icx_status r2 = icx_get_displacement(j, u);
    if (r2) return r2;
  }
  if (rc == ICX_ERR_NOCONV) fail(c, rc, "PCG reached maxit without meeting the tolerance");
  return rc;
}

icx_status icx_get_displacement(const icx_job *jc, double *u) {
  icx_job *j = const_cast<icx_job *>(jc);
  if (!j || !u) return ICX_ERR_ARG;
  if (!j->u_valid) return fail(j->ctx, ICX_ERR_STATE, "no solution available");
  icx_ctx *c = j->ctx;
  DBuf<double> tmp;
  HIPCHK(c, tmp.alloc(std::max(j->numeq, 1)));
  launch_grid_to_dof(j->g, j->nc, j->d_dof.p, j->numeq, j->d_u.p, tmp.p, S(j));
  HIPCHK(c, hipMemcpyAsync(u, tmp.p, j->numeq * sizeof(double), hipMemcpyDeviceToHost, S(j)));
  HIPCHK(c, hipStreamSynchronize(S(j)));
  return ICX_OK;
}

icx_status icx_get_node_values(const icx_job *jc, double *nodeval) {
  icx_job *j = const_cast<icx_job *>(jc);
  if (!j || !nodeval) return ICX_ERR_ARG;
  if (!j->u_valid) return fail(j->ctx, ICX_ERR_STATE, "no solution available");
  icx_ctx *c = j->ctx;
  DBuf<double> tmp, dir;
  HIPCHK(c, tmp.alloc(j->g.nn * 3));
  const double *dptr = nullptr;
  if (j->any_dirich_value) {
    HIPCHK(c, dir.alloc(j->g.nn * 6));
    HIPCHK(c, hipMemcpyAsync(dir.p, j->node_val.data(), j->g.nn * 6 * sizeof(double), hipMemcpyHostToDevice, S(j)));
    dptr = dir.p;
  }
  launch_node_values(j->g, j->nc, j->d_dof.p, j->numeq, j->d_u.p, dptr, tmp.p, S(j));
  HIPCHK(c, hipMemcpyAsync(nodeval, tmp.p, j->g.nn * 3 * sizeof(double), hipMemcpyDeviceToHost, S(j)));
  HIPCHK(c, hipStreamSynchronize(S(j)));
  return ICX_OK;
}

// ---- output step ------------------------------------------------------------------
// node.val of the current solution as a grid vector (free DOFs from u, prescribed
// values from the Dirichlet grid, flushed) in j->ow.ufull
static icx_status output_displacement(icx_job *j) {
  icx_ctx *c = j->ctx;
  const long long N = j->nvec();
  HIPCHK(c, j->ow.ufull.alloc(N));
  launch_full_displacement(N, j->d_u.p, j->any_dirich_value ? j->d_dirich.p : nullptr, j->ow.ufull.p, S(j));
  return ICX_OK;
}
static MisesMat mises_mat(const icx_job *j) {
  MisesMat m{};
  m.state = j->prob.state;
  m.nu = j->prob.poisson;
  for (int d = 0; d < 3; ++d) m.h[d] = j->el.h[d];
  return m;
}

icx_status icx_mises(icx_job *j, double *out) {
  if (!j || !out) return ICX_ERR_ARG;
  if (j->heat()) return fail(j->ctx, ICX_ERR_ARG, "icx_mises: elastic jobs only");
  if (!j->u_valid) return fail(j->ctx, ICX_ERR_STATE, "no solution available");
  icx_ctx *c = j->ctx;
  icx_status rc = output_displacement(j);
  if (rc) return rc;
  HIPCHK(c, j->ow.mises.alloc(j->g.ne));
  if (j->mesh) launch_mesh_mises(j->mop, j->d_E.p, j->ow.ufull.p, j->ow.mises.p, S(j));
  else launch_mises(j->g, mises_mat(j), j->d_E.p, j->ow.ufull.p, j->ow.mises.p, S(j));
  HIPCHK(c, hipMemcpyAsync(out, j->ow.mises.p, j->g.ne * sizeof(double), hipMemcpyDeviceToHost, S(j)));
  HIPCHK(c, hipStreamSynchronize(S(j)));
  return ICX_OK;
}

// The stress constraint of constraint/opt-stress-constraint/sensitivity.cpp:8-155 on the job's
// current design and solve (quad4 box jobs): g, and with dgds one adjoint solve (the FP64
// fused form of the Eigen CG loop, at tol) K lambda = sum_e Fe.  lambda takes the Dirichlet values at constrained DOFs
// (numeqToElement, src/misc/assembling.hpp:271-298); the job's displacement is restored.
static icx_status stress_constraint_dev(icx_job *j, double slim, double pn, double qp, double tol, int maxit,
                                        double *g, double *dgds, int *adjoint_iterations) {
  icx_ctx *c = j->ctx;
  if (j->heat() || j->mesh || j->g.ndim != 2)
    return fail(c, ICX_ERR_ARG, "the stress constraint is for 2D box jobs (the reference exits in 3D)");
  if (!(j->prob.young2 > j->prob.young1)) return fail(c, ICX_ERR_ARG, "young2 need larger parameter than young1");
  if (!j->u_valid) return fail(c, ICX_ERR_STATE, "no solution available");
  if (!(slim > 0.0) || !(pn > 0.0)) return fail(c, ICX_ERR_ARG, "stress_limit and p_norm must be > 0");
  const long long ne = j->g.ne, N = j->nvec();
  DBuf<double> U, sig, vm, sqp, f12, fe, usave;
  HIPCHK(c, U.alloc(N)); HIPCHK(c, sig.alloc(3 * ne)); HIPCHK(c, vm.alloc(ne)); HIPCHK(c, sqp.alloc(ne));
  icx_status rc = output_displacement(j);
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(U.p, j->ow.ufull.p, N * sizeof(double), hipMemcpyDeviceToDevice, S(j)));
  const MisesMat mm = mises_mat(j);
  const StressPar sp{qp, pn, slim, j->prob.young1, j->prob.young2, j->prob.penalty};
  launch_stress_eval(j->g, mm, sp, j->d_design.p, j->d_E.p, U.p, sig.p, vm.p, sqp.p, j->d_scalars.p, j->red(), S(j));
  double psum;
  rc = read_scalars(j, 1, &psum);
  if (rc) return rc;
  *g = std::pow(psum, 1.0 / pn) - 1.0;
  if (!dgds) return ICX_OK;
  HIPCHK(c, f12.alloc(ne)); HIPCHK(c, fe.alloc(8 * ne)); HIPCHK(c, usave.alloc(N));
  HIPCHK(c, hipMemcpyAsync(usave.p, j->d_u.p, N * sizeof(double), hipMemcpyDeviceToDevice, S(j)));
  launch_stress_adjoint_load(j->g, mm, sp, std::pow(psum, (1.0 / pn) - 1.0), j->d_design.p, j->d_E.p, sig.p, vm.p,
                             sqp.p, f12.p, fe.p, j->d_mask.p, j->d_b.p, S(j));
  icx_solve_info info{};
  rc = solve_fused(j, tol, maxit, info);  // lambda in d_u (an adjoint at maxit is kept)
  if (rc == ICX_ERR_NOCONV) rc = ICX_OK;
  if (!rc && adjoint_iterations) *adjoint_iterations = info.iterations;
  if (!rc) rc = output_displacement(j);  // lambda with the Dirichlet values
  if (!rc) launch_stress_sens(j->g, j->op64, sp, j->d_design.p, U.p, j->ow.ufull.p, f12.p, dgds, S(j));
  // the job's displacement comes back on every path: lambda never stands in for u
  const hipError_t e1 = hipMemcpyAsync(j->d_u.p, usave.p, N * sizeof(double), hipMemcpyDeviceToDevice, S(j));
  const hipError_t e2 = hipStreamSynchronize(S(j));
  if (e1 != hipSuccess || e2 != hipSuccess) {
    j->u_valid = false;
    return fail(c, ICX_ERR_HIP, "restoring the displacement after the adjoint solve failed: %s",
                hipGetErrorString(e1 != hipSuccess ? e1 : e2));
  }
  j->u_valid = true;
  return rc;
}

icx_status icx_stress_constraint(icx_job *j, double stress_limit, double p_norm, double qp, double tol, int maxit,
                                 double *g, double *dgds, int *adjoint_iterations) {
  if (!j || !g) return ICX_ERR_ARG;
  ICX_REFUSE_IF_SUBMITTED(j);
  icx_ctx *c = j->ctx;
  DBuf<double> dg;
  if (dgds) HIPCHK(c, dg.alloc(j->g.ne));
  icx_status rc = stress_constraint_dev(j, stress_limit, p_norm, qp, tol, maxit, g, dgds ? dg.p : nullptr,
                                        adjoint_iterations);
  if (rc) return rc;
  if (dgds) {
    HIPCHK(c, hipMemcpyAsync(dgds, dg.p, j->g.ne * sizeof(double), hipMemcpyDeviceToHost, S(j)));
    HIPCHK(c, hipStreamSynchronize(S(j)));
  }
  return ICX_OK;
}

icx_status icx_reactions(icx_job *j, double *out) {
  if (!j || !out) return ICX_ERR_ARG;
  if (!j->u_valid) return fail(j->ctx, ICX_ERR_STATE, "no solution available");
  icx_ctx *c = j->ctx;
  icx_status rc = output_displacement(j);
  if (rc) return rc;
  DBuf<double> ku;
  HIPCHK(c, ku.alloc(j->nvec()));
  HIPCHK(c, j->ow.nodes.alloc(j->g.nn * 3));
  // K u over all DOFs: the operator without the constraint mask
  if (j->heat()) launch_heat_matvec(j->g, j->heat_op, j->d_E.p, nullptr, j->ow.ufull.p, ku.p, nullptr, S(j));
  else if (j->mesh) launch_mesh_matvec(j->mop, j->d_E.p, nullptr, j->ow.ufull.p, ku.p, nullptr, S(j));
  else launch_elastic_matvec<double>(j->g, j->op64, j->d_E.p, nullptr, j->ow.ufull.p, ku.p, nullptr, S(j));
  launch_grid_to_nodes(j->g, j->nc, ku.p, j->ow.nodes.p, S(j));
  HIPCHK(c, hipMemcpyAsync(out, j->ow.nodes.p, j->g.nn * 3 * sizeof(double), hipMemcpyDeviceToHost, S(j)));
  HIPCHK(c, hipStreamSynchronize(S(j)));
  return ICX_OK;
}

icx_status icx_output_begin(icx_job *j) {
  if (!j) return ICX_ERR_ARG;
  if (!j->u_valid) return fail(j->ctx, ICX_ERR_STATE, "no solution available");
  icx_ctx *c = j->ctx;
  OutputWork &o = j->ow;
  icx_status rc = output_displacement(j);
  if (rc) return rc;
  HIPCHK(c, o.nodes.alloc(j->g.nn * 3));
  HIPCHK(c, o.mises.alloc(j->g.ne));
  HIPCHK(c, o.h_disp.alloc(j->g.nn * 3));
  HIPCHK(c, o.h_dens.alloc(j->g.ne));
  HIPCHK(c, o.h_mises.alloc(j->g.ne));
  if (!o.done) HIPCHK(c, hipEventCreateWithFlags(&o.done, hipEventDisableTiming));
  launch_grid_to_nodes(j->g, j->nc, o.ufull.p, o.nodes.p, S(j));
  if (j->mesh) launch_mesh_mises(j->mop, j->d_E.p, o.ufull.p, o.mises.p, S(j));
  else if (!j->heat()) launch_mises(j->g, mises_mat(j), j->d_E.p, o.ufull.p, o.mises.p, S(j));
  HIPCHK(c, hipMemcpyAsync(o.h_disp.p, o.nodes.p, j->g.nn * 3 * sizeof(double), hipMemcpyDeviceToHost, S(j)));
  HIPCHK(c, hipMemcpyAsync(o.h_dens.p, j->d_design.p, j->g.ne * sizeof(double), hipMemcpyDeviceToHost, S(j)));
  if (!j->heat())
    HIPCHK(c, hipMemcpyAsync(o.h_mises.p, o.mises.p, j->g.ne * sizeof(double), hipMemcpyDeviceToHost, S(j)));
  HIPCHK(c, hipEventRecord(o.done, S(j)));
  o.pending = true;
  return ICX_OK;
}

// Legacy VTK, binary (big-endian), with the fields of the reference's output step
// (src/io/export_vtk.hpp:47-156; opt_macro_oc.cpp:214-238): points, cells in the
// reference's node order, cell types (9 quad4, 12 hexa8), displacement (node.val),
// density and mises per element.
namespace {
inline uint32_t be_float(double v) {
  const float x = (float)v;
  uint32_t w;
  std::memcpy(&w, &x, 4);
  return __builtin_bswap32(w);
}
inline uint32_t be_int(int v) { return __builtin_bswap32((uint32_t)v); }
struct BeWriter {
  FILE *f;
  std::vector<uint32_t> buf;
  void text(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vfprintf(f, fmt, ap);
    va_end(ap);
  }
  bool block(const uint32_t *w, size_t n) { return fwrite(w, 4, n, f) == n && fputc('\n', f) != EOF; }
  bool floats(const double *v, size_t n) {
    buf.resize(n);
    for (size_t i = 0; i < n; ++i) buf[i] = be_float(v[i]);
    return block(buf.data(), n);
  }
};
}  // namespace

icx_status icx_output_write(icx_job *j, const char *path) {
  if (!j || !path) return ICX_ERR_ARG;
  icx_ctx *c = j->ctx;
  OutputWork &o = j->ow;
  if (!o.pending) return fail(c, ICX_ERR_STATE, "icx_output_write without icx_output_begin");
  HIPCHK(c, hipEventSynchronize(o.done));
  o.pending = false;
  FILE *f = std::fopen(path, "wb");
  if (!f) return fail(c, ICX_ERR_ARG, "cannot open %s for writing", path);
  BeWriter w{f, {}};
  const long long nn = j->g.nn, ne = j->g.ne;
  const int nq = j->el.nq;
  if (o.geo.empty()) {  // points, cells, cell types: built once per job
    o.geo.reserve(nn * 3 + ne * (nq + 2));
    for (long long n = 0; n < nn; ++n) {
      double x[3];
      node_coord(j, n, x);
      for (int d = 0; d < 3; ++d) o.geo.push_back(be_float(x[d]));
    }
    o.geo_points = o.geo.size();
    for (long long e = 0; e < ne; ++e) {
      o.geo.push_back(be_int(nq));
      for (int k = 0; k < nq; ++k) o.geo.push_back(be_int((int)elem_node(j, e, k)));
    }
    o.geo_cells = o.geo.size() - o.geo_points;
    for (long long e = 0; e < ne; ++e) o.geo.push_back(be_int(j->g.ndim == 3 ? 12 : 9));
  }
  const uint32_t *geo = o.geo.data();
  bool ok = true;
  w.text("# vtk DataFile Version 4.1\nvtk output\nBINARY\nDATASET UNSTRUCTURED_GRID\n");
  w.text("POINTS %lld float\n", nn);
  ok = ok && w.block(geo, o.geo_points);
  w.text("CELLS %lld %lld\n", ne, ne * (nq + 1));
  ok = ok && w.block(geo + o.geo_points, o.geo_cells);
  w.text("CELL_TYPES %lld\n", ne);
  ok = ok && w.block(geo + o.geo_points + o.geo_cells, ne);
  if (j->heat()) {
    // opt-heat/opt_heat.cpp:196-236: temperature, threshold_s (1 where the sum over a
    // node's elements of design / ne exceeds 0.5), design_s
    std::vector<double> temp(nn), nds(nn, 0.0), thr(nn);
    for (long long n = 0; n < nn; ++n) temp[n] = o.h_disp.p[3 * n];
    for (long long e = 0; e < ne; ++e)
      for (int k = 0; k < nq; ++k) nds[elem_node(j, e, k)] += o.h_dens.p[e] / nq;
    for (long long n = 0; n < nn; ++n) thr[n] = nds[n] > 0.5 ? 1.0 : 0.0;
    w.text("POINT_DATA %lld\nSCALARS temperature float\nLOOKUP_TABLE default\n", nn);
    ok = ok && w.floats(temp.data(), nn);
    w.text("SCALARS threshold_s float\nLOOKUP_TABLE default\n");
    ok = ok && w.floats(thr.data(), nn);
    w.text("CELL_DATA %lld\nSCALARS design_s float\nLOOKUP_TABLE default\n", ne);
    ok = ok && w.floats(o.h_dens.p, ne);
  } else {
    w.text("POINT_DATA %lld\nVECTORS displacement float\n", nn);
    ok = ok && w.floats(o.h_disp.p, nn * 3);
    w.text("CELL_DATA %lld\nSCALARS density float\nLOOKUP_TABLE default\n", ne);
    ok = ok && w.floats(o.h_dens.p, ne);
    w.text("SCALARS mises float\nLOOKUP_TABLE default\n");
    ok = ok && w.floats(o.h_mises.p, ne);
  }
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) return fail(c, ICX_ERR_ARG, "write to %s failed", path);
  return ICX_OK;
}

icx_status icx_write_vtk(icx_job *j, const char *path) {
  icx_status rc = icx_output_begin(j);
  if (rc) return rc;
  return icx_output_write(j, path);
}

icx_status icx_compliance(icx_job *j, double *cval) {
  if (!j || !cval) return ICX_ERR_ARG;
  if (!j->u_valid) return fail(j->ctx, ICX_ERR_STATE, "no solution available");
  launch_dot(j->nvec(), j->d_u.p, j->d_fext.p, j->d_scalars.p, j->red(), S(j));
  return read_scalars(j, 1, cval);
}

// displacement the sensitivity reads: node.val of the driver's sensitivity()
// (opt_macro_oc.cpp:193-203 gathers node[..].val, which holds prescribed values on
// constrained DOFs): the solution grid itself, or solution + Dirichlet grid
static icx_status output_displacement(icx_job *j);
static icx_status sens_displacement(icx_job *j, const double **u) {
  *u = j->d_u.p;
  if (!j->any_dirich_value) return ICX_OK;
  icx_status rc = output_displacement(j);
  *u = j->ow.ufull.p;
  return rc;
}

icx_status icx_sensitivity(icx_job *j, double vscale, double *dfds, double *dgds) {
  if (!j) return ICX_ERR_ARG;
  if (!j->u_valid) return fail(j->ctx, ICX_ERR_STATE, "no solution available");
  icx_ctx *c = j->ctx;
  DBuf<double> a, b;
  HIPCHK(c, a.alloc(j->g.ne));
  HIPCHK(c, b.alloc(j->g.ne));
  const double *u;
  icx_status rc = sens_displacement(j, &u);
  if (rc) return rc;
  if (j->heat())
    launch_heat_sensitivity(j->g, j->heat_dK, j->prob.penalty, j->d_design.p, u, j->volume, vscale, a.p, b.p, S(j));
  else if (j->mesh)
    launch_mesh_sensitivity(j->mop, j->prob.young1, j->prob.young2, j->prob.penalty, j->d_design.p, u, j->d_vol.p,
                            vscale, a.p, b.p, S(j));
  else
    launch_sensitivity(j->g, j->op64, j->prob.young1, j->prob.young2, j->prob.penalty, j->d_design.p, u,
                       j->volume, vscale, a.p, b.p, S(j));
  if (dfds) HIPCHK(c, hipMemcpyAsync(dfds, a.p, j->g.ne * sizeof(double), hipMemcpyDeviceToHost, S(j)));
  if (dgds) HIPCHK(c, hipMemcpyAsync(dgds, b.p, j->g.ne * sizeof(double), hipMemcpyDeviceToHost, S(j)));
  HIPCHK(c, hipStreamSynchronize(S(j)));
  return ICX_OK;
}

icx_status icx_element_volume(const icx_job *j, double *vol) {
  if (!j || !vol) return ICX_ERR_ARG;
  if (j->mesh) std::memcpy(vol, j->mvol.data(), j->g.ne * sizeof(double));
  else
    for (long long e = 0; e < j->g.ne; ++e) vol[e] = j->volume;
  return ICX_OK;
}

}  // extern "C"

// =====================================================================================
// filters, projection, updates, design loop
// =====================================================================================
namespace {

// Filter::preNeighbour (src/optimization/filter.hpp:191-234) on an explicit mesh: the
// neighbours of element i are the elements whose centres (the mean of their node
// coordinates, summed as x / ne in node order) lie within the radius, in increasing id,
// with weight r - d; candidates come from a cell grid of cell size r.  aa_i = sum of the
// weights in that order (Filter::densityFilter's normalisation).
static void mesh_neighbours(const icx_job *j, double radius, std::vector<int> &ptr, std::vector<int> &id,
                            std::vector<double> &wt, std::vector<double> &aa) {
  const int nq = j->el.nq, nd = j->g.ndim;
  const long long ne = j->g.ne;
  std::vector<double> cen(3 * ne, 0.0);
  for (long long e = 0; e < ne; ++e)
    for (int d = 0; d < nd; ++d) {
      double v = 0.0;
      for (int a = 0; a < nq; ++a) v += j->mx[3 * (size_t)j->mconn[e * nq + a] + d] / nq;
      cen[3 * e + d] = v;
    }
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  for (int d = 0; d < nd; ++d) {
    lo[d] = hi[d] = cen[d];
    for (long long e = 0; e < ne; ++e) {
      lo[d] = std::min(lo[d], cen[3 * e + d]);
      hi[d] = std::max(hi[d], cen[3 * e + d]);
    }
  }
  int ncell[3] = {1, 1, 1};
  for (int d = 0; d < nd; ++d) ncell[d] = std::max(1, std::min(1 << 20, (int)((hi[d] - lo[d]) / radius) + 1));
  auto cell_of = [&](long long e, int *q) {
    for (int d = 0; d < 3; ++d) {
      q[d] = d < nd ? (int)((cen[3 * e + d] - lo[d]) / radius) : 0;
      q[d] = std::max(0, std::min(ncell[d] - 1, q[d]));
    }
  };
  const long long nc = (long long)ncell[0] * ncell[1] * ncell[2];
  std::vector<int> cptr(nc + 1, 0), cid(ne);
  std::vector<long long> ecell(ne);
  for (long long e = 0; e < ne; ++e) {
    int q[3];
    cell_of(e, q);
    ecell[e] = q[0] + (long long)ncell[0] * (q[1] + (long long)ncell[1] * q[2]);
    ++cptr[ecell[e] + 1];
  }
  for (long long k = 0; k < nc; ++k) cptr[k + 1] += cptr[k];
  {
    std::vector<int> fill(cptr.begin(), cptr.end() - 1);
    for (long long e = 0; e < ne; ++e) cid[fill[ecell[e]]++] = (int)e;  // increasing id per cell
  }
  const unsigned nt = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
  std::vector<std::vector<int>> rid(ne);
  std::vector<std::vector<double>> rdist(ne);
  auto work = [&](unsigned t) {
    std::vector<std::pair<int, double>> cand;
    for (long long i = t; i < ne; i += nt) {
      int q[3];
      cell_of(i, q);
      cand.clear();
      for (int dz = (nd == 3 ? -1 : 0); dz <= (nd == 3 ? 1 : 0); ++dz)
        for (int dy = -1; dy <= 1; ++dy)
          for (int dx = -1; dx <= 1; ++dx) {
            const int x = q[0] + dx, y = q[1] + dy, z = q[2] + dz;
            if (x < 0 || y < 0 || z < 0 || x >= ncell[0] || y >= ncell[1] || z >= ncell[2]) continue;
            const long long k = x + (long long)ncell[0] * (y + (long long)ncell[1] * z);
            for (int m = cptr[k]; m < cptr[k + 1]; ++m) {
              const int jj = cid[m];
              double s2 = 0.0;
              for (int d = 0; d < nd; ++d) {
                const double t2 = cen[3 * i + d] - cen[3 * (size_t)jj + d];
                s2 += t2 * t2;
              }
              const double dist = std::sqrt(s2);
              if (dist <= radius) cand.emplace_back(jj, dist);
            }
          }
      std::sort(cand.begin(), cand.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
      for (const auto &pr : cand) {
        rid[i].push_back(pr.first);
        rdist[i].push_back(pr.second);
      }
    }
  };
  std::vector<std::thread> th;
  for (unsigned t = 1; t < nt; ++t) th.emplace_back(work, t);
  work(0);
  for (auto &t : th) t.join();
  ptr.assign(ne + 1, 0);
  for (long long i = 0; i < ne; ++i) ptr[i + 1] = ptr[i] + (int)rid[i].size();
  id.resize(ptr[ne]);
  wt.resize(ptr[ne]);
  aa.assign(ne, 0.0);
  for (long long i = 0; i < ne; ++i) {
    double a = 0.0;
    for (size_t k = 0; k < rid[i].size(); ++k) {
      id[ptr[i] + k] = rid[i][k];
      wt[ptr[i] + k] = radius - rdist[i][k];
      a += radius - rdist[i][k];
    }
    aa[i] = a;
  }
}

icx_status cone_setup(icx_job *j, double radius) {
  icx_ctx *c = j->ctx;
  ConeWork &w = j->cw;
  if (w.ready && w.radius == radius) return ICX_OK;
  if (j->mesh) {
    std::vector<int> ptr, id;
    std::vector<double> wt, aa;
    mesh_neighbours(j, radius, ptr, id, wt, aa);
    for (double a : aa)
      if (a < 1.0e-10) return fail(c, ICX_ERR_NUMERIC, "cone filter: an element has no neighbour weight (radius too small)");
    HIPCHK(c, w.ptr.alloc(ptr.size())); HIPCHK(c, w.id.alloc(std::max<size_t>(id.size(), 1)));
    HIPCHK(c, w.wt.alloc(std::max<size_t>(wt.size(), 1))); HIPCHK(c, w.aa.alloc(aa.size()));
    HIPCHK(c, hipMemcpyAsync(w.ptr.p, ptr.data(), ptr.size() * sizeof(int), hipMemcpyHostToDevice, S(j)));
    HIPCHK(c, hipMemcpyAsync(w.id.p, id.data(), id.size() * sizeof(int), hipMemcpyHostToDevice, S(j)));
    HIPCHK(c, hipMemcpyAsync(w.wt.p, wt.data(), wt.size() * sizeof(double), hipMemcpyHostToDevice, S(j)));
    HIPCHK(c, hipMemcpyAsync(w.aa.p, aa.data(), aa.size() * sizeof(double), hipMemcpyHostToDevice, S(j)));
    HIPCHK(c, hipStreamSynchronize(S(j)));
    w.n = (int)id.size();
    w.ready = true;
    w.radius = radius;
    return ICX_OK;
  }
  const Grid &g = j->g;
  const double *h = j->el.h;
  const int R[3] = {(int)std::ceil(radius / h[0]) + 1, (int)std::ceil(radius / h[1]) + 1,
                    g.ndim == 3 ? (int)std::ceil(radius / h[2]) + 1 : 0};
  std::vector<int> off;
  std::vector<double> wt;
  // increasing neighbour id = lexicographic (dz, dy, dx)
  for (int dz = -R[2]; dz <= R[2]; ++dz)
    for (int dy = -R[1]; dy <= R[1]; ++dy)
      for (int dx = -R[0]; dx <= R[0]; ++dx) {
        const double a = dx * h[0], b = dy * h[1], cc = dz * h[2];
        double s = a * a + b * b;
        if (g.ndim == 3) s += cc * cc;
        const double d = std::sqrt(s);
        if (d <= radius) {
          off.push_back(dx); off.push_back(dy); off.push_back(dz);
          wt.push_back(radius - d);
        }
      }
  w.n = (int)wt.size();
  HIPCHK(c, w.off.alloc(off.size()));
  HIPCHK(c, w.w.alloc(wt.size()));
  HIPCHK(c, w.aa.alloc(g.ne));
  HIPCHK(c, hipMemcpyAsync(w.off.p, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice, S(j)));
  HIPCHK(c, hipMemcpyAsync(w.w.p, wt.data(), wt.size() * sizeof(double), hipMemcpyHostToDevice, S(j)));
  launch_cone_aa(g, Stencil{w.n, w.off.p, w.w.p}, w.aa.p, S(j));
  HIPCHK(c, hipStreamSynchronize(S(j)));
  w.ready = true;
  w.radius = radius;
  return ICX_OK;
}

Stencil stencil(icx_job *j) { return Stencil{j->cw.n, j->cw.off.p, j->cw.w.p}; }
ConeCsr cone_csr(icx_job *j) { return ConeCsr{j->cw.ptr.p, j->cw.id.p, j->cw.wt.p}; }

// forward (adjoint = 0) or adjoint cone filter of device vectors
void cone_apply(icx_job *j, int adjoint, const double *in, double *out) {
  if (j->mesh) launch_cone_csr(j->g.ne, cone_csr(j), j->cw.aa.p, in, out, adjoint, S(j));
  else if (adjoint) launch_cone_adjoint(j->g, stencil(j), j->cw.aa.p, in, out, S(j));
  else launch_cone_forward(j->g, stencil(j), j->cw.aa.p, in, out, S(j));
}

icx_status helm_setup(icx_job *j, double radius) {
  icx_ctx *c = j->ctx;
  HelmWork &w = j->hw;
  if (w.ready && w.radius == radius) return ICX_OK;
  j->gr_helm.reset();  // captured launches carry the operator of the previous radius
  if (!j->mesh) {
    std::vector<double> K;
    helm_K0(j->el, radius, K);
    std::string why;
    if (!build_helm_op(j->el, K, w.op, why)) return fail(c, ICX_ERR_NUMERIC, "%s", why.c_str());
  }
  const long long nn = j->g.nn;
  HIPCHK(c, w.invd.alloc(nn)); HIPCHK(c, w.b.alloc(nn)); HIPCHK(c, w.x.alloc(nn));
  HIPCHK(c, w.r.alloc(nn)); HIPCHK(c, w.p.alloc(nn)); HIPCHK(c, w.t.alloc(nn)); HIPCHK(c, w.z.alloc(nn));
  if (j->mesh) {
    w.rh2 = std::pow(radius / (2.0 * std::sqrt(3.0)), 2);  // filter_helm.hpp:204
    launch_mesh_helm_invdiag(j->mop, w.rh2, w.invd.p, S(j));
  } else {
    launch_helm_invdiag(j->g, w.op.kdiag, w.invd.p, S(j));
  }
  HIPCHK(c, hipStreamSynchronize(S(j)));
  w.ready = true;
  w.radius = radius;
  return ICX_OK;
}

// device-to-device Helmholtz filter: in/out are element vectors on the device
icx_status helm_apply_dev(icx_job *j, int sens, double tol, const double *din, double *dout, int *iters) {
  HelmWork &w = j->hw;
  const double nq = j->el.nq;
  if (j->mesh) {
    launch_mesh_helm_rhs(j->mop, sens ? nq : 1.0, din, w.b.p, S(j));
  } else {
    const double rhs_scale = sens ? j->volume / (nq * nq) : j->volume / nq;
    launch_helm_rhs(j->g, rhs_scale, din, w.b.p, S(j));
  }
  PcgWork64 pw;  // borrow the Helmholtz buffers
  pw.r.p = w.r.p; pw.p.p = w.p.p; pw.t.p = w.t.p; pw.z.p = w.z.p;
  auto mv = [&](const double *p, double *t) {
    if (j->mesh) launch_mesh_helm_matvec(j->mop, w.rh2, p, t, j->d_state.p, S(j));
    else launch_helm_matvec(j->g, w.op, p, t, j->d_state.p, S(j));
  };
  icx_solve_info info{};
  icx_status rc = pcg_plain(j, j->g.nn, w.b.p, w.x.p, w.invd.p, pw, tol, 10000, mv, info, j->gr_helm);
  pw.r.p = pw.p.p = pw.t.p = pw.z.p = nullptr;  // not owned
  if (rc && rc != ICX_ERR_NOCONV) return rc;
  if (j->mesh) launch_mesh_helm_to_element(j->mop, sens ? 1.0 : nq, w.x.p, dout, S(j));
  else launch_helm_to_element(j->g, sens ? 1.0 : 1.0 / nq, w.x.p, dout, S(j));
  if (iters) *iters = info.iterations;
  return ICX_OK;
}

}  // namespace

extern "C" {

icx_status icx_cone_filter_setup(icx_job *j, double radius) {
  if (!j || !(radius > 0.0)) return ICX_ERR_ARG;
  return cone_setup(j, radius);
}

icx_status icx_cone_filter(icx_job *j, int adjoint, const double *in, double *out) {
  if (!j || !in || !out) return ICX_ERR_ARG;
  if (!j->cw.ready) return fail(j->ctx, ICX_ERR_STATE, "icx_cone_filter_setup first");
  icx_ctx *c = j->ctx;
  const long long ne = j->g.ne;
  DBuf<double> a, b;
  HIPCHK(c, a.alloc(ne)); HIPCHK(c, b.alloc(ne));
  HIPCHK(c, hipMemcpyAsync(a.p, in, ne * sizeof(double), hipMemcpyHostToDevice, S(j)));
  cone_apply(j, adjoint, a.p, b.p);
  HIPCHK(c, hipMemcpyAsync(out, b.p, ne * sizeof(double), hipMemcpyDeviceToHost, S(j)));
  HIPCHK(c, hipStreamSynchronize(S(j)));
  return ICX_OK;
}

icx_status icx_sensitivity_filter(icx_job *j, const double *design, const double *dfds, double *out) {
  if (!j || !design || !dfds || !out) return ICX_ERR_ARG;
  if (!j->cw.ready) return fail(j->ctx, ICX_ERR_STATE, "icx_cone_filter_setup first");
  icx_ctx *c = j->ctx;
  const long long ne = j->g.ne;
  DBuf<double> x, df, o;
  HIPCHK(c, x.alloc(ne)); HIPCHK(c, df.alloc(ne)); HIPCHK(c, o.alloc(ne));
  HIPCHK(c, hipMemcpyAsync(x.p, design, ne * sizeof(double), hipMemcpyHostToDevice, S(j)));
  HIPCHK(c, hipMemcpyAsync(df.p, dfds, ne * sizeof(double), hipMemcpyHostToDevice, S(j)));
  if (j->mesh) launch_cone_sensitivity_csr(ne, cone_csr(j), j->cw.aa.p, x.p, df.p, j->d_vol.p, o.p, S(j));
  else launch_cone_sensitivity(j->g, stencil(j), j->cw.aa.p, x.p, df.p, j->volume, o.p, S(j));
  HIPCHK(c, hipMemcpyAsync(out, o.p, ne * sizeof(double), hipMemcpyDeviceToHost, S(j)));
  HIPCHK(c, hipStreamSynchronize(S(j)));
  return ICX_OK;
}

icx_status icx_helmholtz_setup(icx_job *j, double radius) {
  if (!j || !(radius > 0.0)) return ICX_ERR_ARG;
  return helm_setup(j, radius);
}

icx_status icx_helmholtz_filter(icx_job *j, int sens, double tol, const double *in, double *out, int *iters) {
  if (!j || !in || !out) return ICX_ERR_ARG;
  if (!j->hw.ready) return fail(j->ctx, ICX_ERR_STATE, "icx_helmholtz_setup first");
  icx_ctx *c = j->ctx;
  const long long ne = j->g.ne;
  DBuf<double> a, b;
  HIPCHK(c, a.alloc(ne)); HIPCHK(c, b.alloc(ne));
  HIPCHK(c, hipMemcpyAsync(a.p, in, ne * sizeof(double), hipMemcpyHostToDevice, S(j)));
  icx_status rc = helm_apply_dev(j, sens, tol, a.p, b.p, iters);
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(out, b.p, ne * sizeof(double), hipMemcpyDeviceToHost, S(j)));
  HIPCHK(c, hipStreamSynchronize(S(j)));
  return ICX_OK;
}

icx_status icx_project(icx_job *j, double beta, double T, const double *in, double *out, double *dout) {
  if (!j || !in) return ICX_ERR_ARG;
  icx_ctx *c = j->ctx;
  const long long ne = j->g.ne;
  DBuf<double> a, b, d;
  HIPCHK(c, a.alloc(ne)); HIPCHK(c, b.alloc(ne)); HIPCHK(c, d.alloc(ne));
  HIPCHK(c, hipMemcpyAsync(a.p, in, ne * sizeof(double), hipMemcpyHostToDevice, S(j)));
  launch_project(ne, beta, T, a.p, out ? b.p : nullptr, dout ? d.p : nullptr, S(j));
  if (out) HIPCHK(c, hipMemcpyAsync(out, b.p, ne * sizeof(double), hipMemcpyDeviceToHost, S(j)));
  if (dout) HIPCHK(c, hipMemcpyAsync(dout, d.p, ne * sizeof(double), hipMemcpyDeviceToHost, S(j)));
  HIPCHK(c, hipStreamSynchronize(S(j)));
  return ICX_OK;
}

}  // extern "C"

namespace {

// OC bisection with the volume evaluated on the device (OCSolver.hpp:20-68).  Cone filter
// and no filter: the bisection itself runs on the device (OcBisect: each step is the
// fused volume kernel, its reduction and the bracket update), launched in rounds of
// steps with one read of the state per round; a step queued past the stop exits at once.
// Helmholtz filter: a PDE solve per step, the bracket on the host.
constexpr int OC_ROUND = 16;
icx_status oc_update_dev(icx_job *j, double *dx, const double *dfdx, double g0, const double *dgdx,
                         double xmin, double xmax, double move, int filter, double beta, double T,
                         double *dxnew, double *dtmp, int *bisections) {
  const long long ne = j->g.ne;
  icx_ctx *c = j->ctx;
  const int project = beta > 0.0;
  if (filter != ICX_FILTER_HELMHOLTZ) {
    HIPCHK(c, j->d_oc.alloc(1));
    OcBisect *bs = j->d_oc.p;
    launch_oc_init(bs, g0, S(j));
    OcBisect *h = reinterpret_cast<OcBisect *>(c->h_scalar);  // pinned, 64 doubles
    for (;;) {
      for (int r = 0; r < OC_ROUND; ++r) {
        if (filter == ICX_FILTER_CONE && j->mesh)
          launch_oc_volume_csr(ne, cone_csr(j), j->cw.aa.p, 0.0, move, xmin, xmax, dx, dfdx, dgdx, project, beta, T,
                               j->d_vol.p, j->d_scalars.p, j->red(), S(j), bs);
        else if (filter == ICX_FILTER_CONE)
          launch_oc_volume_cone(j->g, stencil(j), j->cw.aa.p, 0.0, move, xmin, xmax, dx, dfdx, dgdx, project, beta,
                                T, j->volume, j->d_scalars.p, j->red(), S(j), bs);
        else {
          launch_oc_candidate(ne, 0.0, move, xmin, xmax, dx, dfdx, dgdx, dxnew, S(j), bs, true);
          launch_volume(ne, project, beta, T, dxnew, j->volume, j->mesh ? j->d_vol.p : nullptr, j->d_scalars.p,
                        j->red(), S(j));
        }
        launch_oc_decide(bs, j->d_scalars.p, S(j));
      }
      HIPCHK(c, hipMemcpyAsync(h, bs, sizeof(OcBisect), hipMemcpyDeviceToHost, S(j)));
      HIPCHK(c, hipStreamSynchronize(S(j)));
      if (h->done) break;
    }
    launch_oc_candidate(ne, 0.0, move, xmin, xmax, dx, dfdx, dgdx, dxnew, S(j), bs, false);
    HIPCHK(c, hipMemcpyAsync(dx, dxnew, ne * sizeof(double), hipMemcpyDeviceToDevice, S(j)));
    if (bisections) *bisections = h->counter;
    return ICX_OK;
  }
  double l0 = 0.0, l1 = 1.0e4, lambda = 0.0;
  int counter = 0;
  while ((l1 - l0) / (l1 + l0) > 1.0e-5 && 100 > counter) {
    lambda = 0.5 * (l1 + l0);
    launch_oc_candidate(ne, lambda, move, xmin, xmax, dx, dfdx, dgdx, dxnew, S(j));
    icx_status hrc = helm_apply_dev(j, 0, j->opt.ready ? j->opt.par.helm_tol : 1e-10, dxnew, dtmp, nullptr);
    if (hrc) return hrc;
    launch_volume(ne, project, beta, T, dtmp, j->volume, j->mesh ? j->d_vol.p : nullptr, j->d_scalars.p, j->red(),
                  S(j));
    double gk;
    icx_status rc = read_scalars(j, 1, &gk);
    if (rc) return rc;
    if (gk - g0 > 0.0) l0 = lambda;
    else l1 = lambda;
    counter++;
  }
  launch_oc_candidate(ne, lambda, move, xmin, xmax, dx, dfdx, dgdx, dxnew, S(j));
  HIPCHK(c, hipMemcpyAsync(dx, dxnew, ne * sizeof(double), hipMemcpyDeviceToDevice, S(j)));
  if (bisections) *bisections = counter;
  return ICX_OK;
}

MMAVecs mma_vecs(MMAWork &m) {
  return MMAVecs{m.low.p, m.upp.p, m.alpha.p, m.beta.p, m.p0.p, m.q0.p, m.p1.p, m.q1.p, m.xold1.p, m.xold2.p};
}

icx_status mma_alloc(icx_job *j, MMAWork &m) {
  if (m.ready) return ICX_OK;
  icx_ctx *c = j->ctx;
  const long long n = j->g.ne;
  for (DBuf<double> *b : {&m.low, &m.upp, &m.alpha, &m.beta, &m.p0, &m.q0, &m.p1, &m.q1, &m.xold1, &m.xold2}) {
    HIPCHK(c, b->alloc(n));
    HIPCHK(c, hipMemsetAsync(b->p, 0, n * sizeof(double), S(j)));
  }
  m.iter = 0;
  m.ready = true;
  return ICX_OK;
}

// SolveDIP with m = 1, a = 0 (MMASolver.cpp:102-154, 184-398; GCMMASolver.cpp:135-197,
// 227-429): the scalar dual Newton iteration on the host; x(lambda) into dx and its two
// sums on the device.  b is the constraint's constant, cc the y penalty c.
icx_status mma_dual_host(icx_job *j, const MMAVecs &v, double *dx, double b, double cc, double epsimin,
                         int *steps) {
  const long long n = j->g.ne;
  const double a = 0.0;
  icx_status rc = ICX_OK;
  double lam = cc / 2.0, mu = 1.0;
  double epsi = 1.0, err = 1.0;
  int outloop = 0, total = 0;
  double S1 = 0.0, S2 = 0.0, y = 0.0, z = 0.0;
  auto xyz = [&](void) -> icx_status {
    if (lam < 0.0) lam = 0;
    y = std::max(0.0, lam - cc);
    z = std::max(0.0, 10.0 * (lam * a - 1.0));
    launch_mma_dual(n, lam, v, dx, j->d_scalars.p, j->red(), S(j));
    double o[2];
    icx_status r = read_scalars(j, 2, o);
    S1 = o[0];
    S2 = o[1];
    return r;
  };
  rc = xyz();
  if (rc) return rc;
  while (epsi > epsimin && outloop < 100) {
    outloop++;
    int loop = 0;
    while (err > 0.9 * epsi && loop < 100) {
      loop++;
      total++;
      // DualGrad / DualHess at the current lambda (x(lambda), S1, S2 are current)
      double grad = -b - a * z - y + S1;
      grad = -1.0 * grad - epsi / lam;
      if (lam < 0.0) lam = 0.0;
      double hess = S2;
      if (lam > cc) hess += -1.0;
      hess += -mu / lam;
      double corr = 1e-4 * hess / 1;
      if (-1.0 * corr < 1.0e-7) corr = -1.0e-7;
      hess += corr;
      const double s0 = grad / hess;
      const double s1 = -mu + epsi / lam - s0 * mu / lam;
      double theta = 1.005;
      if (theta < -1.01 * s0 / lam) theta = -1.01 * s0 / lam;
      if (theta < -1.01 * s1 / mu) theta = -1.01 * s1 / mu;
      theta = 1.0 / theta;
      lam = lam + theta * s0;
      mu = mu + theta * s1;
      rc = xyz();
      if (rc) return rc;
      const double r0 = -b - a * z - y + mu + S1;
      const double r1 = mu * lam - epsi;
      err = std::max(std::fabs(r0), std::fabs(r1));
    }
    epsi = epsi * 0.1;
  }
  if (steps) *steps = total;
  return ICX_OK;
}

// The same loop on the device (MmaDual): after the first evaluation, rounds of
// (advance, evaluate) pairs with one read of the state per round; pairs queued past the
// end leave at once, so x holds x(lambda) of the last Newton step, as in the host loop.
constexpr int MMA_DUAL_ROUND = 16;
icx_status mma_dual(icx_job *j, const MMAVecs &v, double *dx, double b, double cc, double epsimin, int *steps) {
  icx_ctx *c = j->ctx;
  if (c->mma_host_dual) return mma_dual_host(j, v, dx, b, cc, epsimin, steps);
  const long long n = j->g.ne;
  HIPCHK(c, j->d_mdual.alloc(1));
  MmaDual *d = j->d_mdual.p;
  launch_mma_dual_init(d, b, cc, epsimin, S(j));
  launch_mma_dual(n, 0.0, v, dx, j->d_scalars.p, j->red(), S(j), d);
  static_assert(sizeof(MmaDual) <= 64 * sizeof(double), "pinned scalar buffer");
  MmaDual *h = reinterpret_cast<MmaDual *>(c->h_scalar);  // pinned, 64 doubles
  for (;;) {
    for (int r = 0; r < MMA_DUAL_ROUND; ++r) {
      launch_mma_dual_advance(d, j->d_scalars.p, S(j));
      launch_mma_dual(n, 0.0, v, dx, j->d_scalars.p, j->red(), S(j), d);
    }
    HIPCHK(c, hipMemcpyAsync(h, d, sizeof(MmaDual), hipMemcpyDeviceToHost, S(j)));
    HIPCHK(c, hipStreamSynchronize(S(j)));
    if (h->done) break;
  }
  if (steps) *steps = h->total;
  return ICX_OK;
}

// MMASolver::Update with m = 1 (MMASolver.cpp:66-154, 400-512): GenSub on the device,
// then the dual solve with c = 1e5.
icx_status mma_update_dev(icx_job *j, double *dx, const double *dfdx, double gx, const double *dgdx,
                          double xmin, double xmax, int *steps) {
  MMAWork &m = j->mma;
  icx_status rc = mma_alloc(j, m);
  if (rc) return rc;
  const long long n = j->g.ne;
  const double move = 0.15, albefa = 0.2, raa0 = 0.00001, xmamieps = 1.0e-5;
  const double asyminit = m.asyminit, asymdec = m.asymdec, asyminc = m.asyminc;
  const double epsimin = std::sqrt((double)(n + 1)) * 1e-9;
  m.iter++;
  MMAVecs v = mma_vecs(m);
  launch_mma_gensub(n, m.iter, 1, 1, v, dx, dfdx, dgdx, xmin, xmax, move, albefa, asyminit, asymdec, asyminc,
                    raa0, raa0, xmamieps, j->d_scalars.p, j->red(), S(j));
  double sums[2];
  rc = read_scalars(j, 2, sums);
  if (rc) return rc;
  return mma_dual(j, v, dx, -gx + sums[1], 100000.0, epsimin, steps);
}

// GCMMASolver with m = 1, c = 1000 (GCMMASolver.cpp:41-129, 432-620).  outer = 1:
// OuterUpdate (new asymptotes and x history, raa0 / raa decay to max(1e-6, 0.1 raa));
// outer = 0: InnerUpdate, whose RaaUpdate grows raa0 / raa from the true values
// f0new / fnew at the last returned point, which xmma holds on entry.  Both then build
// the subproblem at xval, solve its dual into xmma and evaluate the approximations there.
icx_status gcmma_update_dev(icx_job *j, bool outer, double *xmma, double f0new, double fnew, const double *xval,
                            double f0x, const double *df0dx, double fx, const double *dfdx, double xmin,
                            double xmax, int *steps) {
  GcmmaWork &g = j->gc;
  icx_status rc = mma_alloc(j, g.v);
  if (rc) return rc;
  const long long n = j->g.ne;
  const double move = 0.15, albefa = 0.2, asyminit = 0.2, asymdec = 0.7, asyminc = 1.2, xmamieps = 1.0e-5;
  const double raaeps = 1e-6, epsimin = std::sqrt((double)(n + 1)) * 1e-9;
  const double xmami = std::max(xmamieps, xmax - xmin);
  MMAVecs v = mma_vecs(g.v);
  if (outer) {
    g.v.iter++;
    g.raa0 = std::max(raaeps, 0.1 * g.raa0);
    g.raa = std::max(raaeps, 0.1 * g.raa);
  } else {
    launch_gcmma_approx(n, v, xmma, xval, xmami, j->d_scalars.p, j->red(), S(j));
    double a[3];
    rc = read_scalars(j, 3, a);
    if (rc) return rc;
    const double raacof = std::max(1e-12, a[2]);
    if (f0new > g.f0app + 0.5 * epsimin) g.raa0 = std::min(1.1 * (g.raa0 + (1.0 / raacof) * (f0new - g.f0app)), 10.0 * g.raa0);
    if (fnew > g.fapp + 0.5 * epsimin) g.raa = std::min(1.1 * (g.raa + (1.0 / raacof) * (fnew - g.fapp)), 10.0 * g.raa);
  }
  launch_mma_gensub(n, g.v.iter, outer ? 1 : 0, 0, v, xval, df0dx, dfdx, xmin, xmax, move, albefa, asyminit,
                    asymdec, asyminc, g.raa0, g.raa, xmamieps, j->d_scalars.p, j->red(), S(j));
  double sums[2];
  rc = read_scalars(j, 2, sums);
  if (rc) return rc;
  g.r0 = f0x - sums[0];
  g.r = fx - sums[1];
  rc = mma_dual(j, v, xmma, -g.r, 1000.0, epsimin, steps);
  if (rc) return rc;
  launch_gcmma_approx(n, v, xmma, xval, xmami, j->d_scalars.p, j->red(), S(j));
  double a[3];
  rc = read_scalars(j, 3, a);
  if (rc) return rc;
  g.f0app = a[0] + g.r0;
  g.fapp = a[1] + g.r;
  return ICX_OK;
}

// ConCheck (GCMMASolver.cpp:115-129)
bool gcmma_conservative(icx_job *j, double f0new, double fnew) {
  const double epsimin = std::sqrt((double)(j->g.ne + 1)) * 1e-9;
  return !(j->gc.f0app + epsimin < f0new) && !(j->gc.fapp + epsimin < fnew);
}

}  // namespace

extern "C" {

icx_status icx_oc_update(icx_job *j, double *x, const double *dfdx, double g, const double *dgdx, double xmin,
                         double xmax, double move, int filter, double beta, double T, int *bisections) {
  if (!j || !x || !dfdx || !dgdx) return ICX_ERR_ARG;
  if (filter == ICX_FILTER_CONE && !j->cw.ready) return fail(j->ctx, ICX_ERR_STATE, "cone filter not set up");
  if (filter == ICX_FILTER_HELMHOLTZ && !j->hw.ready) return fail(j->ctx, ICX_ERR_STATE, "Helmholtz filter not set up");
  icx_ctx *c = j->ctx;
  const long long ne = j->g.ne;
  DBuf<double> dx, df, dg, xn, tmp;
  HIPCHK(c, dx.alloc(ne)); HIPCHK(c, df.alloc(ne)); HIPCHK(c, dg.alloc(ne)); HIPCHK(c, xn.alloc(ne));
  HIPCHK(c, tmp.alloc(ne));
  HIPCHK(c, hipMemcpyAsync(dx.p, x, ne * sizeof(double), hipMemcpyHostToDevice, S(j)));
  HIPCHK(c, hipMemcpyAsync(df.p, dfdx, ne * sizeof(double), hipMemcpyHostToDevice, S(j)));
  HIPCHK(c, hipMemcpyAsync(dg.p, dgdx, ne * sizeof(double), hipMemcpyHostToDevice, S(j)));
  icx_status rc = oc_update_dev(j, dx.p, df.p, g, dg.p, xmin, xmax, move, filter, beta, T, xn.p, tmp.p, bisections);
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(x, dx.p, ne * sizeof(double), hipMemcpyDeviceToHost, S(j)));
  HIPCHK(c, hipStreamSynchronize(S(j)));
  return ICX_OK;
}

icx_status icx_mma_reset(icx_job *j) {
  if (!j) return ICX_ERR_ARG;
  j->mma.iter = 0;
  return ICX_OK;
}

icx_status icx_mma_set_asymptotes(icx_job *j, double init, double decrease, double increase) {
  if (!j) return ICX_ERR_ARG;
  j->mma.asyminit = init;
  j->mma.asymdec = decrease;
  j->mma.asyminc = increase;
  return ICX_OK;
}

icx_status icx_mma_update(icx_job *j, double *x, const double *dfdx, double gx, const double *dgdx, double xmin,
                          double xmax, int *newton_steps) {
  if (!j || !x || !dfdx || !dgdx) return ICX_ERR_ARG;
  icx_ctx *c = j->ctx;
  const long long ne = j->g.ne;
  DBuf<double> dx, df, dg;
  HIPCHK(c, dx.alloc(ne)); HIPCHK(c, df.alloc(ne)); HIPCHK(c, dg.alloc(ne));
  HIPCHK(c, hipMemcpyAsync(dx.p, x, ne * sizeof(double), hipMemcpyHostToDevice, S(j)));
  HIPCHK(c, hipMemcpyAsync(df.p, dfdx, ne * sizeof(double), hipMemcpyHostToDevice, S(j)));
  HIPCHK(c, hipMemcpyAsync(dg.p, dgdx, ne * sizeof(double), hipMemcpyHostToDevice, S(j)));
  icx_status rc = mma_update_dev(j, dx.p, df.p, gx, dg.p, xmin, xmax, newton_steps);
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(x, dx.p, ne * sizeof(double), hipMemcpyDeviceToHost, S(j)));
  HIPCHK(c, hipStreamSynchronize(S(j)));
  return ICX_OK;
}

icx_status icx_gcmma_reset(icx_job *j) {
  if (!j) return ICX_ERR_ARG;
  j->gc.v.iter = 0;
  return ICX_OK;
}

static icx_status gcmma_call(icx_job *j, bool outer, double *xmma, double f0new, double fnew, const double *xval,
                             double f0x, const double *df0dx, double fx, const double *dfdx, double xmin,
                             double xmax, int *newton_steps) {
  if (!j || !xmma || !xval || !df0dx || !dfdx) return ICX_ERR_ARG;
  icx_ctx *c = j->ctx;
  if (!outer && j->gc.v.iter == 0) return fail(c, ICX_ERR_STATE, "GCMMA inner update before an outer update");
  const long long ne = j->g.ne;
  DBuf<double> xm, xv, df, dg;
  HIPCHK(c, xm.alloc(ne)); HIPCHK(c, xv.alloc(ne)); HIPCHK(c, df.alloc(ne)); HIPCHK(c, dg.alloc(ne));
  if (!outer) HIPCHK(c, hipMemcpyAsync(xm.p, xmma, ne * sizeof(double), hipMemcpyHostToDevice, S(j)));
  HIPCHK(c, hipMemcpyAsync(xv.p, xval, ne * sizeof(double), hipMemcpyHostToDevice, S(j)));
  HIPCHK(c, hipMemcpyAsync(df.p, df0dx, ne * sizeof(double), hipMemcpyHostToDevice, S(j)));
  HIPCHK(c, hipMemcpyAsync(dg.p, dfdx, ne * sizeof(double), hipMemcpyHostToDevice, S(j)));
  icx_status rc = gcmma_update_dev(j, outer, xm.p, f0new, fnew, xv.p, f0x, df.p, fx, dg.p, xmin, xmax, newton_steps);
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(xmma, xm.p, ne * sizeof(double), hipMemcpyDeviceToHost, S(j)));
  HIPCHK(c, hipStreamSynchronize(S(j)));
  return ICX_OK;
}

icx_status icx_gcmma_outer_update(icx_job *j, double *xmma, const double *xval, double f0x, const double *df0dx,
                                  double fx, const double *dfdx, double xmin, double xmax, int *newton_steps) {
  return gcmma_call(j, true, xmma, 0.0, 0.0, xval, f0x, df0dx, fx, dfdx, xmin, xmax, newton_steps);
}

icx_status icx_gcmma_inner_update(icx_job *j, double *xmma, double f0xnew, double fxnew, const double *xval,
                                  double f0x, const double *df0dx, double fx, const double *dfdx, double xmin,
                                  double xmax, int *newton_steps) {
  return gcmma_call(j, false, xmma, f0xnew, fxnew, xval, f0x, df0dx, fx, dfdx, xmin, xmax, newton_steps);
}

icx_status icx_gcmma_concheck(const icx_job *jc, double f0xnew, double fxnew, int *conservative) {
  icx_job *j = const_cast<icx_job *>(jc);
  if (!j || !conservative) return ICX_ERR_ARG;
  if (j->gc.v.iter == 0) return fail(j->ctx, ICX_ERR_STATE, "GCMMA check before an outer update");
  *conservative = gcmma_conservative(j, f0xnew, fxnew) ? 1 : 0;
  return ICX_OK;
}

// ---- device-resident design loop ------------------------------------------------------------
icx_status icx_opt_init(icx_job *j, const icx_opt_params *par, const double *x0) {
  if (!j || !par || !x0) return ICX_ERR_ARG;
  ICX_REFUSE_IF_SUBMITTED(j);
  if (j->numeq < 0) return fail(j->ctx, ICX_ERR_STATE, "icx_make_dof first");
  if (j->heat()) return fail(j->ctx, ICX_ERR_ARG, "the device design loop drives elastic jobs");
  if (par->driver < 0 || par->driver > 2) return fail(j->ctx, ICX_ERR_ARG, "driver must be 0 (OC), 1 (MMA) or 2 (GCMMA)");
  icx_ctx *c = j->ctx;
  OptWork &o = j->opt;
  o.par = *par;
  o.step = 0;
  o.V0 = 0.0;
  const long long ne = j->g.ne;
  for (DBuf<double> *b : {&o.x, &o.fil, &o.thr, &o.dthr, &o.dfds, &o.dgds, &o.tmp, &o.tmp2}) HIPCHK(c, b->alloc(ne));
  HIPCHK(c, hipMemcpyAsync(o.x.p, x0, ne * sizeof(double), hipMemcpyHostToDevice, S(j)));
  icx_status rc = ICX_OK;
  if (par->filter == ICX_FILTER_CONE) rc = cone_setup(j, par->filter_radius);
  else if (par->filter == ICX_FILTER_HELMHOLTZ) rc = helm_setup(j, par->filter_radius);
  if (rc) return rc;
  if (par->driver == 1) {
    rc = mma_alloc(j, j->mma);
    if (rc) return rc;
    j->mma.iter = 0;
  } else if (par->driver == 2) {
    rc = mma_alloc(j, j->gc.v);
    if (rc) return rc;
    GcmmaWork &g = j->gc;  // a new GCMMASolver (GCMMASolver.cpp:41-56)
    g.v.iter = 0;
    g.raa0 = g.raa = 1.0;
    g.r0 = g.r = g.f0app = g.fapp = 0.0;
    HIPCHK(c, o.xnew.alloc(ne));
  }
  HIPCHK(c, hipStreamSynchronize(S(j)));
  o.ready = true;
  return ICX_OK;
}

}  // extern "C"

namespace {

// filter, projection (beta of the current step), design -> moduli: the design of x
// (opt_macro_oc.cpp:87-101, opt_mma_Helmholtz.cpp:86-98, opt_macro_gcmma.cpp:96-101)
icx_status opt_design(icx_job *j, const double *x) {
  OptWork &o = j->opt;
  const icx_opt_params &P = o.par;
  icx_ctx *c = j->ctx;
  const long long ne = j->g.ne;
  if (P.filter == ICX_FILTER_CONE) cone_apply(j, 0, x, o.fil.p);
  else if (P.filter == ICX_FILTER_HELMHOLTZ) {
    icx_status rc = helm_apply_dev(j, 0, P.helm_tol, x, o.fil.p, nullptr);
    if (rc) return rc;
  } else HIPCHK(c, hipMemcpyAsync(o.fil.p, x, ne * sizeof(double), hipMemcpyDeviceToDevice, S(j)));
  if (P.projection) launch_project(ne, o.beta, 0.5, o.fil.p, o.thr.p, o.dthr.p, S(j));
  else HIPCHK(c, hipMemcpyAsync(o.thr.p, o.fil.p, ne * sizeof(double), hipMemcpyDeviceToDevice, S(j)));
  HIPCHK(c, hipMemcpyAsync(j->d_design.p, o.thr.p, ne * sizeof(double), hipMemcpyDeviceToDevice, S(j)));
  launch_simp(ne, j->prob.young1, j->prob.young2, j->prob.penalty, j->d_design.p, j->d_E.p, nullptr, S(j));
  j->design_set = true;
  return ICX_OK;
}

// the structural solve of the design loop in its configured form
icx_status opt_solve(icx_job *j, icx_solve_info &info) {
  const icx_opt_params &P = j->opt.par;
  icx_status rc = build_rhs(j);
  if (rc) return rc;
  if (P.form == ICX_SOLVE_PLAIN) rc = solve_plain(j, P.tol, P.maxit, info);
  else if (P.form == ICX_SOLVE_FUSED) rc = solve_fused(j, P.tol, P.maxit, info);
  else rc = solve_mixed(j, P.tol, P.maxit, info);
  return rc == ICX_ERR_NOCONV ? ICX_OK : rc;
}

// objective u.fext and volume sum rho_hat v of the job's current solve and design
icx_status opt_objective_volume(icx_job *j, double *obj, double *vol) {
  launch_dot(j->nvec(), j->d_u.p, j->d_fext.p, j->d_scalars.p, j->red(), S(j));
  launch_volume(j->g.ne, 0, 0.0, 0.5, j->opt.thr.p, j->volume, j->mesh ? j->d_vol.p : nullptr, j->d_scalars.p + 1,
                j->red(), S(j));
  double sc[2];
  icx_status rc = read_scalars(j, 2, sc);
  *obj = sc[0];
  *vol = sc[1];
  return rc;
}

// The GCMMA step of opt_macro_gcmma.cpp:150-215: outer update from x with (f, df, g, dg),
// a solve of the trial design, and up to 15 inner updates while the approximations are
// not conservative; x takes the last trial design.
icx_status gcmma_step(icx_job *j, double f, const double *df, double g, const double *dg, icx_iter_record &r) {
  OptWork &o = j->opt;
  const icx_opt_params &P = o.par;
  int steps = 0, total = 0;
  double fnew = 0.0, gnew = 0.0;
  auto trial = [&]() -> icx_status {
    icx_status rc = opt_design(j, o.xnew.p);
    if (rc) return rc;
    icx_solve_info info{};
    rc = opt_solve(j, info);
    if (rc) return rc;
    r.trial_cg_iterations += info.iterations;
    double vol;
    rc = opt_objective_volume(j, &fnew, &vol);
    gnew = vol - o.V0;
    return rc;
  };
  icx_status rc = gcmma_update_dev(j, true, o.xnew.p, 0.0, 0.0, o.x.p, f, df, g, dg, P.xmin, P.xmax, &steps);
  if (rc) return rc;
  total += steps;
  if ((rc = trial())) return rc;
  bool conserv = gcmma_conservative(j, fnew, gnew);
  int inner = 0;
  for (; !conserv && inner < 15; ++inner) {
    rc = gcmma_update_dev(j, false, o.xnew.p, fnew, gnew, o.x.p, f, df, g, dg, P.xmin, P.xmax, &steps);
    if (rc) return rc;
    total += steps;
    if ((rc = trial())) return rc;
    conserv = gcmma_conservative(j, fnew, gnew);
  }
  r.inner_loops = inner;
  r.update_steps = total;
  HIPCHK(j->ctx, hipMemcpyAsync(o.x.p, o.xnew.p, j->g.ne * sizeof(double), hipMemcpyDeviceToDevice, S(j)));
  return ICX_OK;
}

}  // namespace

extern "C" {

icx_status icx_opt_step_begin(icx_job *j) {
  if (!j) return ICX_ERR_ARG;
  ICX_REFUSE_IF_SUBMITTED(j);
  OptWork &o = j->opt;
  if (!o.ready) return fail(j->ctx, ICX_ERR_STATE, "icx_opt_init first");
  if (o.in_step) return fail(j->ctx, ICX_ERR_STATE, "design step already begun");
  o.t0 = std::chrono::steady_clock::now();
  const int k = ++o.step;
  // projection, beta = min(2^(k/40), 16) (opt_macro_oc.cpp:94, opt_mma_Helmholtz.cpp:90);
  // the GCMMA driver: min(int(max((k - 100) / 40, 1)), 8) (opt_macro_gcmma.cpp:91-92)
  if (o.par.driver == 2) o.beta = std::min((double)(int)std::max((k - 100) / 40.0, 1.0), 8.0);
  else o.beta = std::min(std::pow(2, (int)(k / 40)), 16.0);
  icx_status rc = opt_design(j, o.x.p);
  if (rc) return rc;
  o.in_step = true;
  return ICX_OK;
}

icx_status icx_opt_step_end(icx_job *j, const icx_solve_info *info, icx_iter_record *rec) {
  if (!j) return ICX_ERR_ARG;
  ICX_REFUSE_IF_SUBMITTED(j);
  OptWork &o = j->opt;
  if (!o.in_step) return fail(j->ctx, ICX_ERR_STATE, "icx_opt_step_begin first");
  o.in_step = false;
  icx_ctx *c = j->ctx;
  const icx_opt_params &P = o.par;
  const long long ne = j->g.ne;
  const int k = o.step;
  icx_iter_record r{};
  r.step = k;
  r.cg_iterations = info ? info->iterations : 0;
  j->u_valid = true;
  // objective and volume; const_h: V (A), V / V0 - 1 (B), V - V0 (GCMMA,
  // opt_macro_gcmma.cpp:109-125)
  icx_status rc = opt_objective_volume(j, &r.objective, &r.volume);
  if (rc) return rc;
  if (k == 1) o.V0 = r.volume;
  if (P.driver == 0) r.constraint = r.volume;
  else if (P.driver == 1) r.constraint = k == 1 ? 0.0 : r.volume / o.V0 - 1.0;
  else r.constraint = r.volume - o.V0;
  // sensitivity and chain rule
  const double *usens;
  {
    icx_status rs = sens_displacement(j, &usens);
    if (rs) return rs;
  }
  if (j->mesh)
    launch_mesh_sensitivity(j->mop, j->prob.young1, j->prob.young2, j->prob.penalty, j->d_design.p, usens,
                            j->d_vol.p, P.driver == 1 ? o.V0 : 1.0, o.dfds.p, o.dgds.p, S(j));
  else
    launch_sensitivity(j->g, j->op64, j->prob.young1, j->prob.young2, j->prob.penalty, j->d_design.p, usens,
                       j->volume, P.driver == 1 ? o.V0 : 1.0, o.dfds.p, o.dgds.p, S(j));
  // dfds *= dthr, dgds *= dthr (projection derivative; 1 without projection)
  if (P.projection) {
    launch_mul(ne, o.dthr.p, o.dfds.p, S(j));
    launch_mul(ne, o.dthr.p, o.dgds.p, S(j));
  }
  // filter adjoint
  if (P.filter == ICX_FILTER_CONE) {
    cone_apply(j, 1, o.dfds.p, o.tmp.p);
    cone_apply(j, 1, o.dgds.p, o.tmp2.p);
  } else if (P.filter == ICX_FILTER_HELMHOLTZ) {
    rc = helm_apply_dev(j, 1, P.helm_tol, o.dfds.p, o.tmp.p, nullptr);
    if (rc) return rc;
    rc = helm_apply_dev(j, 1, P.helm_tol, o.dgds.p, o.tmp2.p, nullptr);
    if (rc) return rc;
  } else {
    HIPCHK(c, hipMemcpyAsync(o.tmp.p, o.dfds.p, ne * sizeof(double), hipMemcpyDeviceToDevice, S(j)));
    HIPCHK(c, hipMemcpyAsync(o.tmp2.p, o.dgds.p, ne * sizeof(double), hipMemcpyDeviceToDevice, S(j)));
  }
  // update
  int steps = 0;
  if (P.driver == 0) {
    rc = oc_update_dev(j, o.x.p, o.tmp.p, o.V0, o.tmp2.p, P.xmin, P.xmax, P.move, P.filter,
                       P.projection ? o.beta : 0.0, 0.5, o.fil.p, o.dfds.p, &steps);
  } else if (P.driver == 1) {
    rc = mma_update_dev(j, o.x.p, o.tmp.p, r.constraint, o.tmp2.p, P.xmin, P.xmax, &steps);
  } else {
    rc = gcmma_step(j, r.objective, o.tmp.p, r.constraint, o.tmp2.p, r);
    steps = r.update_steps;
  }
  if (rc) return rc;
  r.update_steps = steps;
  HIPCHK(c, hipStreamSynchronize(S(j)));
  r.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - o.t0).count();
  if (rec) *rec = r;
  return ICX_OK;
}

icx_status icx_opt_step(icx_job *j, icx_iter_record *rec) {
  icx_status rc = icx_opt_step_begin(j);
  if (rc) return rc;
  icx_solve_info info{};
  rc = opt_solve(j, info);
  if (rc) {
    j->opt.in_step = false;
    return rc;
  }
  return icx_opt_step_end(j, &info, rec);
}

icx_status icx_opt_get_design(const icx_job *jc, double *x) {
  icx_job *j = const_cast<icx_job *>(jc);
  if (!j || !x) return ICX_ERR_ARG;
  if (!j->opt.ready) return ICX_ERR_STATE;
  icx_ctx *c = j->ctx;
  HIPCHK(c, hipMemcpyAsync(x, j->opt.x.p, j->g.ne * sizeof(double), hipMemcpyDeviceToHost, S(j)));
  HIPCHK(c, hipStreamSynchronize(S(j)));
  return ICX_OK;
}

icx_status icx_submit(icx_job *j, double tol, int maxit) { return icx_submit_form(j, ICX_SOLVE_MIXED, tol, maxit); }

icx_status icx_submit_form(icx_job *j, int form, double tol, int maxit) {
  if (j && j->mesh) return fail(j->ctx, ICX_ERR_ARG, "mesh jobs: %s is for box jobs only", "submitted solves");
  if (!j) return ICX_ERR_ARG;
  icx_ctx *c = j->ctx;
  if (form != ICX_SOLVE_MIXED && form != ICX_SOLVE_FUSED)
    return fail(c, ICX_ERR_ARG, "submitted solves: form ICX_SOLVE_MIXED or ICX_SOLVE_FUSED");
  if (j->numeq < 0 || !j->design_set)
    return fail(c, ICX_ERR_STATE, "icx_submit needs icx_make_dof and icx_set_design");
  if (!(tol > 0.0) || maxit < 1) return fail(c, ICX_ERR_ARG, "tol must be > 0 and maxit >= 1");
  if (j->heat()) return fail(c, ICX_ERR_ARG, "batched solves: elastic jobs only");
  if (j->sub.active) return fail(c, ICX_ERR_STATE, "job already has a submitted solve");
  (void)hipSetDevice(c->device);
  if (!c->batch) c->batch = new BatchRun();
  auto &b = j->sub;
  b.t0 = std::chrono::steady_clock::now();
  b.tol = tol;
  b.maxit = maxit;
  b.refinements = 0;
  b.replacing = false;
  b.has_info = false;
  b.form = form;
  icx_status rc = build_rhs(j);
  if (rc) return rc;
  SolverState st;
  rc = form == ICX_SOLVE_FUSED ? fused_begin(j, tol, maxit, st) : mixed_begin(j, tol, maxit, st, use_sr(j));
  if (rc) return rc;
  b.active = true;
  c->batch->active.push_back(j);
  c->batch->dirty = true;
  return ICX_OK;
}

icx_status icx_cancel(icx_job *j) {
  if (!j) return ICX_ERR_ARG;
  icx_ctx *c = j->ctx;
  if (!j->sub.active) return fail(c, ICX_ERR_STATE, "icx_cancel: the job has no submitted solve");
  HIPCHK(c, hipStreamSynchronize(c->stream));  // queued chunks that touch the job end first
  batch_remove(c, j);
  j->sub.has_info = false;
  j->u_valid = false;
  return ICX_OK;
}

icx_status icx_poll(icx_ctx *c, int chunks, icx_job **finished, int cap, int *nfinished) {
  if (!c || !nfinished || (cap > 0 && !finished)) return ICX_ERR_ARG;
  *nfinished = 0;
  if (!c->batch || c->batch->active.empty()) return ICX_OK;
  (void)hipSetDevice(c->device);
  BatchRun &B = *c->batch;
  std::vector<icx_job *> done;
  icx_status rc = ICX_OK;
  for (int ch = 0; ch < std::max(chunks, 1) && done.empty() && !B.active.empty(); ++ch) {
    if (B.dirty && (rc = batch_rebuild(c, B))) return rc;
    // states first: a solve may already be stopped (zero load, converged start)
    if ((rc = batch_states(c, B))) return rc;
    bool runnable = false;
    for (size_t q = 0; q < B.active.size(); ++q) runnable |= B.h_states[q].done == 0;
    if (runnable) {
      auto chunk = [&]() { batch_chunk(c, B); };
      if (!c->use_graphs || c->prof_kernel) {
        chunk();
      } else {
        if (!B.graph.exec) {
          hipGraph_t graph = nullptr;
          HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
          chunk();
          HIPCHK(c, hipStreamEndCapture(c->stream, &graph));
          const hipError_t e = hipGraphInstantiate(&B.graph.exec, graph, nullptr, nullptr, 0);
          (void)hipGraphDestroy(graph);
          if (e != hipSuccess) {
            B.graph.exec = nullptr;
            c->use_graphs = false;
          }
        }
        if (B.graph.exec) HIPCHK(c, hipGraphLaunch(B.graph.exec, c->stream));
        else chunk();
      }
      if ((rc = batch_states(c, B))) return rc;
    }
    // stopped solves: converged at set-up, a fused-form stop (the owed x step, then
    // done), or a residual replacement is due
    bool replaced = false, flushed = false;
    for (size_t q = 0; q < B.active.size(); ++q) {
      icx_job *j = B.active[q];
      const SolverState &st = B.h_states[q];
      if (j->sub.form == ICX_SOLVE_FUSED) {
        if (st.done) {
          fused_flush(j, st);
          flushed = true;
          batch_finish(j, st, (st.done == 1 || st.done == 3) ? ICX_OK : ICX_ERR_NOCONV);
          done.push_back(j);
        }
        continue;
      }
      if (st.done == 1) {
        batch_finish(j, st, ICX_OK);
        done.push_back(j);
      } else if (st.done >= 2) {
        j->sub.was = st.done;
        j->sub.replacing = true;
        mixed_replace(j);
        ++j->sub.refinements;
        replaced = true;
      }
    }
    if (replaced) {
      if ((rc = batch_states(c, B))) return rc;
      for (size_t q = 0; q < B.active.size(); ++q) {
        icx_job *j = B.active[q];
        if (!j->sub.replacing) continue;
        j->sub.replacing = false;
        const SolverState &st = B.h_states[q];
        if (st.done == 1) {
          batch_finish(j, st, ICX_OK);
          done.push_back(j);
        } else if (j->sub.was == 2 || st.done == 2) {
          batch_finish(j, st, ICX_ERR_NOCONV);
          done.push_back(j);
        }
      }
    }
    if (flushed) HIPCHK(c, hipStreamSynchronize(c->stream));  // x complete before the caller reads it
    for (icx_job *j : done) {
      auto it = std::find(B.active.begin(), B.active.end(), j);
      if (it != B.active.end()) B.active.erase(it);
      B.dirty = true;
    }
  }
  const int n = (int)done.size();
  for (int q = 0; q < n && q < cap; ++q) finished[q] = done[q];
  *nfinished = n;
  if (n > cap) return fail(c, ICX_ERR_ARG, "%d solves finished, room for %d", n, cap);
  return ICX_OK;
}

icx_status icx_job_last_info(const icx_job *j, icx_solve_info *info) {
  if (!j || !info) return ICX_ERR_ARG;
  if (!j->sub.has_info) return fail(j->ctx, ICX_ERR_STATE, "no finished submitted solve");
  *info = j->sub.info;
  return j->sub.rc;
}

icx_status icx_solve_batch(icx_ctx *ctx, icx_job **jobs, int njobs, int form, double tol, int maxit,
                           icx_solve_info *infos) {
  if (!ctx || (!jobs && njobs)) return ICX_ERR_ARG;
  icx_status worst = ICX_OK;
  // mixed and (box elastic) fused-form solves run batched; the plain form, and fused
  // solves of mesh or heat jobs (which refuse it alone too), one after the other
  bool batched = form == ICX_SOLVE_MIXED || form == ICX_SOLVE_FUSED;
  for (int i = 0; i < njobs && form == ICX_SOLVE_FUSED; ++i)
    if (jobs[i] && (jobs[i]->mesh || jobs[i]->heat())) batched = false;
  if (!batched) {
    for (int i = 0; i < njobs; ++i) {
      icx_status rc = icx_solve(jobs[i], form, tol, maxit, nullptr, infos ? &infos[i] : nullptr);
      if (rc && rc != ICX_ERR_NOCONV) return rc;
      if (rc) worst = rc;
    }
    return worst;
  }
  for (int i = 0; i < njobs; ++i) {
    if (!jobs[i] || jobs[i]->ctx != ctx) return fail(ctx, ICX_ERR_ARG, "job %d is not a job of this context", i);
    icx_status rc = icx_submit_form(jobs[i], form, tol, maxit);
    if (rc) return rc;
  }
  int left = njobs;
  std::vector<icx_job *> fin(njobs > 0 ? njobs : 1);
  while (left > 0) {
    int nf = 0;
    icx_status rc = icx_poll(ctx, 1 << 20, fin.data(), njobs, &nf);
    if (rc) return rc;
    if (nf == 0 && (!ctx->batch || ctx->batch->active.empty())) break;
    left -= nf;
  }
  for (int i = 0; i < njobs; ++i) {
    icx_solve_info inf{};
    icx_status rc = icx_job_last_info(jobs[i], &inf);
    if (infos) infos[i] = inf;
    if (rc && rc != ICX_ERR_NOCONV) return rc;
    if (rc) worst = rc;
  }
  return worst;
}

icx_status icx_job_footprint(const icx_job *j, size_t *bytes) {
  if (!j || !bytes) return ICX_ERR_ARG;
  size_t b = 0;
  auto add = [&](const auto &d) { b += d.bytes(); };
  add(j->d_dof); add(j->d_mask); add(j->d_fext); add(j->d_dirich); add(j->d_b); add(j->d_u);
  add(j->d_design); add(j->d_E); add(j->d_vec); add(j->d_vec2); add(j->d_partial); add(j->d_pcgpart);
  add(j->d_state); add(j->d_scalars);
  for (auto *d : {&j->pw.r, &j->pw.p, &j->pw.t, &j->pw.z, &j->pw.invd}) add(*d);
  for (auto *d : {&j->fw.r, &j->fw.pa, &j->fw.pb, &j->fw.Ap, &j->fw.invd, &j->fw.invs, &j->fw.xa, &j->fw.ba})
    add(*d);
  for (auto *d : {&j->mw.r, &j->mw.pa, &j->mw.pb, &j->mw.Ap, &j->mw.Ap2, &j->mw.xlo, &j->mw.invd, &j->mw.invs, &j->mw.E,
                  &j->mw.b32, &j->mw.r1, &j->mw.ap1})
    add(*d);
  add(j->mw.Kx); add(j->mw.part);
  for (auto *d : {&j->hw.invd, &j->hw.b, &j->hw.x, &j->hw.r, &j->hw.p, &j->hw.t, &j->hw.z}) add(*d);
  add(j->cw.off); add(j->cw.w); add(j->cw.aa); add(j->cw.ptr); add(j->cw.id); add(j->cw.wt);
  for (const MMAWork *m : {&j->mma, &j->gc.v})
    for (auto *d : {&m->low, &m->upp, &m->alpha, &m->beta, &m->p0, &m->q0, &m->p1, &m->q1, &m->xold1, &m->xold2})
      add(*d);
  for (auto *d : {&j->opt.x, &j->opt.fil, &j->opt.thr, &j->opt.dthr, &j->opt.dfds, &j->opt.dgds, &j->opt.tmp,
                  &j->opt.tmp2, &j->opt.xnew})
    add(*d);
  // mesh jobs and the output step
  add(j->d_conn); add(j->d_adj_ptr); add(j->d_slot); add(j->d_xyz); add(j->d_ye); add(j->d_vol);
  add(j->ow.ufull); add(j->ow.nodes); add(j->ow.mises); add(j->d_oc); add(j->d_mdual);
  *bytes = b;
  return ICX_OK;
}

}  // extern "C"
