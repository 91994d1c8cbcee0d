// Host-side check of the index arithmetic kernel B uses on the seam layers of the FP32
// hexa8 kernel A: udiv (multiply-shift division for dividends < 2^29) against integer
// division, and seam_flags against a brute-force layer test, over random box meshes
// and tile depths, for component-plane and node-interleaved vectors.  Exit status 0: no
// mismatch.
#include <hip/hip_runtime.h>
#include "../icarat-contact_amd/csrc/icx_internal.h"
#include <algorithm>
#include <cstdio>
#include <random>
using namespace icx;

int main() {
  std::mt19937_64 rng(20261203);
  long long bad = 0, checked = 0;
  // udiv: every dividend below 4096 for small divisors, random and boundary dividends
  // for divisors up to 2^27 (the node limit)
  for (long long d = 1; d < 600; ++d) {
    const Magic g = magic_u29(d);
    for (uint32_t n = 0; n < 4096; ++n, ++checked) bad += udiv(n, g) != n / d;
  }
  for (int t = 0; t < 100000; ++t) {
    const long long d = 1 + (long long)(rng() % (1ULL << 27));
    const Magic g = magic_u29(d);
    const uint32_t top = (1u << 29) - 1;
    const uint32_t ns[] = {(uint32_t)(rng() % (1u << 29)), top, (uint32_t)(top / d * d),
                           (uint32_t)(top / d * d - 1), (uint32_t)d, (uint32_t)(d - 1)};
    for (uint32_t n : ns) { bad += udiv(n, g) != n / d; ++checked; }
  }
  // seam flags: DOF i lies on a seam iff its node layer k = TK s with 0 < s < tiles
  long long bad2 = 0, dofs = 0, seams = 0;
  for (int t = 0; t < 120; ++t) {
    const int sx = 2 + (int)(rng() % 30), sy = 2 + (int)(rng() % 30), nz = 1 + (int)(rng() % 40);
    const int TK = 1 + (int)(rng() % 9), tiles = (nz + TK - 1) / TK;
    if (tiles < 2) continue;
    const long long layer = (long long)sx * sy, nn = layer * (nz + 1), N = 3 * nn;
    SeamGeom sg{};
    sg.nn = magic_u29(nn); sg.L = magic_u29(layer); sg.TK = magic_u29(TK);
    sg.nnodes = (uint32_t)nn; sg.layer = (uint32_t)layer; sg.tk = (uint32_t)TK; sg.kmax = (uint32_t)(tiles - 1) * TK;
    for (long long i0 = 0; i0 < N; i0 += 1 + (long long)(rng() % 3)) {
      const int nw = (int)std::min<long long>(4, N - i0);
      const uint32_t f = seam_flags((uint32_t)i0, nw, sg);
      for (int w = 0; w < nw; ++w) {
        const long long k = ((i0 + w) % nn) / layer;
        const bool ref = k > 0 && k % TK == 0 && k / TK < tiles;
        bad2 += ref != (bool)((f >> w) & 1);
        seams += ref;
        ++dofs;
      }
      // the same vector node-interleaved (entry 3 n + c of node n)
      const uint32_t fa = seam_flags_aos((uint32_t)i0, nw, sg);
      for (int w = 0; w < nw; ++w) {
        const long long k = ((i0 + w) / 3) / layer;
        const bool ref = k > 0 && k % TK == 0 && k / TK < tiles;
        bad2 += ref != (bool)((fa >> w) & 1);
        ++dofs;
      }
    }
  }
  printf("udiv: %lld mismatches in %lld; seam flags: %lld mismatches in %lld DOFs (%lld on seams)\n", bad, checked,
         bad2, dofs, seams);
  return (bad || bad2 || seams == 0) ? 1 : 0;
}
