/*
 * icx_profile.h -- measurement hooks of libicx (no reference counterpart).
 *
 * While a profile window is open, every launch of the selected kernel on the
 * context's stream is bracketed by a pair of HIP events; icx_profile_end returns
 * the summed kernel time and the launch count.  Used by bench.py to measure the
 * dominant kernel's average launch duration inside the timed region.
 */
#ifndef ICX_PROFILE_H
#define ICX_PROFILE_H
#include "icx.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
  ICX_PROF_MATVEC64 = 1,  /* FP64 operator kernel of the plain form */
  ICX_PROF_FUSED_A = 2,   /* fused direction-update + operator kernel (fused/mixed forms) */
  ICX_PROF_FUSED_B = 3,   /* fused residual-update kernel (fused/mixed forms) */
  ICX_PROF_MATVEC32 = 4   /* FP32 operator kernel of icx_matvec (Haar VALU or dense MFMA form) */
};

icx_status icx_profile_begin(icx_ctx *ctx, int kernel);
icx_status icx_profile_end(icx_ctx *ctx, double *seconds, long long *launches);

/* Tuning options of a context (no counterpart in the reference).
 * ICX_OPT_GRAPHS (default 1): replay each job's PCG iteration chunk from a captured
 *   HIP graph instead of launching its kernels one by one; results are identical.
 * ICX_OPT_TILE_DEPTH (default 0 = automatic): node layers along z per block of the
 *   hexa8 operator kernels, 1..64 (raised as needed to keep the block count within
 *   the partial-sum buffers).  Applies to every job of the context; results of the
 *   operator are identical, PCG dot products change only in rounding.  For Haar-diagonal
 *   hexa8 heat jobs: the nodes each thread marches along z (2, 4, 8 or 16; else 4).
 * ICX_OPT_PIPELINE (default 1): while the host waits for the solver state after a chunk
 *   of PCG iterations, the next chunk is already queued (it does nothing if the solve
 *   has stopped); 0 polls chunk by chunk.  Results are identical.
 * ICX_OPT_SINGLE_REDUCTION (default 0): 1 runs hexa8 mixed-precision solves (icx_solve and
 *   submitted solves) with one kernel and one reduction phase per PCG iteration instead
 *   of two (the next r.z is formed from an expansion in dot products reduced with the
 *   previous iteration's).  Same iterates up to rounding, same stopping tests; on one
 *   GPU the two-kernel iteration is faster.  Set it before submitting solves.
 * ICX_OPT_HEAT_DENSE (default 0): 1 makes heat jobs apply their element matrices as dense
 *   rows even where both are diagonal in the element's Haar basis (diagonal conductivity
 *   tensors), the path taken otherwise; for checking one operator against the other.
 * ICX_OPT_MMA_HOST_DUAL (default 0): 1 runs the Newton loop of the MMA / GCMMA dual on
 *   the host, reading the two sums of every step, instead of on the device with one read
 *   of its state per 16 steps; the same arithmetic, for checking one against the other.
 * ICX_OPT_MFMA_MATVEC (default 0): 1 makes icx_matvec / icx_matvec_noalloc at FP32 on
 *   hexa8 box jobs contract the dense element stiffness on the matrix cores
 *   (v_mfma_f32_32x32x2f32) instead of applying its Haar blocks on the vector units;
 *   FP32 results differ in rounding only.  A measurement of the two forms; the solve
 *   forms always use the Haar kernels.
 */
enum { ICX_OPT_GRAPHS = 1, ICX_OPT_TILE_DEPTH = 2, ICX_OPT_PIPELINE = 3, ICX_OPT_SINGLE_REDUCTION = 4,
       ICX_OPT_HEAT_DENSE = 5, ICX_OPT_MMA_HOST_DUAL = 6, ICX_OPT_MFMA_MATVEC = 7 };
icx_status icx_set_option(icx_ctx *ctx, int option, long long value);

#ifdef __cplusplus
}
#endif
#endif
