// efl_reduce.h - deterministic reductions over the blocks of one launch.
//
// Every block of a reduction launch stores its partial at
// red[component * red_stride + local block] (thread 0).  A later kernel folds
// the partials of an instance in a fixed order - thread t sums partials t,
// t + 256, ..., then a tree over 256 threads - so the result does not depend
// on scheduling.  Folding happens in the kernel that needs the scalar (every
// block folds for itself: the partials are a few KB and sit in L2), which
// needs neither a separate finalize launch nor any communication between
// the blocks of one launch.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "efl_internal.h"

namespace efl {

constexpr int RED_NT = 256;  // threads of every block that takes part

// partial components (red[])
constexpr int RED_PQ = 0, RED_RR = 1, RED_BB = 2, RED_RZ = 2;  // bb is folded before any r . z exists
constexpr int RED_ZQ = 3;  // z . q_prev of the flexible (learned) preconditioner
constexpr int RED_ZAZ = 3;  // z . A z of the multigrid preconditioner's last launch (never flexible)

// scal[] slots
constexpr int SC_PQ = 1, SC_RR = 2, SC_BB = 3, SC_ALPHA = 4, SC_BETA = 5, SC_DONE = 6;
constexpr int SC_NRZ = 7;            // number of r . z partials of the last preconditioner
constexpr int SC_NRR = 8;            // number of ||r||^2 partials of the last k_pcg_axpy
constexpr int SC_ITER = 9, SC_REL = 10, SC_CONV = 11;
constexpr int SC_RZ0 = 12;           // r . z of iteration k at SC_RZ0 + (k & 1)
constexpr int SC_KCUR = 14;          // the iteration k_pcg_axpy just ran
constexpr int SC_STALL = 15;         // flexible CG: relative residual at the last stall checkpoint
constexpr int SC_RZF = 16;           // k_prefold: r . z (and z . q_prev) of the last preconditioner
constexpr int SC_PQF = 18;           // k_prefold: p . q of the last k_pcg_apply
constexpr int SC_MISSED = 20;        // asynchronous solves that ended without meeting the tolerance (k_cg_tally)
constexpr int SC_AL0 = 21;           // alpha of iteration k at SC_AL0 + (k & 1)
constexpr int kScalars = 24;         // scal[] length (the host reads the first 16)

// Partial counts up to this are folded by every block that needs the
// scalar; an instance with more (DevInst::prefold) has them folded once by
// k_prefold, one block per instance, and its consumers read scal[].
constexpr int kInlineFold = 2048;

// Fold the nblk partials of components c0 .. c0 + N - 1 (all threads of a
// block of RED_NT call it; every thread gets the sums): thread t sums
// partials t, t + 256, ... in order, a butterfly over the wavefront's lanes
// (the same tree in every lane), then the 4 wavefront sums in order.  The
// partials are loaded FOLD_U at a time (absent ones as 0, which leaves a sum
// or a max of non-negative values unchanged: the accumulator starts at +0
// and never becomes -0), so one round trip covers up to FOLD_U * 256.
constexpr int FOLD_U = 8;
template <int N, class DI>
__device__ __forceinline__ void fold_partials(DI I, int c0, int nblk, bool is_max, double (&out)[N]) {
  __shared__ double sh[N][RED_NT / 64];
  const int t = threadIdx.x;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const double* red = I->red + (size_t)(c0 + k) * I->red_stride;
    double acc = 0.0;
    for (int i0 = t; i0 < nblk; i0 += RED_NT * FOLD_U) {
      double v[FOLD_U];
#pragma unroll
      for (int u = 0; u < FOLD_U; ++u) v[u] = i0 + u * RED_NT < nblk ? red[i0 + u * RED_NT] : 0.0;
#pragma unroll
      for (int u = 0; u < FOLD_U; ++u) acc = is_max ? fmax(acc, v[u]) : acc + v[u];
    }
    for (int off = 32; off > 0; off >>= 1) {
      const double o = __shfl_xor(acc, off, 64);
      acc = is_max ? fmax(acc, o) : acc + o;
    }
    if ((t & 63) == 0) sh[k][t >> 6] = acc;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < N; ++k) {
    double acc = sh[k][0];
    for (int w = 1; w < RED_NT / 64; ++w) acc = is_max ? fmax(acc, sh[k][w]) : acc + sh[k][w];
    out[k] = acc;
  }
}

// True once the CG solve of this instance has stopped: every kernel of an
// iteration returns at once (the host may queue iterations ahead).
template <class DI>
__device__ __forceinline__ bool cg_done(DI I) { return I->scal[SC_DONE] != 0.0; }

// The stopping test of iteration k = scal[SC_KCUR] with ||r||^2 = rr: stop
// when ||r|| / ||b|| <= rel_tol at a checked iteration (every check_every,
// and the last) or when max_cycles iterations have run.  Every block that
// evaluates it gets the same answer (it reads only values of earlier
// launches); the one with `record` stores the outcome.
template <class DI>
__device__ __forceinline__ bool cg_stop(DI I, double rr, bool record) {
  double* sc = I->scal;
  const uint32_t k = (uint32_t)sc[SC_KCUR];
  bool stop = k >= I->cg_maxc;
  bool checked = false, ok = false;
  double rel = 0.0;
  if (I->cg_tol > 0.0f && (k % I->cg_every == 0u || k >= I->cg_maxc)) {
    const double bb = sc[SC_BB];
    rel = bb > 0.0 ? sqrt(rr / bb) : sqrt(rr);
    ok = rel <= (double)I->cg_tol;
    checked = true;
    stop = stop || ok;
  }
  if (record) {
    sc[SC_RR] = rr;
    sc[SC_ITER] = (double)k;
    if (checked) {
      sc[SC_REL] = rel;
      sc[SC_CONV] = ok ? 1.0 : 0.0;
    }
    if (stop) sc[SC_DONE] = 1.0;
  }
  return stop;
}

}  // namespace efl
