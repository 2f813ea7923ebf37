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

// scal[] slots
constexpr int SC_PQ = 1, SC_RR = 2, SC_BB = 3, SC_ALPHA = 4, SC_BETA = 5, SC_DONE = 6;
constexpr int SC_NRZ = 7;            // number of r . z partials of the last preconditioner
constexpr int SC_ITER = 9, SC_REL = 10, SC_CONV = 11;
constexpr int SC_RZ0 = 12;           // r . z of iteration k at SC_RZ0 + (k & 1)

// Fold the nblk partials of components c0 .. c0 + N - 1 (all threads of a
// block of RED_NT call it; every thread gets the sums).
template <int N>
__device__ __forceinline__ void fold_partials(const DevInst* I, int c0, int nblk, bool is_max, double (&out)[N]) {
  __shared__ double sh[N][RED_NT];
  const int t = threadIdx.x;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    double acc = 0.0;
    for (int i = t; i < nblk; i += RED_NT) {
      const double v = I->red[(size_t)(c0 + k) * I->red_stride + i];
      acc = is_max ? fmax(acc, v) : acc + v;
    }
    sh[k][t] = acc;
  }
  __syncthreads();
  for (int st = RED_NT / 2; st > 0; st >>= 1) {
    if (t < st)
#pragma unroll
      for (int k = 0; k < N; ++k) sh[k][t] = is_max ? fmax(sh[k][t], sh[k][t + st]) : sh[k][t] + sh[k][t + st];
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < N; ++k) out[k] = sh[k][0];
}

// True once the CG solve of this instance has stopped: every kernel of an
// iteration returns at once (the host may queue iterations ahead).
__device__ __forceinline__ bool cg_done(const DevInst* I) { return I->scal[SC_DONE] != 0.0; }

}  // namespace efl
