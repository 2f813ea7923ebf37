// efl_internal.h - device-side data layout shared by the HIP kernels and the
// host scheduler of libefl.
//
// One instance = one fluid entity (setup_components.rs:38-330).  Its grids
// live in one device allocation, row-major [y][x]; the MAC face grids are
// u: (W+1) x H and v: W x (H+1).  Multi-channel textures are interleaved
// (area fractions as float4, seeds as float2), matching the rgba32f / rg32f
// texel order of the reference.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/efl.h"

namespace efl {

constexpr int kMaxLevels = 14;
// Reduction tile: kernels that leave one partial per tile (the CG kernels,
// and the level-0 multigrid launches that fuse r . z or ||r||^2) all tile the
// grid this way, so a fold over DevInst::cg_tiles partials reads exactly
// what any of them wrote.  Each side checks it with a static_assert.
constexpr int kRedTileW = 64, kRedTileH = 16;
constexpr int kMaxBatch = EFL_MAX_BATCH;

struct Level {  // one multigrid level (multi_grid.rs:99-118)
  int w, h;
  float* x;
  float* b;
  float* ls;
  float4* af;
  float* r;
  float* xs;  // second buffer of x: the tiled smoothers read one and write the other
  // per-cell stencil, fixed during a solve (k_mg_setup): neighbour
  // coefficients and flags (the Gauss-Seidel denominator is their in-order
  // sum, formed where it is used)
  float4* cf;
  int32_t* fl;
};

// Per-instance descriptor, copied to device memory; kernels read it through
// a pointer (uniform across a block, so it stays in scalar registers).
struct DevInst {
  int W, H;
  // SimulationUniform (fluid_uniform.wgsl:3-10)
  float dx, dt, rho, gx, gy;
  float T[16];
  float sizex, sizey;
  // grids
  float *u0, *u1, *v0, *v1, *us, *vs;
  int32_t* sid;
  int32_t *uval0, *uval1, *vval0, *vval1;
  float *div, *p0, *p1, *ls0, *ls1, *lss;
  float2* grad;
  float4* afs;
  uint32_t *lab0, *lab;
  float *phi_tmp;      // FIM phase buffer
  uint32_t* lab_tmp;   // FIM phase labels
  float2 *seed0, *seed1;
  int32_t* bins;       // 4 x 256 accumulators, zero between passes
  int32_t* bins_last;  // 4 x 256, the last pass's sums (the SOLID_BINS field)
  float4* fts;         // 256 x (fx, fy, torque, 0)
  // sources (uniform-sized), forces, obstacles
  uint32_t nsrc_init, nsrc;
  efl_source* src_init;
  efl_source* src;
  uint32_t nforces;
  efl_force* forces;
  uint32_t nobs;
  const efl_obstacle* obs;
  // multigrid pyramid; level 0 aliases p0/div/ls0/afs.  lev_pcg0 is level 0
  // with x = pcg_z, b = pcg_r, used when the V-cycle preconditions CG.
  int nlev;
  Level lev[kMaxLevels];
  Level lev_pcg0;
  // PCG vectors (W x H) and scalars
  float *pcg_r, *pcg_z, *pcg_q;
  float* pcg_r2;    // second residual buffer of the fused CG iteration (Batch::sel SEL_R2)
  float* pcg_p[2];  // search direction of iteration k in pcg_p[k & 1]
  double* red;         // reduction partials: 4 x red_stride doubles
  int red_stride;      // >= blocks of one full-grid launch
  double* scal;        // scalars, slots SC_* (efl_reduce.h)
  uint32_t* counter;   // k_mg_coarse's group barrier: [0] arrivals, [1] generation, [2] a barrier timed out
  // CG stopping rule, evaluated on the device (projection settings)
  float cg_tol;
  uint32_t cg_maxc, cg_every;
  uint32_t cg_flex;    // flexible beta (learned, nonlinear preconditioner)
  uint32_t prefold;    // more CG partials than kInlineFold: k_prefold folds them
  int cg_tiles;        // 64 x 16 tiles of the CG kernels (partials of one of their reductions)
  float* diag;         // [0] max |D| (as float bits via atomic max), ...
  // fluid diagnostics: per 16 x 16 block volume, max and min speed
  // (3 arrays of (W/16)(H/16) floats), then {volume, min, max}
  float* fdp;
  float* fdr;
};

// Kernels never write a DevInst (the host uploads it between launches):
// read through the constant address space, its fields come in scalar loads,
// which keeps the descriptor reads out of the waits for the data they point
// to.  Values kernels do write (scal[], red[]) are read through the pointers.
using CDevInst = const __attribute__((address_space(4))) DevInst;
__device__ __forceinline__ CDevInst* cdev(const DevInst* I) { return (CDevInst*)I; }
// a level descriptor out of the constant address space, field by field
__device__ __forceinline__ Level lev_copy(const __attribute__((address_space(4))) Level& c) {
  Level r;
  r.w = c.w;
  r.h = c.h;
  r.x = c.x;
  r.b = c.b;
  r.ls = c.ls;
  r.af = c.af;
  r.r = c.r;
  r.xs = c.xs;
  r.cf = c.cf;
  r.fl = c.fl;
  return r;
}
// A scal[] value an earlier launch wrote, in a scalar load: only for slots
// the reading launch does not write before it reads them (the scalar cache
// does not see the launch's own vector stores).
__device__ __forceinline__ double sld(const double* p) { return *(const __attribute__((address_space(4))) double*)p; }

// Fluid diagnostics (volume, min / max cell-centred speed) of one grid:
// the fields, the per-16x16-block partials (3 (W/16)(H/16) floats) and the
// 3 results, passed by value so that an instance's buffers or a caller's
// fields with the library's static scratch can be used alike.
struct FluidDiagArgs {
  const float *ls, *u, *v;
  int W, H;
  float* part;
  float* out;
};

// Arguments of the rigid-body integrator (one launch per obstacle list).
struct BodyStep {
  efl_obstacle* obs;
  const efl_body* bodies;
  int n;
  int ninst;
  const float4* fts[kMaxBatch];
  float inv_dx[kMaxBatch];
  float dx[kMaxBatch];
  float dt, gwx, gwy, half_w, half_h, rho_fluid;
  float4* state;  // per body: previous (ax, ay, angular acceleration)
};

// One instance's own body list for k_integrate_own_bodies (the fields of
// BodyStep for a single instance).
struct BodyList {
  efl_obstacle* obs;
  const efl_body* bodies;
  float4* state;
  const float4* fts;
  int n;
  float dx, gwx, gwy, half_w, half_h, rho_fluid;
};
constexpr int kMaxBodyLists = 32;  // per launch (kernel argument size)
struct BodyLists {
  int n;
  float dt;
  BodyList list[kMaxBodyLists];
};

// Level::fl bits, written by k_mg_setup
constexpr int MG_IN0 = 0;           // bits 0-3: neighbour k (-x, +x, -y, +y) inside the level
constexpr int MG_FL0 = 4;           // bits 4-7: neighbour k inside the level and fluid (ls < 0)
constexpr int MG_GSZ = 1 << 8;      // the GS update gives 0 (ls >= 0, closed, |denominator| < 1e-6)
constexpr int MG_RZ = 1 << 9;       // the residual is 0 (ls >= 0 or closed)
constexpr int MG_LSNEG = 1 << 10;   // ls < 0 (prolongation target, z mask)
constexpr int MG_CLOSED = 1 << 11;  // all four area fractions are 0
constexpr int MG_UNIT = 1 << 12;    // the stencil is (1, 1, 1, 1)

// Whether a cell's stencil has to be read: an MG_UNIT stencil is known, and
// one whose update and residual are both forced to 0 (MG_GSZ and MG_RZ) is
// never used; mg_cf_implied stands in for it then (same results).
__host__ __device__ __forceinline__ bool mg_cf_needed(int fl) {
  return !(fl & MG_UNIT) && (fl & (MG_GSZ | MG_RZ)) != (MG_GSZ | MG_RZ);
}
__host__ __device__ __forceinline__ float4 mg_cf_implied(int fl) {
  const float v = (fl & MG_UNIT) ? 1.0f : 0.0f;
  return make_float4(v, v, v, v);
}

// Workgroups go to the 8 XCDs round robin (block i to XCD i mod 8), and
// each XCD has its own L2.  Renumbering gives XCD k a contiguous band of
// tiles, so halos that neighbouring tiles share are read through one L2.
__device__ __forceinline__ int xcd_tile(int blk, int nb) {
  const int q = nb >> 3;
  return blk < 8 * q ? (blk & 7) * q + (blk >> 3) : blk;
}

// The slot whose block range holds block blk: the last i < n with
// start[i] <= blk (start[0] = 0, non-decreasing; a slot without blocks has
// start[i] = start[i + 1] and is passed over).  A binary search: 6 dependent
// scalar loads for 64 slots where a scan took up to 63.
__device__ __forceinline__ int slot_of(const int* start, int n, int blk) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= blk) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Launch map of a batched kernel: the flat grid of blocks is split into
// per-instance ranges; inside a range, blocks tile the instance's extent.
// Batch::sel bits of the level-0 / level-1 CG preconditioner launches (bits
// 0..2 are the fused multigrid's buffer selection): SEL_AXPY on the first
// level-0 k_mg_down of an iteration: it also runs the iteration itself (p,
// A p, x and r; r read from the buffer SEL_R2 names, written to the other);
// on the level-1 k_mg_down after it: it runs the stopping test.  SEL_R2:
// level-0 b (the residual) is in pcg_r2; a fused solve's residual changes
// buffer every iteration (the first reads pcg_r), so on its SEL_AXPY launch
// SEL_R2 also marks an even iteration, which names the buffers of p (p of
// iteration k in pcg_p[k & 1]) and the scalar slots the launch reads and
// writes.  SEL_LAZY (any tiled multigrid launch): the level
// has 2^20 cells or more, so its launches are bound by the bytes they stream
// from HBM (smaller ones by latency); a launch whose instances all carry it
// reads only the stencils the flags need (mg_cf_needed).
constexpr int SEL_AXPY = 8, SEL_R2 = 16, SEL_LAZY = 32;

// Launch of k_mg_coarse (efl_mg.hip): for each instance a group of
// workgroups, one per tile of its level l0 (the first level of at most
// kCoarseTiles 64 x 16 tiles), runs the V-cycle from l0 down through the
// bottom and back in one launch, the phases separated by a barrier of the
// group.  sel holds the Batch::sel bits of each phase in order: the down
// launches of l0 .. lb - 1, the bottom, the up launches of lb - 1 .. l0.
constexpr int kCoarseTiles = 16;
constexpr int kCoarseMaxLevels = 6;  // tiled levels of one fused cycle (lb - l0)
struct CoarseBatch {
  int n;
  int start[kMaxBatch + 1];  // first workgroup of each instance's group
  const DevInst* inst[kMaxBatch];
  int8_t l0[kMaxBatch], lb[kMaxBatch], coarse[kMaxBatch];
  uint8_t sel[kMaxBatch][2 * kCoarseMaxLevels + 1];
};

struct Batch {
  int n;
  int nblocks;  // start[n], beside n so that one scalar load brings both
  int start[kMaxBatch + 1];
  int tiles_x[kMaxBatch];
  int ex[kMaxBatch], ey[kMaxBatch];
  int level[kMaxBatch];
  int sel[kMaxBatch];  // fused multigrid: bit 0 x is in Level::xs, bit 1 the coarse x is
  const DevInst* inst[kMaxBatch];
};

}  // namespace efl
