// efl_mg.hip - fused multigrid for gfx950: the V-cycle of
// projection/multi_grid.rs:390-475 (and its symmetric variant used as the
// CG preconditioner) in two launches per level plus one for the bottom.
//
//   mg_setup   once per projection and level: the stencil of every cell
//              (the 4 neighbour coefficients and the flag word that
//              gauss_seidel.wgsl:44-88 and residual.wgsl:12-57 derive from
//              the air level set and the area fractions, which do not change
//              during a solve; the denominator is the coefficients' in-order
//              sum) and, for the next level, restriction.wgsl's level set and
//              area fractions.
//   mg_down    a block owns a 64 x 16 tile of one level; it stages x of the
//              tile plus a halo in LDS and keeps the stencils of its cells in
//              registers, runs the pre-smoothing half-sweeps (each one
//              shrinks the region whose values are exact by one cell),
//              writes the tile and, with `restrict`, forms the residual of
//              its cells and restricts it to the 32 x 8 coarse cells under
//              the tile; r is never stored.
//   mg_up      prolongation.wgsl applied while staging, then the
//              post-smoothing half-sweeps; at level 0 of the CG
//              preconditioner it also masks z and forms the block partial
//              of r . z.
//   bottom     one workgroup runs the V-cycle below the first level of at
//              most 1024 cells in LDS (reference order, no launches).
//
// Red-black Gauss-Seidel inside a colour is order independent, and the
// precomputed stencil holds exactly the values the WGSL recomputes at every
// sweep (same expressions, same summation order, -ffp-contract=off), so
// these kernels produce the values of the one-kernel-per-half-sweep
// reference dataflow bit for bit.
//
// A block's halo belongs to other blocks' tiles, so a tiled launch never
// writes the x it reads: x and Level::xs form a pair, each launch reads one
// (Batch::sel) and writes every in-level cell of its tiles to the other.
// Writing in place would let a block that starts late stage neighbour
// values that were already smoothed.  Batch::sel bit 2 says that x is zero
// (after a restriction, or z at the start of the preconditioner): then it is
// not read at all.
#include <hip/hip_runtime.h>
#include <cstring>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "efl_internal.h"
#include "efl_launch.h"
#include "efl_reduce.h"

namespace efl {

constexpr int MT_W = 64;   // tile width (one wavefront per tile row)
constexpr int MT_H = 16;   // tile height of the launches that leave reduction partials
// Launches without reductions (every level of a plain multigrid or Gauss-
// Seidel solve, and the CG preconditioner below level 0) take the tile
// height of the instance's tuning: MT_H or MT_H8.
constexpr int MT_H8 = 8;
constexpr int MNT = 256;   // threads per block
static_assert(MNT == RED_NT, "the r . z reduction uses RED_NT threads per block");
static_assert(MT_W == kRedTileW && MT_H == kRedTileH,
              "k_mg_up / k_mg_down leave one r . z / ||r||^2 partial per tile; folds count DevInst::cg_tiles");

// Level::fl bits (efl_internal.h)
constexpr int F_IN0 = MG_IN0, F_FL0 = MG_FL0, F_GSZ = MG_GSZ, F_RZ = MG_RZ, F_LSNEG = MG_LSNEG, F_CLOSED = MG_CLOSED;

__device__ __forceinline__ bool mg_inb(int x, int y, int w, int h) {
  return (unsigned)x < (unsigned)w && (unsigned)y < (unsigned)h;
}
__device__ __forceinline__ bool closed4(float4 f) { return f.x == 0.0f && f.y == 0.0f && f.z == 0.0f && f.w == 0.0f; }

// Level arrays are reached through pointers held in device memory, which the
// compiler can only treat as generic (flat) addresses; these name the global
// address space so the accesses become global_load / global_store.
typedef float f4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float ldg(const float* p, size_t i) {
  return ((const __attribute__((address_space(1))) float*)p)[i];
}
__device__ __forceinline__ int ldg(const int32_t* p, size_t i) {
  return ((const __attribute__((address_space(1))) int32_t*)p)[i];
}
__device__ __forceinline__ float4 ldg(const float4* p, size_t i) {
  const f4v v = ((const __attribute__((address_space(1))) f4v*)p)[i];
  return make_float4(v.x, v.y, v.z, v.w);
}
// Loads at a 32-bit element index from a uniform base (global_load with an
// SGPR base and a 32-bit VGPR offset: one VGPR per address instead of two).
// Every level array is far below 4 GiB.
template <class T>
__device__ __forceinline__ T ld32(const T* p, uint32_t i) {
  using G = const __attribute__((address_space(1))) char;
  return *(const __attribute__((address_space(1))) T*)((G*)p + i * (uint32_t)sizeof(T));
}
__device__ __forceinline__ float4 ld32(const float4* p, uint32_t i) {
  using G = const __attribute__((address_space(1))) char;
  const f4v v = *(const __attribute__((address_space(1))) f4v*)((G*)p + i * 16u);
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void stg(float* p, size_t i, float v) {
  ((__attribute__((address_space(1))) float*)p)[i] = v;
}
__device__ __forceinline__ void stg(int32_t* p, size_t i, int v) {
  ((__attribute__((address_space(1))) int32_t*)p)[i] = v;
}
__device__ __forceinline__ void stg(float4* p, size_t i, float4 v) {
  const f4v w = {v.x, v.y, v.z, v.w};
  ((__attribute__((address_space(1))) f4v*)p)[i] = w;
}

__device__ __forceinline__ const Level& mg_level(const DevInst* I, int l, int pcg) {
  return (pcg && l == 0) ? I->lev_pcg0 : I->lev[l];
}
__device__ __forceinline__ Level mg_level(CDevInst* I, int l, int pcg) {
  return (pcg && l == 0) ? lev_copy(I->lev_pcg0) : lev_copy(I->lev[l]);
}

// dx * dx * rho / dt and dt / rho / dx / dx, evaluated left to right as the
// WGSL does
__device__ __forceinline__ float gs_c0(const DevInst* I, float scale) {
  float dx = scale * I->dx;
  return dx * dx * I->rho / I->dt;
}
__device__ __forceinline__ float resid_factor(const DevInst* I, float scale) {
  float dx = I->dx * scale;
  return I->dt / I->rho / dx / dx;
}
__device__ __forceinline__ float gs_c0(CDevInst* I, float scale) {
  float dx = scale * I->dx;
  return dx * dx * I->rho / I->dt;
}
__device__ __forceinline__ float resid_factor(CDevInst* I, float scale) {
  float dx = I->dx * scale;
  return I->dt / I->rho / dx / dx;
}

// bit `bit` of fl set ? a : b, as v_bfe_i32 (the bit spread to a lane mask)
// and v_bfi_b32 in one block: the flag tests stay per-lane VALU work, where
// the compiler would otherwise hoist them out of the unrolled sweeps as
// 64-bit lane masks and spill those through VGPR lanes.
static_assert(F_IN0 == 0 && F_GSZ == 1 << 8 && F_RZ == 1 << 9, "flag bits fsel is called with");
template <int BIT>
__device__ __forceinline__ float fsel(int fl, float a, float b) {
  float r;
  asm("v_bfe_i32 %0, %1, %2, 1\n\tv_bfi_b32 %0, %0, %3, %4" : "=&v"(r) : "v"(fl), "I"(BIT), "v"(a), "v"(b));
  return r;
}

// the Gauss-Seidel denominator from the stencil: the in-level coefficients
// summed in neighbour order, as k_mg_setup forms it (F_IN0 is bit 0)
__device__ __forceinline__ float b_denominator(float4 c, int fl) {
  float d = 0.0f;
  d = fsel<0>(fl, d + c.x, d);
  d = fsel<1>(fl, d + c.y, d);
  d = fsel<2>(fl, d + c.z, d);
  d = fsel<3>(fl, d + c.w, d);
  return d;
}

// gauss_seidel.wgsl:44-88 with the stencil precomputed: nume = c0 b plus the
// fluid neighbours' c_k x_k in neighbour order, then nume / denominator and
// the relaxation.  A skipped term leaves the sum untouched.
__device__ __forceinline__ float gs_upd(float b, float4 c, float dn, int fl, float xo, float x0, float x1, float x2,
                                        float x3, float weight, float c0) {
  float nume = c0 * b;
  nume = fsel<F_FL0 + 0>(fl, nume + c.x * x0, nume);
  nume = fsel<F_FL0 + 1>(fl, nume + c.y * x1, nume);
  nume = fsel<F_FL0 + 2>(fl, nume + c.z * x2, nume);
  nume = fsel<F_FL0 + 3>(fl, nume + c.w * x3, nume);
  const float pn = nume / dn;
  const float v = weight * pn + (1.0f - weight) * xo;
  return fsel<8>(fl, 0.0f, v);  // F_GSZ
}

// residual.wgsl:12-57 with the stencil precomputed: fluid neighbours
// subtract c_k (x - x_k) factor, air neighbours c_k x factor.  The operand
// (x - x_k or x) is selected first, so each face costs one product chain;
// the rounding is that of the two separate expressions.
__device__ __forceinline__ float res_upd(float b, float4 c, int fl, float xc, float x0, float x1, float x2, float x3,
                                         float factor) {
  float res = b;
  res = fsel<F_IN0 + 0>(fl, res - c.x * fsel<F_FL0 + 0>(fl, xc - x0, xc) * factor, res);
  res = fsel<F_IN0 + 1>(fl, res - c.y * fsel<F_FL0 + 1>(fl, xc - x1, xc) * factor, res);
  res = fsel<F_IN0 + 2>(fl, res - c.z * fsel<F_FL0 + 2>(fl, xc - x2, xc) * factor, res);
  res = fsel<F_IN0 + 3>(fl, res - c.w * fsel<F_FL0 + 3>(fl, xc - x3, xc) * factor, res);
  return fsel<9>(fl, 0.0f, res);  // F_RZ
}

// ---------------------------------------------------------------------------
// setup: stencils of one level (+ level set and area fractions of the next)
// ---------------------------------------------------------------------------
constexpr int SU_W = 64, SU_H = 4;

// Where a level's level set and area fractions are read (global memory or
// LDS), and where the next level's are written (global memory, or global
// memory and LDS).
struct SetupGlobalSrc {
  const float* ls;
  const float4* af;
  __device__ float l(size_t j) const { return ldg(ls, j); }
  __device__ float4 a(size_t j) const { return ldg(af, j); }
};
struct SetupLdsSrc {
  const float* ls;
  const float4* af;
  __device__ float l(size_t j) const { return ls[j]; }
  __device__ float4 a(size_t j) const { return af[j]; }
};
struct SetupGlobalDst {
  __device__ void put(const Level& C, size_t cid, float pr, float4 a) const {
    stg(C.ls, cid, pr);
    stg(C.af, cid, a);
  }
};
struct SetupLdsDst {
  float* ls;
  float4* af;
  __device__ void put(const Level& C, size_t cid, float pr, float4 a) const {
    stg(C.ls, cid, pr);
    stg(C.af, cid, a);
    ls[cid] = pr;
    af[cid] = a;
  }
};

// One cell (gx, gy) of level l: its stencil and flags, and, when `coarse`
// and the cell is a (0, 0) child, the next level's level set and area
// fractions.
template <class Src, class Dst>
__device__ __forceinline__ void setup_cell(const DevInst* I, int l, const Level& L, int gx, int gy, bool coarse,
                                           const Src& src, const Dst& dst) {
  const size_t id = (size_t)gy * L.w + gx;
  const float lc = src.l(id);
  const float4 f = src.a(id);
  const bool in[4] = {gx > 0, gx + 1 < L.w, gy > 0, gy + 1 < L.h};
  const size_t nb[4] = {id - 1, id + 1, id - L.w, id + L.w};
  const float fv[4] = {f.x, f.y, f.z, f.w};
  float cv[4];
  float denom = 0.0f;
  int fl = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    cv[k] = 0.0f;
    if (in[k]) {
      const float lv = src.l(nb[k]);
      const bool fluid = lv < 0.0f;
      cv[k] = fluid ? fv[k] : fv[k] * (1.0f - lv / lc);
      denom += cv[k];
      fl |= (1 << (F_IN0 + k)) | (fluid ? 1 << (F_FL0 + k) : 0);
    }
  }
  const bool closed = closed4(f);
  if (lc >= 0.0f || closed || fabsf(denom) < 1.0e-6f) fl |= F_GSZ;
  if (lc >= 0.0f || closed) fl |= F_RZ;
  if (lc < 0.0f) fl |= F_LSNEG;
  if (closed) fl |= F_CLOSED;
  if (cv[0] == 1.0f && cv[1] == 1.0f && cv[2] == 1.0f && cv[3] == 1.0f) fl |= MG_UNIT;
  stg(L.cf, id, make_float4(cv[0], cv[1], cv[2], cv[3]));
  stg(L.fl, id, fl);
  // restriction.wgsl:10-60 of the level set and the area fractions, by the
  // thread of the coarse cell's (0, 0) child
  if (!coarse || (gx & 1) || (gy & 1)) return;
  const Level C = I->lev[l + 1];
  const int cx = gx >> 1, cy = gy >> 1;
  if (cx >= C.w || cy >= C.h) return;
  const int dxs[4] = {0, 0, 1, 1}, dys[4] = {0, 1, 0, 1};  // (0,0) (0,1) (1,0) (1,1)
  float l4[4];
  float4 a4[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int x = gx + dxs[k], y = gy + dys[k];
    const bool inl = mg_inb(x, y, L.w, L.h);
    const size_t j = (size_t)y * L.w + x;
    l4[k] = inl ? src.l(j) : 0.0f;
    a4[k] = inl ? src.a(j) : make_float4(0.f, 0.f, 0.f, 0.f);
  }
  float pr = l4[0];
  if (fabsf(l4[1]) < fabsf(pr)) pr = l4[1];
  if (fabsf(l4[2]) < fabsf(pr)) pr = l4[2];
  if (fabsf(l4[3]) < fabsf(pr)) pr = l4[3];
  const size_t cid = (size_t)cy * C.w + cx;
  dst.put(C, cid, pr,
          make_float4(0.5f * (a4[0].x + a4[1].x), 0.5f * (a4[2].y + a4[3].y), 0.5f * (a4[0].z + a4[2].z),
                      0.5f * (a4[1].w + a4[3].w)));
}

__global__ __launch_bounds__(SU_W* SU_H) void k_mg_setup(Batch b) {
  int blk = blockIdx.x;
  const int i = slot_of(b.start, b.n, blk);
  const int lb = blk - b.start[i];
  const DevInst* I = b.inst[i];
  const int l = b.level[i];
  const Level L = I->lev[l];
  const int gx = (lb % b.tiles_x[i]) * SU_W + (int)(threadIdx.x % SU_W);
  const int gy = (lb / b.tiles_x[i]) * SU_H + (int)(threadIdx.x / SU_W);
  if (gx >= L.w || gy >= L.h) return;
  setup_cell(I, l, L, gx, gy, (b.sel[i] & 1) != 0, SetupGlobalSrc{L.ls, L.af}, SetupGlobalDst{});
}

// Levels b.level[n] .. nlev - 1 of instance n in one workgroup, level after
// level.  The first level reads its inputs from global memory; each level's
// restriction writes the next level's to global memory and to LDS, where the
// next level reads them after the barrier (levels at odd offsets from the
// first in the SC_B buffers, at even offsets in the SC_A ones, which the host
// checks they fit).  For the coarse levels, where a tiled launch per level
// costs more than the work.
constexpr int SC_NT = 1024, SC_B = 2048, SC_A = 512;
__global__ __launch_bounds__(SC_NT) void k_mg_setup_coarse(Batch b) {
  __shared__ float lsB[SC_B], lsA[SC_A];
  __shared__ float4 afB[SC_B], afA[SC_A];
  const DevInst* I = b.inst[blockIdx.x];
  const int nlev = I->nlev, l0 = b.level[blockIdx.x];
  {
    const Level L = I->lev[l0];
    const int n = L.w * L.h;
    for (int c = threadIdx.x; c < n; c += SC_NT)
      setup_cell(I, l0, L, c % L.w, c / L.w, l0 + 1 < nlev, SetupGlobalSrc{L.ls, L.af}, SetupLdsDst{lsB, afB});
    __syncthreads();
  }
  for (int l = l0 + 1; l < nlev; ++l) {
    const Level L = I->lev[l];
    const int n = L.w * L.h;
    const bool odd = ((l - l0) & 1) != 0;
    const SetupLdsSrc src = odd ? SetupLdsSrc{lsB, afB} : SetupLdsSrc{lsA, afA};
    const SetupLdsDst dst = odd ? SetupLdsDst{lsA, afA} : SetupLdsDst{lsB, afB};
    for (int c = threadIdx.x; c < n; c += SC_NT) setup_cell(I, l, L, c % L.w, c / L.w, l + 1 < nlev, src, dst);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// tiled smoothers
// ---------------------------------------------------------------------------

struct MSite {
  const DevInst* I;
  int blk;       // tile index in the launch
  int tx0, ty0;  // tile origin (cells)
  int lvl;
  int ex, ey;    // update extent
  int slot;
  int sel;       // buffer selectors (Batch::sel)
};

// Slot 0's fields are read together with the slot count, unconditionally, so
// that a launch of one instance (the usual case) has its tile origin and
// instance after one round of scalar loads; other slots are looked up after.
template <int TH = MT_H>
__device__ __forceinline__ MSite mg_locate(const Batch& b) {
  const int n = b.n, nb = b.nblocks;
  int st = b.start[0], tx = b.tiles_x[0];
  MSite s;
  s.I = b.inst[0];
  s.lvl = b.level[0];
  s.ex = b.ex[0];
  s.ey = b.ey[0];
  s.sel = b.sel[0];
  s.slot = 0;
  // keeps the compiler from sinking these loads into the paths below
  asm volatile("" ::"s"(st), "s"(tx), "s"(s.I), "s"(s.lvl), "s"(s.ex), "s"(s.ey), "s"(s.sel));
  const int blk = xcd_tile((int)blockIdx.x, nb);
  if (n > 1) {
    const int i = slot_of(b.start, n, blk);
    st = b.start[i];
    tx = b.tiles_x[i];
    s.I = b.inst[i];
    s.lvl = b.level[i];
    s.ex = b.ex[i];
    s.ey = b.ey[i];
    s.sel = b.sel[i];
    s.slot = i;
  }
  const int lb = blk - st;
  s.blk = blk;
  s.tx0 = (lb % tx) * MT_W;
  s.ty0 = (lb / tx) * TH;
  return s;
}

// The cells of a region [tile + halo] a thread owns: slots 0..NIT-1 hold
// cells of the colour swept first, NIT..2NIT-1 the other colour; slot k of a
// colour is the thread's k-th cell in a row-major walk over that colour's
// cells.  Stencil, b and x of its cells live in registers; x is mirrored in
// LDS for the neighbours.  The tile is TW x TH cells (TW + 2 HALO even).
template <int HALO, int TH = MT_H, int TW = MT_W>
struct Tile {
  static constexpr int RW = TW + 2 * HALO, RH = TH + 2 * HALO, HW = RW / 2;
  static_assert(RW % 2 == 0, "rows of whole colour pairs");
  static constexpr int NC = HW * RH;                 // cells of one colour
  static constexpr int NIT = (NC + MNT - 1) / MNT;   // slots per colour
  static constexpr int NSL = 2 * NIT;
  float b[NSL], x[NSL], dn[NSL];  // dn: the Gauss-Seidel denominator of the cell's stencil
  float4 c[NSL];
  int fl[NSL];
  int lx[NSL], ly[NSL];
  bool valid[NSL], inl[NSL];

  // ox, oy: level coordinates of LDS cell (0, 0)
  __device__ __forceinline__ void place(int ox, int oy, int first, int w, int h) {
    const int base = (ox + oy) & 1;
#pragma unroll
    for (int s = 0; s < NSL; ++s) {
      const int colour = s < NIT ? first : 1 - first;
      const int u = (int)threadIdx.x + (s % NIT) * MNT;
      valid[s] = u < NC;
      ly[s] = u / HW;
      lx[s] = 2 * (u - ly[s] * HW) + ((colour + ly[s] + base) & 1);
      inl[s] = valid[s] && mg_inb(ox + lx[s], oy + ly[s], w, h);
    }
  }

  // Stencils, b and x from the level (x = 0 when xzero); with C, the
  // coarse x (cxin) the prolongation adds.  The loads are straight-line (a cell outside the level reads index
  // 0) and their values are masked only in stage(), so no load waits for an
  // earlier one and every wait covers just the loads it needs.  They are
  // issued in the order they are consumed: the flags, x and the coarse x
  // (the LDS copy), then b and the coefficients.  With lazy_cf the
  // coefficients are left for load_cf.
  float xc[NSL];
  bool par[NSL];
  __device__ __forceinline__ void load(const Level& L, const float* xin, bool xzero, const Level* C,
                                       const float* cxin, int ox, int oy, bool lazy_cf = false) {
#pragma unroll
    for (int s = 0; s < NSL; ++s) {
      const int gx = ox + lx[s], gy = oy + ly[s];
      const uint32_t id = inl[s] ? (uint32_t)(gy * L.w + gx) : 0u;
      fl[s] = ld32(L.fl, id);
      x[s] = xzero ? 0.0f : ld32(xin, id);
      par[s] = false;
      xc[s] = 0.0f;
      if (C) {
        const int cx = gx >> 1, cy = gy >> 1;
        par[s] = inl[s] && cx < C->w && cy < C->h;
        xc[s] = ld32(cxin, par[s] ? (uint32_t)(cy * C->w + cx) : 0u);
      }
    }
#pragma unroll
    for (int s = 0; s < NSL; ++s) {
      const uint32_t id = inl[s] ? (uint32_t)((oy + ly[s]) * L.w + ox + lx[s]) : 0u;
      b[s] = ld32(L.b, id);
      if (!lazy_cf) c[s] = ld32(L.cf, id);
    }
  }
  // Every in-level cell's coefficients (load with lazy_cf left them out).
  __device__ __forceinline__ void load_c(const Level& L, int ox, int oy) {
#pragma unroll
    for (int s = 0; s < NSL; ++s) c[s] = ld32(L.cf, inl[s] ? (uint32_t)((oy + ly[s]) * L.w + ox + lx[s]) : 0u);
  }
  // The coefficients where the flags need them (mg_cf_needed; stage() puts
  // mg_cf_implied in the others): on a level that streams from HBM most
  // cells (open fluid, air) need none, a third of the bytes.  A cell that
  // needs none reads index 0, which the wavefront's other such lanes share.
  // These loads wait for the flags, so they go after every other load.
  __device__ __forceinline__ void load_cf(const Level& L, int ox, int oy) {
#pragma unroll
    for (int s = 0; s < NSL; ++s) {
      const bool need = inl[s] && mg_cf_needed(fl[s]);
      c[s] = ld32(L.cf, need ? (uint32_t)((oy + ly[s]) * L.w + ox + lx[s]) : 0u);
    }
  }
  // Cells outside the level take flags, x, b and coefficients 0 (after
  // load_cf, lazy_cf: the implied coefficients where none were read); then
  // the prolongation (x += coarse x on fluid cells with a parent), the LDS
  // copy of x and the denominators.
  __device__ __forceinline__ void stage(const Level* C, float* sx, bool lazy_cf = false) {
#pragma unroll
    for (int s = 0; s < NSL; ++s) {
      if (lazy_cf && inl[s] && !mg_cf_needed(fl[s])) c[s] = mg_cf_implied(fl[s]);
      if (!inl[s]) {
        fl[s] = 0;
        x[s] = b[s] = 0.0f;
        c[s] = make_float4(0.f, 0.f, 0.f, 0.f);
      }
      if (C && par[s] && (fl[s] & F_LSNEG)) x[s] = xc[s] + x[s];
      if (valid[s]) sx[ly[s] * RW + lx[s]] = x[s];
    }
#pragma unroll
    for (int s = 0; s < NSL; ++s) dn[s] = b_denominator(c[s], fl[s]);
  }

  // The LDS index of slot s, or of an interior cell (read, never used) when
  // the slot is inactive, so that every slot's neighbour loads can be issued
  // unconditionally and all at once.
  __device__ __forceinline__ int lds_at(int s, bool act) const { return act ? ly[s] * RW + lx[s] : RW + 1; }

  // 2 * NS half-sweeps (first colour = the one placed in slots 0..NIT-1).
  // Half-sweep hs updates cells at least hs + 1 from the region border (their
  // neighbours are still exact) inside the level and [0, ex) x [0, ey).
  template <int NS>
  __device__ __forceinline__ void sweep(int ox, int oy, int w, int h, int ex, int ey, float weight, float c0,
                                        float* sx) {
    const int bx0 = -ox, by0 = -oy, bx1 = min(ex, w) - ox, by1 = min(ey, h) - oy;
#pragma unroll
    for (int hs = 0; hs < 2 * NS; ++hs) {
      const int m = hs + 1;
      const int lo_x = max(m, bx0), hi_x = min(RW - m, bx1), lo_y = max(m, by0), hi_y = min(RH - m, by1);
      const int s0 = (hs & 1) ? NIT : 0;
      // The cells a half-sweep updates and the neighbours they read are of
      // different colours: the neighbours of every slot are loaded first
      // (the loads overlap), then each slot updates and stores.
      bool act[NIT];
      int q[NIT];
      float nb[NIT][4];
#pragma unroll
      for (int k = 0; k < NIT; ++k) {
        const int s = s0 + k;
        act[k] = valid[s] && lx[s] >= lo_x && lx[s] < hi_x && ly[s] >= lo_y && ly[s] < hi_y;
        q[k] = lds_at(s, act[k]);
        nb[k][0] = sx[q[k] - 1];
        nb[k][1] = sx[q[k] + 1];
        nb[k][2] = sx[q[k] - RW];
        nb[k][3] = sx[q[k] + RW];
      }
#pragma unroll
      for (int k = 0; k < NIT; ++k) {
        const int s = s0 + k;
        if (act[k]) {
          const float nv = gs_upd(b[s], c[s], dn[s], fl[s], x[s], nb[k][0], nb[k][1], nb[k][2], nb[k][3], weight, c0);
          x[s] = nv;
          sx[q[k]] = nv;
        }
      }
      __syncthreads();
    }
  }

  __device__ __forceinline__ bool in_tile(int s) const {
    return valid[s] && lx[s] >= HALO && lx[s] < HALO + TW && ly[s] >= HALO && ly[s] < HALO + TH;
  }

  // The residual of every tile cell (0 outside the level) into sr, row-major
  // over the tile; neighbour loads first, as in sweep.
  __device__ __forceinline__ void residual(float factor, const float* sx, float* sr) const {
    bool act[NSL];
    int q[NSL];
    float nb[NSL][4];
#pragma unroll
    for (int s = 0; s < NSL; ++s) {
      act[s] = in_tile(s);
      q[s] = lds_at(s, act[s]);
      nb[s][0] = sx[q[s] - 1];
      nb[s][1] = sx[q[s] + 1];
      nb[s][2] = sx[q[s] - RW];
      nb[s][3] = sx[q[s] + RW];
    }
#pragma unroll
    for (int s = 0; s < NSL; ++s) {
      if (!act[s]) continue;
      const float r = inl[s] ? res_upd(b[s], c[s], fl[s], x[s], nb[s][0], nb[s][1], nb[s][2], nb[s][3], factor) : 0.0f;
      sr[(ly[s] - HALO) * TW + lx[s] - HALO] = r;
    }
  }
};

// Every in-level cell of the tile goes to xout, updated or not, so xout is a
// complete copy of x after the launch (row-major from LDS: coalesced).
template <int HALO, int TH = MT_H>
__device__ __forceinline__ void write_tile(const Level& L, float* xout, int x0, int y0, const float* sx) {
  constexpr int RW = MT_W + 2 * HALO;
  static_assert((MT_W * TH) % MNT == 0, "whole rows per thread round");
#pragma unroll
  for (int k = 0; k < (MT_W * TH) / MNT; ++k) {
    int t = threadIdx.x + k * MNT;
    int lx = t % MT_W, ly = t / MT_W;
    int gx = x0 + lx, gy = y0 + ly;
    if (gx < L.w && gy < L.h) stg(xout, (size_t)gy * L.w + gx, sx[(ly + HALO) * RW + lx + HALO]);
  }
}

// ---------------------------------------------------------------------------
// down: NS pre-smoothing sweeps (+ residual + restriction of b).  The halo is
// one cell wider than the sweeps need so that the residual of every tile
// cell sees exact neighbours.
// ---------------------------------------------------------------------------
//
// With SEL_AXPY at level 0 (the AX instantiation) the launch first runs a
// whole CG iteration in the Chronopoulos-Gear form (the preconditioner's
// r . z = gamma and z . A z = delta, partials of the k_mg_up launch that
// ended it): every block folds gamma and delta, beta = gamma / gamma_prev,
// alpha = gamma / (delta - beta gamma / alpha_prev) (= gamma / p . A p in
// exact arithmetic; the first iteration takes p = z and gamma / delta),
// forms p = z + beta p_prev over its region and a one-cell ring in LDS and
// q = A p over the region (the operator of residual.wgsl with b = 0,
// negated), stages b = r - alpha q (r from the buffer SEL_R2 names), writes
// the new r of its tile to the other buffer, p of its tile to
// pcg_p[iter & 1] and x += alpha p, and leaves the partial of ||r||^2; the
// stopping test then runs in the level-1 launch that follows (SEL_AXPY
// there).  A neighbour's halo reads the old r and p, which no block of the
// launch writes, so the tiles need no order.  One launch thus replaces
// k_pcg_apply + k_pcg_axpy: the only global sums an iteration needs are
// those the preconditioner's last launch already leaves.
template <int NS, int HALO, bool AX = false, int TH = MT_H, bool LZ = false>
__global__ __launch_bounds__(MNT) void k_mg_down(Batch b, int pcg, int first, float weight, int restrict_,
                                                 int cg_check) {
  static_assert(!AX || TH == MT_H, "the fused CG update leaves one partial per 64 x 16 tile");
  using T = Tile<HALO, TH>;
  // AX: z and p_prev over the region and a one-cell ring (PW x PH), then p
  // in place of z; the residual of the restriction reuses the first array
  constexpr int PW = T::RW + 2, PH = T::RH + 2, PN = PW * PH, PIT = (PN + MNT - 1) / MNT;
  constexpr int SRN = AX ? (PN > MT_W * TH ? PN : MT_W * TH) : MT_W * TH;
  __shared__ float sx[T::RW * T::RH];
  __shared__ float sr[SRN];
  __shared__ float spo[AX ? PN : 1];
  __shared__ double wsum[MNT / 64];
  const MSite s = mg_locate<TH>(b);
  const DevInst* I = s.I;
  CDevInst* CI = cdev(I);
  double* const scal = CI->scal;
  // The done flag (written by an earlier launch, usually not in this XCD's
  // L2) is read first but tested only after the tile loads are in flight, so
  // the two latencies overlap; returning then is uniform for the block.  The
  // iteration and r . z of the CG update come from earlier launches too.
  const double done_flag = sld(scal + SC_DONE);
  Level L = mg_level(CI, s.lvl, pcg);
  // CG bits count in preconditioner launches only.  SEL_R2: the residual is
  // in pcg_r2; in a fused solve that marks the even iterations, so the
  // iteration's buffers and scalar slots follow from it without waiting for
  // the iteration count (which only the fold needs).
  const bool rsel = pcg && s.lvl == 0 && (s.sel & SEL_R2);
  if (rsel) L.b = CI->pcg_r2;
  const bool fax = pcg && (s.sel & SEL_AXPY) != 0;
  const bool axpy = AX && fax && s.lvl == 0;  // AX: the instantiation that carries the update
  const int even = rsel ? 1 : 0;
  const int iter = axpy ? (int)sld(scal + SC_ITER) + 1 : 0;
  // gamma and alpha of iteration k - 1 (slot (k - 2) & 1; this launch
  // writes slot (k - 1) & 1); unused in the first
  const double rz_prev = axpy ? sld(scal + SC_RZ0 + 1 - even) : 0.0;
  const double al_prev = axpy ? sld(scal + SC_AL0 + 1 - even) : 0.0;
  const float* xin = (s.sel & 1) ? L.xs : L.x;
  float* xout = (s.sel & 1) ? L.x : L.xs;
  const int ox = s.tx0 - HALO, oy = s.ty0 - HALO;
  T t;
  constexpr int XU = (MT_W * TH) / MNT;  // tile cells per thread, row-major over the tile
  float xv[XU];
  t.place(ox, oy, first, L.w, L.h);
  constexpr bool lazy = LZ;  // SEL_LAZY on every instance of the launch
  float* const p0 = CI->p0;
  // AX: z and p_prev over the region and its ring go to LDS (cells outside
  // the level: 0).  Their loads are issued before the tile's, and stored as
  // soon as they land while the tile's are still in flight, so that they do
  // not hold registers beside the tile's.  The first iteration reads p_prev
  // too (any value: it is not used), so the loads stay straight-line.
  float zv[PIT], pv[PIT];
  if (axpy) {
    const float* z = CI->pcg_z;
    const float* pold = even ? CI->pcg_p[1] : CI->pcg_p[0];  // p of iteration k - 1 in pcg_p[(k - 1) & 1]
#pragma unroll
    for (int k = 0; k < PIT; ++k) {
      const int u = (int)threadIdx.x + k * MNT;
      const int gx = ox - 1 + u % PW, gy = oy - 1 + u / PW;
      const uint32_t id = u < PN && mg_inb(gx, gy, L.w, L.h) ? (uint32_t)(gy * L.w + gx) : 0u;
      zv[k] = ld32(z, id);
      pv[k] = ld32(pold, id);
    }
  }
  t.load(L, xin, (s.sel & 4) != 0, nullptr, nullptr, ox, oy, lazy);
  if (axpy) {
#pragma unroll
    for (int k = 0; k < PIT; ++k) {
      const int u = (int)threadIdx.x + k * MNT;
      const int gx = ox - 1 + u % PW, gy = oy - 1 + u / PW;
      const bool inside = u < PN && mg_inb(gx, gy, L.w, L.h);
      if (u < PN) {
        sr[u] = inside ? zv[k] : 0.0f;
        spo[u] = inside ? pv[k] : 0.0f;
      }
    }
#pragma unroll
    for (int k = 0; k < XU; ++k) {  // x of the tile's cells (coalesced; outside the level: index 0, unused)
      const int u = (int)threadIdx.x + k * MNT, gx = s.tx0 + u % MT_W, gy = s.ty0 + u / MT_W;
      xv[k] = ld32(p0, gx < L.w && gy < L.h ? (uint32_t)(gy * L.w + gx) : 0u);
    }
  }
  if (lazy) t.load_cf(L, ox, oy);
  t.stage(nullptr, sx, lazy);
  if (pcg && done_flag != 0.0) return;
  if (axpy) {
    const int tile = s.blk - b.start[s.slot];
    double gd[2];  // gamma = r . z, delta = z . A z
    if (CI->prefold) {  // folded once by k_prefold (the same fold)
      gd[0] = sld(scal + SC_RZF);
      gd[1] = sld(scal + SC_RZF + 1);
    } else {
      fold_partials<2>(I, RED_RZ, CI->cg_tiles, false, gd);
    }
    const double beta = iter > 1 ? gd[0] / rz_prev : 0.0;
    const double den = iter > 1 ? gd[1] - beta * gd[0] / al_prev : gd[1];  // p . A p
    if (!(den > 0.0)) {  // breakdown (e.g. a non-finite state): stop before the update
      if (tile == 0 && threadIdx.x == 0) scal[SC_DONE] = 1.0;
      return;
    }
    const double alpha = gd[0] / den;
    const float fa = (float)alpha, fb = (float)beta;
    // p = z + beta p_prev (the first iteration: p = z), in place of z
#pragma unroll
    for (int k = 0; k < PIT; ++k) {
      const int u = (int)threadIdx.x + k * MNT;
      if (u < PN && iter > 1) sr[u] = sr[u] + fb * spo[u];
    }
    __syncthreads();
    const float factor = resid_factor(CI, 1.0f);
    float* rout = rsel ? CI->pcg_r : CI->pcg_r2;
    double acc = 0.0;
    // q = A p at every region cell in the level, two slots at a time
    // (neighbour loads first; more at once would cost the occupancy that
    // keeps all of a 1024^2 level's tiles resident)
    static_assert(T::NSL % 2 == 0, "slots in pairs");
#pragma unroll
    for (int h = 0; h < T::NSL / 2; ++h) {
      float nb[2][5];
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int k = h * 2 + j;
        const int q = t.inl[k] ? (t.ly[k] + 1) * PW + t.lx[k] + 1 : PW + 1;
        nb[j][0] = sr[q];
        nb[j][1] = sr[q - 1];
        nb[j][2] = sr[q + 1];
        nb[j][3] = sr[q - PW];
        nb[j][4] = sr[q + PW];
      }
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int k = h * 2 + j;
        if (!t.inl[k]) continue;
        const float q = 0.0f - res_upd(0.0f, t.c[k], t.fl[k], nb[j][0], nb[j][1], nb[j][2], nb[j][3], nb[j][4], factor);
        const float rv = t.b[k] - fa * q;
        t.b[k] = rv;
        if (!t.in_tile(k)) continue;
        stg(rout, (size_t)(oy + t.ly[k]) * L.w + ox + t.lx[k], rv);
        acc += (double)rv * (double)rv;
      }
    }
    float* pnew = even ? CI->pcg_p[0] : CI->pcg_p[1];
#pragma unroll
    for (int k = 0; k < XU; ++k) {
      const int u = (int)threadIdx.x + k * MNT, lx = u % MT_W, ly = u / MT_W, gx = s.tx0 + lx, gy = s.ty0 + ly;
      const float pc = sr[(ly + HALO + 1) * PW + lx + HALO + 1];
      if (gx < L.w && gy < L.h) {
        stg(p0, (size_t)gy * L.w + gx, xv[k] + fa * pc);
        stg(pnew, (size_t)gy * L.w + gx, pc);
      }
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
      double tot = wsum[0];
      for (int w = 1; w < MNT / 64; ++w) tot += wsum[w];
      CI->red[(size_t)RED_RR * CI->red_stride + tile] = tot;
      if (tile == 0) {
        scal[SC_RZ0 + even] = gd[0];
        scal[SC_AL0 + even] = alpha;
        scal[SC_BETA] = beta;
        scal[SC_PQ] = den;
        scal[SC_ALPHA] = alpha;
        scal[SC_KCUR] = (double)iter;
        scal[SC_NRR] = (double)(b.start[s.slot + 1] - b.start[s.slot]);
      }
    }
  }
  // the CG stopping test of the iteration that just ran (pre-folded
  // instances: k_cg_check)
  if (!CI->prefold && (fax ? s.lvl == 1 : cg_check != 0)) {
    double rr[1];  // one partial per 64 x 16 tile (k_pcg_axpy's, or this kernel's at level 0)
    fold_partials<1>(I, RED_RR, CI->cg_tiles, false, rr);
    if (cg_stop(I, rr[0], s.blk == b.start[s.slot] && threadIdx.x == 0)) return;
  }
  __syncthreads();
  const float scale = (float)(1 << s.lvl);
  t.template sweep<NS>(ox, oy, L.w, L.h, s.ex, s.ey, weight, gs_c0(CI, scale), sx);
  write_tile<HALO, TH>(L, xout, s.tx0, s.ty0, sx);
  if (!restrict_) return;
  const float factor = resid_factor(CI, scale);
  t.residual(factor, sx, sr);
  __syncthreads();
  // one coarse cell per thread: 32 x TH/2 coarse cells under the tile
  const Level Cc = lev_copy(CI->lev[s.lvl + 1]);
  const int ux = (int)threadIdx.x % 32, uy = (int)threadIdx.x / 32;
  const int cx = s.tx0 / 2 + ux, cy = s.ty0 / 2 + uy;
  if (uy >= TH / 2 || cx >= Cc.w || cy >= Cc.h) return;
  const int q00 = 2 * uy * MT_W + 2 * ux;
  // (0,0), (0,1), (1,0), (1,1) as restriction.wgsl sums them
  const float br = 0.25f * (((sr[q00] + sr[q00 + MT_W]) + sr[q00 + 1]) + sr[q00 + MT_W + 1]);
  stg(Cc.b, (size_t)cy * Cc.w + cx, br);
}

// The end of an up launch, after its sweeps: x of the tile to xout; with rz
// (level 0 of the CG preconditioner: 64 x 16 tiles, a halo one cell wider
// than the sweeps need) z masked to the pressure unknowns instead, and the
// block partials of r . z and z . A z.
template <int NS, int HALO, int TH>
__device__ __forceinline__ void up_finish(Tile<HALO, TH>& t, const Batch& b, const MSite& s, CDevInst* CI,
                                          const Level& L, float* xout, int rz, float* sx, double (*wsum)[MNT / 64]) {
  using T = Tile<HALO, TH>;
  constexpr bool zaz = HALO > 2 * NS;  // the r . z instantiation
  if (TH != MT_H || !zaz || !rz) {  // r . z partials come from 64 x 16 tiles only
    write_tile<HALO, TH>(L, xout, s.tx0, s.ty0, sx);
    return;
  }
  double* const scal = CI->scal;
  const int ox = s.tx0 - HALO, oy = s.ty0 - HALO;
  // z masked to the unknowns over the region, in registers and LDS
#pragma unroll
  for (int k = 0; k < T::NSL; ++k) {
    if (!(t.fl[k] & F_LSNEG) || (t.fl[k] & F_CLOSED)) t.x[k] = 0.0f;
    if (t.valid[k]) sx[t.ly[k] * T::RW + t.lx[k]] = t.x[k];
  }
  __syncthreads();
  const float factor = resid_factor(CI, 1.0f);
  float nb[T::NSL][4];
#pragma unroll
  for (int k = 0; k < T::NSL; ++k) {
    const int q = t.lds_at(k, t.in_tile(k));
    nb[k][0] = sx[q - 1];
    nb[k][1] = sx[q + 1];
    nb[k][2] = sx[q - T::RW];
    nb[k][3] = sx[q + T::RW];
  }
  double acc = 0.0, acd = 0.0;
#pragma unroll
  for (int k = 0; k < T::NSL; ++k) {
    if (!t.in_tile(k) || !t.inl[k]) continue;
    const float z = t.x[k];
    stg(xout, (size_t)(oy + t.ly[k]) * L.w + ox + t.lx[k], z);
    acc += (double)t.b[k] * (double)z;
    const float az = 0.0f - res_upd(0.0f, t.c[k], t.fl[k], z, nb[k][0], nb[k][1], nb[k][2], nb[k][3], factor);
    acd += (double)z * (double)az;
  }
  for (int off = 32; off > 0; off >>= 1) {
    acc += __shfl_xor(acc, off, 64);
    acd += __shfl_xor(acd, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    wsum[0][threadIdx.x >> 6] = acc;
    wsum[1][threadIdx.x >> 6] = acd;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double tot = wsum[0][0], tod = wsum[1][0];
    for (int w = 1; w < MNT / 64; ++w) {
      tot += wsum[0][w];
      tod += wsum[1][w];
    }
    const int tile = s.blk - b.start[s.slot];
    CI->red[(size_t)RED_RZ * CI->red_stride + tile] = tot;
    CI->red[(size_t)RED_ZAZ * CI->red_stride + tile] = tod;
    if (tile == 0) scal[SC_NRZ] = (double)(b.start[s.slot + 1] - b.start[s.slot]);
  }
}


// ---------------------------------------------------------------------------
// up: prolongation while staging, NS post-smoothing sweeps.  With rz (level 0
// of the CG preconditioner) z is masked to the pressure unknowns and the block
// partials of r . z and z . A z go to red[RED_RZ][tile] and red[RED_ZAZ][tile]
// (z . A z: the halo is one cell wider than the sweeps need, so z is exact on
// the tile and a one-cell ring; the AX launch of the next CG iteration takes
// its step length from both sums).
// ---------------------------------------------------------------------------
template <int NS, int HALO, int TH = MT_H, bool LZ = false>
__global__ __launch_bounds__(MNT) void k_mg_up(Batch b, int pcg, int first, float weight, int rz) {
  using T = Tile<HALO, TH>;
  __shared__ float sx[T::RW * T::RH];
  __shared__ double wsum[2][MNT / 64];
  const MSite s = mg_locate<TH>(b);
  CDevInst* CI = cdev(s.I);
  double* const scal = CI->scal;
  const double done_flag = sld(scal + SC_DONE);  // tested once the tile loads are in flight
  Level L = mg_level(CI, s.lvl, pcg);
  if (pcg && s.lvl == 0 && (s.sel & SEL_R2)) L.b = CI->pcg_r2;
  const Level C = lev_copy(CI->lev[s.lvl + 1]);
  const float* xin = (s.sel & 1) ? L.xs : L.x;
  float* xout = (s.sel & 1) ? L.x : L.xs;
  const int ox = s.tx0 - HALO, oy = s.ty0 - HALO;
  T t;
  t.place(ox, oy, first, L.w, L.h);
  constexpr bool lazy = LZ;  // SEL_LAZY on every instance of the launch
  t.load(L, xin, (s.sel & 4) != 0, &C, (s.sel & 2) ? C.xs : C.x, ox, oy, lazy);
  if (lazy) t.load_cf(L, ox, oy);
  t.stage(&C, sx, lazy);
  if (pcg && done_flag != 0.0) return;
  __syncthreads();
  const float scale = (float)(1 << s.lvl);
  t.template sweep<NS>(ox, oy, L.w, L.h, s.ex, s.ey, weight, gs_c0(CI, scale), sx);
  up_finish<NS, HALO, TH>(t, b, s, CI, L, xout, rz, sx, wsum);
}

// Up of two levels in one launch: level lvl + 1's prolongation (from
// lvl + 2) and two post-smoothing sweeps over the parents of this tile's
// region and a 4-cell halo, in LDS and never stored (only the prolongation
// to lvl reads them), then lvl's work as k_mg_up does it.  Every parent lies
// 4 or more cells inside the coarse region, so its value is the one k_mg_up
// on level lvl + 1 would have stored, and the results are bit for bit those
// of the two separate launches, one kernel boundary fewer.  Batch::sel bits
// 8..10 hold level lvl + 1's buffer state as bits 0..2 hold lvl's.
template <int NS, int HALO, int TH = MT_H, bool LZ = false>
__global__ __launch_bounds__(MNT) void k_mg_up2(Batch b, int pcg, int first, float weight, int rz) {
  using T = Tile<HALO, TH>;
  constexpr int HC = 4;         // two coarse sweeps
  constexpr int OD = HALO & 1;  // tiles start at even cells: an odd halo starts the region at an odd cell
  constexpr int TWC = (T::RW + OD) / 2 + OD, THC = (T::RH + OD) / 2 + OD;  // the parents of the region
  using TC = Tile<HC, THC, TWC>;
  __shared__ float sx[T::RW * T::RH];
  __shared__ float sxc[TC::RW * TC::RH];
  __shared__ double wsum[2][MNT / 64];
  const MSite s = mg_locate<TH>(b);
  CDevInst* CI = cdev(s.I);
  const double done_flag = sld(CI->scal + SC_DONE);  // tested once the tile loads are in flight
  Level L = mg_level(CI, s.lvl, pcg);
  if (pcg && s.lvl == 0 && (s.sel & SEL_R2)) L.b = CI->pcg_r2;
  const Level C = lev_copy(CI->lev[s.lvl + 1]);
  const Level C2 = lev_copy(CI->lev[s.lvl + 2]);
  const int selc = s.sel >> 8;
  const float* xin = (s.sel & 1) ? L.xs : L.x;
  float* xout = (s.sel & 1) ? L.x : L.xs;
  const int ox = s.tx0 - HALO, oy = s.ty0 - HALO;
  const int cox = (ox >> 1) - HC, coy = (oy >> 1) - HC;  // LDS cell (0, 0) of the coarse region
  TC tc;
  tc.place(cox, coy, first, C.w, C.h);
  tc.load(C, (selc & 1) ? C.xs : C.x, (selc & 4) != 0, &C2, (selc & 2) ? C2.xs : C2.x, cox, coy);
  T t;
  t.place(ox, oy, first, L.w, L.h);
  constexpr bool lazy = LZ;  // SEL_LAZY on every instance of the launch (fine level)
  t.load(L, xin, (s.sel & 4) != 0, nullptr, nullptr, ox, oy, true);  // the coefficients below
  tc.stage(&C2, sxc);
  if (pcg && done_flag != 0.0) return;
  __syncthreads();
  tc.template sweep<2>(cox, coy, C.w, C.h, C.w, C.h, weight, gs_c0(CI, (float)(2 << s.lvl)), sxc);
  // the fine coefficients once the coarse tile's registers are free
  if (lazy) t.load_cf(L, ox, oy);
  else t.load_c(L, ox, oy);
  // the parents' values (the sweep ended with a barrier), then the staging
  // with the prolongation as k_mg_up does it
#pragma unroll
  for (int k = 0; k < T::NSL; ++k) {
    const int cx = (ox + t.lx[k]) >> 1, cy = (oy + t.ly[k]) >> 1;
    t.par[k] = t.inl[k] && cx < C.w && cy < C.h;
    t.xc[k] = sxc[t.valid[k] ? (cy - coy) * TC::RW + cx - cox : 0];
  }
  t.stage(&C, sx, lazy);
  __syncthreads();
  t.template sweep<NS>(ox, oy, L.w, L.h, s.ex, s.ey, weight, gs_c0(CI, (float)(1 << s.lvl)), sx);
  up_finish<NS, HALO, TH>(t, b, s, CI, L, xout, rz, sx, wsum);
}

// ---------------------------------------------------------------------------
// Bottom of the V-cycle in one workgroup of 1024 threads: levels l0 .. nlev-1
// staged in LDS (stencil, x, b); the level table lives in LDS too.  Levels of
// more than BOT_SMALL cells are swept by the whole workgroup (a barrier after
// every half-sweep); the smaller ones by wavefront 0 alone, which needs only
// an LDS fence between half-sweeps, while the other wavefronts wait at one
// barrier.  One workgroup owns the whole level, so x is read from and
// written back to the buffer Batch::sel names.
// ---------------------------------------------------------------------------
constexpr int BOT_NT = 1024;
constexpr int BOT_SMALL = 256;

struct BotLds {
  float4* c;
  float *x, *b;
  int* fl;
};

// G threads take part: a whole workgroup (G > 64: a workgroup barrier) or
// wavefront 0 alone (G = 64: an LDS fence and a wavefront barrier)
template <int G>
__device__ __forceinline__ void b_sync() {
  if (G > 64) {
    __syncthreads();
  } else {  // one wavefront: its LDS accesses complete before any later one
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
  }
}

// G threads walk a level as a TX x TY grid (no index divisions).  Within a
// colour the update order does not matter, so every grid gives the same bits.
template <int G, int TX_ = (G >= 1024 ? 32 : (G >= 256 ? 16 : 8))>
struct BGrid {
  static constexpr int TX = TX_, TY = G / TX;
  static __device__ __forceinline__ int tx() { return (int)threadIdx.x % TX; }
  static __device__ __forceinline__ int ty() { return (int)threadIdx.x / TX; }
};

// Run f<TX>() with the narrowest row width TX in {32, 16, 8} (at most G) that
// covers n columns, so that the active threads fill whole wavefronts and the
// others skip the loop: the VALU time of a phase follows the cells, not G.
template <int G, typename F>
__device__ __forceinline__ void b_cols(int n, F&& f) {
  if (G >= 256 && n > 16) f(BGrid<G, 32>{});
  else if (G >= 128 && n > 8) f(BGrid<G, 16>{});
  else f(BGrid<G, 8>{});
}

template <int G>
__device__ __forceinline__ void b_sweep(int w, int h, int off, const BotLds& m, int parity, float c0) {
  const int hw = (w + 1) / 2;  // cells of one colour in a row, at most
  b_cols<G>(hw, [&](auto grid) {
    using T = decltype(grid);
    for (int gy = T::ty(); gy < h; gy += T::TY)
      for (int i = T::tx(); i < hw; i += T::TX) {
        const int gx = 2 * i + ((parity + gy) & 1);
        if (gx >= w) continue;
        const int q = off + gy * w + gx;
        const int j0 = gx > 0 ? q - 1 : q, j1 = gx + 1 < w ? q + 1 : q;
        const int j2 = gy > 0 ? q - w : q, j3 = gy + 1 < h ? q + w : q;
        const float4 c = m.c[q];
        const int fl = m.fl[q];
        m.x[q] = gs_upd(m.b[q], c, b_denominator(c, fl), fl, m.x[q], m.x[j0], m.x[j1], m.x[j2], m.x[j3], 1.0f, c0);
      }
  });
  b_sync<G>();
}

// residual + restriction of b into the next level (each coarse cell forms
// the residuals of its four children); the coarse x starts at 0
template <int G>
__device__ __forceinline__ void b_restrict(int w, int h, int off, int cw, int ch, int coff, const BotLds& m,
                                           float factor) {
  b_cols<G>(cw, [&](auto grid) {
  using T = decltype(grid);
  for (int cy = T::ty(); cy < ch; cy += T::TY)
    for (int cx = T::tx(); cx < cw; cx += T::TX) {
      const int dxs[4] = {0, 0, 1, 1}, dys[4] = {0, 1, 0, 1};  // (0,0) (0,1) (1,0) (1,1)
      float r4[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int gx = 2 * cx + dxs[k], gy = 2 * cy + dys[k];
        const bool in = mg_inb(gx, gy, w, h);
        const int q = in ? off + gy * w + gx : off;
        const int j0 = gx > 0 ? q - 1 : q, j1 = gx + 1 < w ? q + 1 : q;
        const int j2 = gy > 0 ? q - w : q, j3 = gy + 1 < h ? q + w : q;
        const float r = res_upd(m.b[q], m.c[q], m.fl[q], m.x[q], m.x[j0], m.x[j1], m.x[j2], m.x[j3], factor);
        r4[k] = in ? r : 0.0f;
      }
      const int q = coff + cy * cw + cx;
      m.b[q] = 0.25f * (((r4[0] + r4[1]) + r4[2]) + r4[3]);
      m.x[q] = 0.0f;
    }
  });
  b_sync<G>();
}

template <int G>
__device__ __forceinline__ void b_prolong(int w, int h, int off, int cw, int ch, int coff, const BotLds& m) {
  b_cols<G>(w, [&](auto grid) {
    using T = decltype(grid);
    for (int gy = T::ty(); gy < h; gy += T::TY)
      for (int gx = T::tx(); gx < w; gx += T::TX) {
        const int cx = gx >> 1, cy = gy >> 1;
        const int q = off + gy * w + gx;
        if (cx < cw && cy < ch && (m.fl[q] & F_LSNEG)) m.x[q] = m.x[coff + cy * cw + cx] + m.x[q];
      }
  });
  b_sync<G>();
}

struct BotCfg {
  const DevInst* I;
  const int *lw, *lh, *loff;
  int nl, pre, post, coarse, symmetric;
};

// the coarsest level: coarse red-black sweeps (and as many black-red ones
// when the cycle is symmetric)
template <int G>
__device__ __forceinline__ void b_coarsest(const BotCfg& g, const BotLds& m) {
  const int l = g.nl - 1;
  const float c0 = gs_c0(g.I, (float)(1 << l));
  for (int k = 0; k < g.coarse; ++k) {
    b_sweep<G>(g.lw[l], g.lh[l], g.loff[l], m, 1, c0);
    b_sweep<G>(g.lw[l], g.lh[l], g.loff[l], m, 0, c0);
  }
  if (g.symmetric)
    for (int k = 0; k < g.coarse; ++k) {
      b_sweep<G>(g.lw[l], g.lh[l], g.loff[l], m, 0, c0);
      b_sweep<G>(g.lw[l], g.lh[l], g.loff[l], m, 1, c0);
    }
}
// pre-smoothing of level l and restriction into l + 1
template <int G>
__device__ __forceinline__ void b_down(const BotCfg& g, const BotLds& m, int l) {
  const float scale = (float)(1 << l);
  const float c0 = gs_c0(g.I, scale);
  for (int k = 0; k < g.pre; ++k) {
    b_sweep<G>(g.lw[l], g.lh[l], g.loff[l], m, 1, c0);
    b_sweep<G>(g.lw[l], g.lh[l], g.loff[l], m, 0, c0);
  }
  b_restrict<G>(g.lw[l], g.lh[l], g.loff[l], g.lw[l + 1], g.lh[l + 1], g.loff[l + 1], m, resid_factor(g.I, scale));
}
// prolongation from l + 1 into l and post-smoothing of level l
template <int G>
__device__ __forceinline__ void b_up(const BotCfg& g, const BotLds& m, int l) {
  b_prolong<G>(g.lw[l], g.lh[l], g.loff[l], g.lw[l + 1], g.lh[l + 1], g.loff[l + 1], m);
  const float c0 = gs_c0(g.I, (float)(1 << l));
  for (int k = 0; k < g.post; ++k) {
    if (g.symmetric) {
      b_sweep<G>(g.lw[l], g.lh[l], g.loff[l], m, 0, c0);
      b_sweep<G>(g.lw[l], g.lh[l], g.loff[l], m, 1, c0);
    } else {
      b_sweep<G>(g.lw[l], g.lh[l], g.loff[l], m, 1, c0);
      b_sweep<G>(g.lw[l], g.lh[l], g.loff[l], m, 0, c0);
    }
  }
}

// The last two levels when they are 16 x 16 and 8 x 8 (square power-of-two
// grids), on wavefront 0 in registers: lane (lx, ly) = (lane & 7, lane >> 3)
// owns the 2 x 2 block (2 lx + dx, 2 ly + dy) of the 16 x 16 level, kept as
// k = 2 dy + dx, and cell (lx, ly) of the 8 x 8 level; the neighbours across
// a block edge come from the adjacent lanes by cross-lane shuffles.  The
// updates are the ones of b_down / b_coarsest / b_up with the same operands
// in the same order (a neighbour outside the level is the cell itself, as in
// b_sweep), so the results equal the LDS path's bit for bit; only the LDS
// round trip and the fence per half-sweep are gone.
__device__ __forceinline__ float b_shfl(float v, int src) { return __shfl(v, src & 63, 64); }

__device__ void b_tail16(const BotCfg& g, const BotLds& m, int ls) {
  const int lane = threadIdx.x & 63, lx = lane & 7, ly = lane >> 3;
  const int off = g.loff[ls], coff = g.loff[ls + 1];
  const bool hasL = lx > 0, hasR = lx < 7, hasU = ly > 0, hasD = ly < 7;
  float x[4], bb[4], dn[4];
  float4 c[4];
  int fl[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int q = off + (2 * ly + (k >> 1)) * 16 + 2 * lx + (k & 1);
    x[k] = m.x[q];
    bb[k] = m.b[q];
    c[k] = m.c[q];
    fl[k] = m.fl[q];
    dn[k] = b_denominator(c[k], fl[k]);
  }
  const int cq = coff + ly * 8 + lx;
  const float4 cc = m.c[cq];
  const int cfl = m.fl[cq];
  const float cdn = b_denominator(cc, cfl);
  const float c0f = gs_c0(g.I, (float)(1 << ls)), c0c = gs_c0(g.I, (float)(1 << (ls + 1)));
  // one colour of the 16 x 16 level: parity 0 = cells k 0, 3; parity 1 = k 1, 2
  auto fine = [&](int parity) {
    if (parity == 0) {
      const float l1 = b_shfl(x[1], lane - 1), u2 = b_shfl(x[2], lane - 8);
      const float r2 = b_shfl(x[2], lane + 1), d1 = b_shfl(x[1], lane + 8);
      const float n0 = gs_upd(bb[0], c[0], dn[0], fl[0], x[0], hasL ? l1 : x[0], x[1], hasU ? u2 : x[0], x[2], 1.0f, c0f);
      const float n3 = gs_upd(bb[3], c[3], dn[3], fl[3], x[3], x[2], hasR ? r2 : x[3], x[1], hasD ? d1 : x[3], 1.0f, c0f);
      x[0] = n0;
      x[3] = n3;
    } else {
      const float r0 = b_shfl(x[0], lane + 1), u3 = b_shfl(x[3], lane - 8);
      const float l3 = b_shfl(x[3], lane - 1), d0 = b_shfl(x[0], lane + 8);
      const float n1 = gs_upd(bb[1], c[1], dn[1], fl[1], x[1], x[0], hasR ? r0 : x[1], hasU ? u3 : x[1], x[3], 1.0f, c0f);
      const float n2 = gs_upd(bb[2], c[2], dn[2], fl[2], x[2], hasL ? l3 : x[2], x[3], x[0], hasD ? d0 : x[2], 1.0f, c0f);
      x[1] = n1;
      x[2] = n2;
    }
  };
  for (int k = 0; k < g.pre; ++k) {
    fine(1);
    fine(0);
  }
  // residuals of the four children, restricted in the order (0,0) (0,1) (1,0) (1,1)
  float cb, cx = 0.0f;
  {
    const float rf = resid_factor(g.I, (float)(1 << ls));
    const float l1 = b_shfl(x[1], lane - 1), l3 = b_shfl(x[3], lane - 1);
    const float r0 = b_shfl(x[0], lane + 1), r2 = b_shfl(x[2], lane + 1);
    const float u2 = b_shfl(x[2], lane - 8), u3 = b_shfl(x[3], lane - 8);
    const float d0 = b_shfl(x[0], lane + 8), d1 = b_shfl(x[1], lane + 8);
    const float ra = res_upd(bb[0], c[0], fl[0], x[0], hasL ? l1 : x[0], x[1], hasU ? u2 : x[0], x[2], rf);
    const float rb = res_upd(bb[2], c[2], fl[2], x[2], hasL ? l3 : x[2], x[3], x[0], hasD ? d0 : x[2], rf);
    const float rc = res_upd(bb[1], c[1], fl[1], x[1], x[0], hasR ? r0 : x[1], hasU ? u3 : x[1], x[3], rf);
    const float rd = res_upd(bb[3], c[3], fl[3], x[3], x[2], hasR ? r2 : x[3], x[1], hasD ? d1 : x[3], rf);
    cb = 0.25f * (((ra + rb) + rc) + rd);
  }
  // the coarsest level: lane (lx, ly) is cell (lx, ly), colour (lx + ly) & 1
  auto coarse = [&](int parity) {
    const float l = b_shfl(cx, lane - 1), r = b_shfl(cx, lane + 1), u = b_shfl(cx, lane - 8), d = b_shfl(cx, lane + 8);
    const float n = gs_upd(cb, cc, cdn, cfl, cx, hasL ? l : cx, hasR ? r : cx, hasU ? u : cx, hasD ? d : cx, 1.0f, c0c);
    if (((lx + ly) & 1) == parity) cx = n;
  };
  for (int k = 0; k < g.coarse; ++k) {
    coarse(1);
    coarse(0);
  }
  if (g.symmetric)
    for (int k = 0; k < g.coarse; ++k) {
      coarse(0);
      coarse(1);
    }
  // prolongation and post-smoothing
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (fl[k] & F_LSNEG) x[k] = cx + x[k];
  for (int k = 0; k < g.post; ++k) {
    if (g.symmetric) {
      fine(0);
      fine(1);
    } else {
      fine(1);
      fine(0);
    }
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) m.x[off + (2 * ly + (k >> 1)) * 16 + 2 * lx + (k & 1)] = x[k];
  m.x[cq] = cx;
  m.b[cq] = cb;
}

// Register blocks: the levels of the bottom solve that the whole workgroup
// sweeps (more than BOT_SMALL cells) with thread t owning the 2 x 2 block t
// (row-major over ceil(w / 2) x ceil(h / 2) blocks): cell k = 2 dy + dx at
// (2 bx + dx, 2 by + dy), its x, b and flags in registers (x also in LDS for
// the neighbours in other blocks), its stencil in LDS.  A block holds both
// colours, so a half-sweep updates two of its cells, reading the other two
// from registers and the rest from LDS; the four children of a coarse cell
// are one block, so the restriction needs no other thread's residuals.
// Every update, residual, restriction and prolongation is b_sweep's /
// b_restrict's / b_prolong's expression with the same operands in the same
// order (a neighbour outside the level is the cell itself), so the results
// equal the LDS loops' bit for bit.
// RC: the block's four stencils stay in registers too (the first, largest
// level: its half-sweeps then read only the neighbours' x from LDS, not 32
// KB of stencils per half-sweep); otherwise they are read from LDS.
template <bool RC>
struct RbLevel {
  static constexpr bool kRegC = RC;
  float x[4], b[4], dn[4];  // dn: the Gauss-Seidel denominator of the cell's stencil
  float4 c[RC ? 4 : 1];
  int fl[4];
  bool in[4];
  bool own;
  int bx, by, w, h, off;  // block, level size, offset of the level in the LDS arrays
};
template <class LV>
__device__ __forceinline__ float4 rb_stencil(const LV& L, const BotLds& m, int q, int k) {
  if constexpr (LV::kRegC) return L.c[k];
  else return m.c[q];
}

template <class LV>
__device__ __forceinline__ int rb_cell(const LV& L, int k) {
  return (2 * L.by + (k >> 1)) * L.w + 2 * L.bx + (k & 1);
}
// the block of this thread on a level, x, b and flags from the LDS copy
template <class LV>
__device__ __forceinline__ void rb_load(LV& L, const BotLds& m, int w, int h, int off) {
  const int bw = (w + 1) / 2, bh = (h + 1) / 2, t = (int)threadIdx.x;
  L.w = w;
  L.h = h;
  L.off = off;
  L.own = t < bw * bh;
  L.bx = L.own ? t % bw : 0;
  L.by = L.own ? t / bw : 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    L.in[k] = L.own && 2 * L.bx + (k & 1) < w && 2 * L.by + (k >> 1) < h;
    const int q = off + (L.in[k] ? rb_cell(L, k) : 0);
    L.x[k] = L.in[k] ? m.x[q] : 0.0f;
    L.b[k] = L.in[k] ? m.b[q] : 0.0f;
    L.fl[k] = L.in[k] ? m.fl[q] : 0;
    const float4 cq = m.c[q];
    if constexpr (LV::kRegC) L.c[k] = cq;
    L.dn[k] = b_denominator(cq, L.fl[k]);
  }
}
// the four neighbours (-x, +x, -y, +y) of cell k: those in the block from
// registers, the others from LDS (the cell itself when outside the level; a
// cell outside the level, whose results are dropped, reads around cell 0)
template <class LV>
__device__ __forceinline__ int rb_nbrs(const LV& L, const float* sx, int k, float (&n)[4]) {
  const int dx = k & 1, dy = k >> 1, gx = L.in[k] ? 2 * L.bx + dx : 0, gy = L.in[k] ? 2 * L.by + dy : 0;
  const int q = L.off + gy * L.w + gx;
  if (dx == 0) {
    n[0] = sx[gx > 0 ? q - 1 : q];
    n[1] = L.x[k + 1];
  } else {
    n[0] = L.x[k - 1];
    n[1] = sx[gx + 1 < L.w ? q + 1 : q];
  }
  if (dy == 0) {
    n[2] = sx[gy > 0 ? q - L.w : q];
    n[3] = L.x[k + 2];
  } else {
    n[2] = L.x[k - 2];
    n[3] = sx[gy + 1 < L.h ? q + L.w : q];
  }
  return q;
}
// one half-sweep: parity 0 = cells 0 and 3, parity 1 = cells 1 and 2 (a
// compile-time choice, so that the block's registers are indexed statically)
template <int G, int P, class LV>
__device__ __forceinline__ void rb_half(LV& L, const BotLds& m, float c0) {
  constexpr int k0 = P ? 1 : 0, k1 = P ? 2 : 3;
  if (L.own) {
    float n0[4], n1[4];
    const int q0 = rb_nbrs(L, m.x, k0, n0), q1 = rb_nbrs(L, m.x, k1, n1);
    const float4 c0v = rb_stencil(L, m, q0, k0), c1v = rb_stencil(L, m, q1, k1);
    const float v0 = gs_upd(L.b[k0], c0v, L.dn[k0], L.fl[k0], L.x[k0], n0[0], n0[1], n0[2], n0[3], 1.0f, c0);
    const float v1 = gs_upd(L.b[k1], c1v, L.dn[k1], L.fl[k1], L.x[k1], n1[0], n1[1], n1[2], n1[3], 1.0f, c0);
    if (L.in[k0]) {
      L.x[k0] = v0;
      m.x[q0] = v0;
    }
    if (L.in[k1]) {
      L.x[k1] = v1;
      m.x[q1] = v1;
    }
  }
  b_sync<G>();
}
// residuals of the block's cells restricted into the coarse level at coff
// (its b, and its x set to 0), as b_restrict
template <int G, class LV>
__device__ __forceinline__ void rb_restrict(const LV& L, const BotLds& m, int coff, int cw, int ch, float factor) {
  if (L.own && L.bx < cw && L.by < ch) {
    float r[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      float n[4];
      const int q = rb_nbrs(L, m.x, k, n);
      const float v = res_upd(L.b[k], rb_stencil(L, m, q, k), L.fl[k], L.x[k], n[0], n[1], n[2], n[3], factor);
      r[k] = L.in[k] ? v : 0.0f;
    }
    const int q = coff + L.by * cw + L.bx;
    m.b[q] = 0.25f * (((r[0] + r[2]) + r[1]) + r[3]);  // children (0,0) (0,1) (1,0) (1,1)
    m.x[q] = 0.0f;
  }
  b_sync<G>();
}
// the coarse x added to the block's fluid cells that have a parent, as b_prolong
template <int G, class LV>
__device__ __forceinline__ void rb_prolong(LV& L, const BotLds& m, int coff, int cw, int ch) {
  if (L.own && L.bx < cw && L.by < ch) {
    const float cx = m.x[coff + L.by * cw + L.bx];
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (L.in[k] && (L.fl[k] & F_LSNEG)) {
        L.x[k] = cx + L.x[k];
        m.x[L.off + rb_cell(L, k)] = L.x[k];
      }
  }
  b_sync<G>();
}
template <int G, class LV>
__device__ __forceinline__ void rb_pre(LV& L, const BotLds& m, int pre, float c0) {
  for (int k = 0; k < pre; ++k) {
    rb_half<G, 1>(L, m, c0);
    rb_half<G, 0>(L, m, c0);
  }
}
template <int G, class LV>
__device__ __forceinline__ void rb_post(LV& L, const BotLds& m, int post, int symmetric, float c0) {
  for (int k = 0; k < post; ++k) {
    if (symmetric) {
      rb_half<G, 0>(L, m, c0);
      rb_half<G, 1>(L, m, c0);
    } else {
      rb_half<G, 1>(L, m, c0);
      rb_half<G, 0>(L, m, c0);
    }
  }
}

// The bottom of the V-cycle from level l0 on, by the NT threads of one
// workgroup, in LDS (smem: bottom_lds_bytes of the levels' cells).  sel: bit
// 0 x of l0 is in Level::xs, bit 2 it is zero (not read); sweeps: pre, post,
// coarsest.  x of l0 goes back to the buffer it came from.  done (CG): the
// value of the done flag read at the kernel's start, tested once the staging
// loads are in flight (the two latencies overlap): if set, nothing happens.
template <int NT>
__device__ void bottom_run(const DevInst* I, int l0, int sel, int pcg, int symmetric, int pre, int post, int coarse,
                           float4* smem, double done = 0.0) {
  __shared__ int lw[kMaxLevels + 1], lh[kMaxLevels + 1], loff[kMaxLevels + 1];
  __shared__ const float4* lcf[kMaxLevels];
  __shared__ const int32_t* lfl[kMaxLevels];
  __shared__ int sh_ls;
  const int nl = I->nlev;
  // the level table: one thread per level loads its descriptor (the loads
  // overlap), then thread 0 forms the offsets
  if ((int)threadIdx.x < nl - l0) {
    const int l = l0 + (int)threadIdx.x;
    const Level& L = mg_level(I, l, pcg);
    lw[l] = L.w;
    lh[l] = L.h;
    lcf[l] = L.cf;
    lfl[l] = L.fl;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int total = 0;
    for (int l = l0; l < nl; ++l) {
      loff[l] = total;
      total += lw[l] * lh[l];
    }
    loff[nl] = total;
    int ls = l0;  // first level small enough for wavefront 0
    while (ls < nl && lw[ls] * lh[ls] > BOT_SMALL) ++ls;
    sh_ls = ls;
  }
  __syncthreads();
  const int total = loff[nl];
  BotLds m;
  m.c = smem;
  m.x = (float*)(m.c + total);
  m.b = m.x + total;
  m.fl = (int*)(m.b + total);
  const bool xzero = (sel & 4) != 0;
  // stencils of every level as one flat range (cell t of the range belongs to
  // the level whose offsets bracket it), four cells per thread per round so
  // that their loads are in flight together
  {
    const Level& L0 = mg_level(I, l0, pcg);
    const float* xsrc = (sel & 1) ? L0.xs : L0.x;
    const float* bsrc = L0.b;
    const int n0 = lw[l0] * lh[l0];
    constexpr int SU = 4;
    for (int t0 = threadIdx.x; t0 < total; t0 += SU * NT) {
      float4 c[SU];
      int f[SU], at[SU];
      float xv[SU], bv[SU];
#pragma unroll
      for (int u = 0; u < SU; ++u) {
        const int t = t0 + u * NT;
        at[u] = t;
        if (t >= total) continue;
        int l = l0;
        while (t >= loff[l + 1]) ++l;
        c[u] = ldg(lcf[l], t - loff[l]);
        f[u] = ldg(lfl[l], t - loff[l]);
        if (t < n0) {
          xv[u] = xzero ? 0.0f : ldg(xsrc, t);
          bv[u] = ldg(bsrc, t);
        }
      }
#pragma unroll
      for (int u = 0; u < SU; ++u) {
        const int t = at[u];
        if (t >= total) continue;
        m.c[t] = c[u];
        m.fl[t] = f[u];
        if (t < n0) {
          m.x[t] = xv[u];
          m.b[t] = bv[u];
        }
      }
    }
  }
  if (pcg && done != 0.0) return;  // uniform for the workgroup
  __syncthreads();
  BotCfg g;
  g.I = I;
  g.lw = lw;
  g.lh = lh;
  g.loff = loff;
  g.nl = nl;
  g.pre = pre;
  g.post = post;
  g.coarse = coarse;
  g.symmetric = symmetric;
  const int ls = sh_ls;  // first level small enough for wavefront 0
  // one or two levels for the whole workgroup, each of at most NT 2 x 2
  // blocks, then a tail for wavefront 0: those levels in register blocks
  bool rb = NT == 1024 && ls < nl && (ls - l0 == 1 || ls - l0 == 2);
  for (int l = l0; l < ls; ++l) rb = rb && ((lw[l] + 1) / 2) * ((lh[l] + 1) / 2) <= NT;
  RbLevel<true> A;  // the first level: stencils in registers
  RbLevel<false> B;
  if (rb) {
    rb_load(A, m, lw[l0], lh[l0], loff[l0]);
    rb_pre<NT>(A, m, pre, gs_c0(I, (float)(1 << l0)));
    rb_restrict<NT>(A, m, loff[l0 + 1], lw[l0 + 1], lh[l0 + 1], resid_factor(I, (float)(1 << l0)));
    if (ls - l0 == 2) {
      rb_load(B, m, lw[l0 + 1], lh[l0 + 1], loff[l0 + 1]);
      rb_pre<NT>(B, m, pre, gs_c0(I, (float)(1 << (l0 + 1))));
      rb_restrict<NT>(B, m, loff[ls], lw[ls], lh[ls], resid_factor(I, (float)(1 << (l0 + 1))));
    }
  } else {
    for (int l = l0; l < ls && l < nl - 1; ++l) b_down<NT>(g, m, l);
  }
  if (ls >= nl) {
    b_coarsest<NT>(g, m);  // the coarsest level is itself large
  } else {
    const bool tail16 = nl == ls + 2 && lw[ls] == 16 && lh[ls] == 16 && lw[ls + 1] == 8 && lh[ls + 1] == 8;
    if (threadIdx.x < 64) {
      if (tail16) {
        b_tail16(g, m, ls);
      } else {
        for (int l = ls; l < nl - 1; ++l) b_down<64>(g, m, l);
        b_coarsest<64>(g, m);
        for (int l = nl - 2; l >= ls; --l) b_up<64>(g, m, l);
      }
    }
    __syncthreads();
  }
  const Level L = mg_level(I, l0, pcg);
  float* xb = (sel & 1) ? L.xs : L.x;
  if (rb) {
    if (ls - l0 == 2) {
      rb_prolong<NT>(B, m, loff[ls], lw[ls], lh[ls]);
      rb_post<NT>(B, m, post, symmetric, gs_c0(I, (float)(1 << (l0 + 1))));
    }
    rb_prolong<NT>(A, m, loff[l0 + 1], lw[l0 + 1], lh[l0 + 1]);
    rb_post<NT>(A, m, post, symmetric, gs_c0(I, (float)(1 << l0)));
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (A.in[k]) stg(xb, rb_cell(A, k), A.x[k]);
    return;
  }
  for (int l = min(ls, nl - 1) - 1; l >= l0; --l) b_up<NT>(g, m, l);
  for (int t = threadIdx.x; t < L.w * L.h; t += NT) stg(xb, t, m.x[t]);
}

__global__ __launch_bounds__(BOT_NT) void k_bottom(Batch b, int pcg, int symmetric) {
  extern __shared__ float4 smem_f4[];
  const int slot = blockIdx.x;
  const DevInst* I = b.inst[slot];
  const double done = I->scal[SC_DONE];
  // sweep counts ride in the batch: ex = pre, ey = post | coarsest << 16
  bottom_run<BOT_NT>(I, b.level[slot], b.sel[slot], pcg, symmetric, b.ex[slot], b.ey[slot] & 0xffff, b.ey[slot] >> 16,
                     smem_f4, done);
}

// ---------------------------------------------------------------------------
// fused coarse V-cycle: the levels of at most kCoarseTiles tiles and the
// bottom in one launch (k_mg_coarse).  Each phase is the work of one
// k_mg_down / k_bottom / k_mg_up launch of the separate-launch cycle, with the
// same expressions in the same order (so the same bits); a barrier of the
// instance's workgroup group stands in for each kernel boundary.
// ---------------------------------------------------------------------------
// Barrier of the G workgroups of one group: arrivals in cnt[0], the
// generation in cnt[1].  The release fence before arriving and the acquire
// after leaving are agent-scope, so tiles written by any workgroup before it
// are read fresh after it.  A waiter that polls kBarrierSpin times without
// the generation moving records a failure in cnt[2] and leaves (false); a
// launch that finds cnt[2] set does nothing, so no wait is unbounded.
constexpr int kBarrierSpin = 1 << 14;
__device__ bool group_barrier(uint32_t* cnt, uint32_t G) {
  __shared__ int ok;
  __syncthreads();
  if (threadIdx.x == 0) {
    int good = 1;
    __threadfence();
    const uint32_t gen = __hip_atomic_load(cnt + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (atomicAdd(cnt, 1u) == G - 1) {
      __hip_atomic_store(cnt, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(cnt + 1, gen + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      for (int k = 0; __hip_atomic_load(cnt + 1, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == gen; ++k) {
        if (k >= kBarrierSpin) {
          __hip_atomic_store(cnt + 2, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          good = 0;
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
    }
    __threadfence();
    ok = good;
  }
  __syncthreads();
  return ok != 0;
}

// origin of tile `tile` of a level, or one far outside it when the level has
// fewer tiles (such a workgroup loads, updates and stores nothing)
__device__ __forceinline__ void coarse_tile(const Level& L, int tile, int& tx0, int& ty0) {
  const int tiles_x = (L.w + MT_W - 1) / MT_W, ntiles = tiles_x * ((L.h + MT_H - 1) / MT_H);
  if (tile < ntiles) {
    tx0 = (tile % tiles_x) * MT_W;
    ty0 = (tile / tiles_x) * MT_H;
  } else {
    tx0 = ty0 = 1 << 20;
  }
}

// k_mg_down's work on level lvl >= 1: NS sweeps, residual, restriction of b
template <int NS, int HALO>
__device__ void coarse_down(const DevInst* I, int lvl, int sel, int tile, float* sx, float* sr) {
  using T = Tile<HALO>;
  const Level L = I->lev[lvl];
  int tx0, ty0;
  coarse_tile(L, tile, tx0, ty0);
  const float* xin = (sel & 1) ? L.xs : L.x;
  float* xout = (sel & 1) ? L.x : L.xs;
  const int ox = tx0 - HALO, oy = ty0 - HALO;
  T t;
  t.place(ox, oy, 1, L.w, L.h);
  t.load(L, xin, (sel & 4) != 0, nullptr, nullptr, ox, oy);
  t.stage(nullptr, sx);
  __syncthreads();
  const float scale = (float)(1 << lvl);
  t.template sweep<NS>(ox, oy, L.w, L.h, L.w, L.h, 1.0f, gs_c0(I, scale), sx);
  write_tile<HALO>(L, xout, tx0, ty0, sx);
  const float factor = resid_factor(I, scale);
  t.residual(factor, sx, sr);
  __syncthreads();
  const Level Cc = I->lev[lvl + 1];
  const int ux = (int)threadIdx.x % 32, uy = (int)threadIdx.x / 32;
  const int cx = tx0 / 2 + ux, cy = ty0 / 2 + uy;
  if (cx < Cc.w && cy < Cc.h) {
    const int q00 = 2 * uy * MT_W + 2 * ux;
    const float br = 0.25f * (((sr[q00] + sr[q00 + MT_W]) + sr[q00 + 1]) + sr[q00 + MT_W + 1]);
    stg(Cc.b, (size_t)cy * Cc.w + cx, br);
  }
}

// k_mg_up's work on level lvl >= 1: prolongation while staging, NS sweeps
template <int NS, int HALO>
__device__ void coarse_up(const DevInst* I, int lvl, int sel, int tile, int first, float* sx) {
  using T = Tile<HALO>;
  const Level L = I->lev[lvl];
  const Level C = I->lev[lvl + 1];
  int tx0, ty0;
  coarse_tile(L, tile, tx0, ty0);
  const float* xin = (sel & 1) ? L.xs : L.x;
  float* xout = (sel & 1) ? L.x : L.xs;
  const int ox = tx0 - HALO, oy = ty0 - HALO;
  T t;
  t.place(ox, oy, first, L.w, L.h);
  t.load(L, xin, (sel & 4) != 0, &C, (sel & 2) ? C.xs : C.x, ox, oy);
  t.stage(&C, sx);
  __syncthreads();
  t.template sweep<NS>(ox, oy, L.w, L.h, L.w, L.h, 1.0f, gs_c0(I, (float)(1 << lvl)), sx);
  write_tile<HALO>(L, xout, tx0, ty0, sx);
}

// One group per instance (workgroups start[i] .. start[i+1]-1).  With pcg
// the group returns at once when the solve has stopped, and the first down
// phase runs the CG stopping test when it carries it (SEL_AXPY: level 1 of
// an iteration whose update rode on level 0), as k_mg_down does; every
// workgroup evaluates it from the same partials and so leaves together.
__global__ __launch_bounds__(MNT) void k_mg_coarse(CoarseBatch b, int pcg, int symmetric) {
  extern __shared__ float4 smem_f4[];  // the bottom's LDS (used by workgroup 0 of a group)
  __shared__ float sx[Tile<5>::RW * Tile<5>::RH];
  __shared__ float sr[MT_W * MT_H];
  const int i = slot_of(b.start, b.n, (int)blockIdx.x);
  const int w = (int)blockIdx.x - b.start[i];
  const uint32_t G = (uint32_t)(b.start[i + 1] - b.start[i]);
  const DevInst* I = b.inst[i];
  uint32_t* cnt = I->counter;
  if (__hip_atomic_load(cnt + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) return;
  if (pcg && I->scal[SC_DONE] != 0.0) return;
  const int l0 = b.l0[i], lb = b.lb[i];
  const uint8_t* sel = b.sel[i];
  int ph = 0;
  for (int l = l0; l < lb; ++l, ++ph) {
    if (l == l0 && pcg && (sel[ph] & SEL_AXPY)) {  // the stopping test of the iteration that just ran
      double rr[1];
      fold_partials<1>(I, RED_RR, I->cg_tiles, false, rr);
      if (cg_stop(I, rr[0], w == 0 && threadIdx.x == 0)) return;
    }
    coarse_down<2, 5>(I, l, sel[ph], w, sx, sr);
    if (!group_barrier(cnt, G)) return;
  }
  if (w == 0) bottom_run<MNT>(I, lb, sel[ph], pcg, symmetric, 2, 2, b.coarse[i], smem_f4);
  ++ph;
  if (!group_barrier(cnt, G)) return;
  const int first = symmetric ? 0 : 1;  // post-smoothing black-red in the symmetric cycle
  for (int l = lb - 1; l >= l0; --l, ++ph) {
    coarse_up<2, 4>(I, l, sel[ph], w, first, sx);
    if (l > l0 && !group_barrier(cnt, G)) return;
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
#define MG_LAUNCH(kern, nb, st, ...) hipLaunchKernelGGL(kern, dim3(nb), dim3(MNT), 0, st, __VA_ARGS__)

void launch_mg_setup_coarse(hipStream_t st, const Batch& b) {
  if (b.n > 0) hipLaunchKernelGGL(k_mg_setup_coarse, dim3(b.n), dim3(SC_NT), 0, st, b);
}
int mg_setup_coarse_cells() { return 64 * 64; }
int mg_setup_coarse_lds_cells(int offset) { return (offset & 1) ? SC_B : SC_A; }
void launch_mg_setup(hipStream_t st, const Batch& b) {
  int nb = b.start[b.n];
  if (nb <= 0) return;
  hipLaunchKernelGGL(k_mg_setup, dim3(nb), dim3(SU_W * SU_H), 0, st, b);
}

// nsweeps 0..3 pre-smoothing sweeps (+ residual/restriction); 4 on 64 x 8 tiles
// with cg_check, the launch first runs the CG stopping test (every block)
// The slots of b whose sel has `bit` (with) or lacks it, their tile ranges
// renumbered from 0.
Batch batch_subset(const Batch& b, int bit, bool with) {
  Batch o;
  std::memset(&o, 0, sizeof(o));
  int start = 0;
  for (int n = 0; n < b.n; ++n) {
    if (((b.sel[n] & bit) != 0) != with) continue;
    const int k = o.n++;
    o.start[k] = start;
    o.tiles_x[k] = b.tiles_x[n];
    o.ex[k] = b.ex[n];
    o.ey[k] = b.ey[n];
    o.level[k] = b.level[n];
    o.sel[k] = b.sel[n];
    o.inst[k] = b.inst[n];
    start += b.start[n + 1] - b.start[n];
  }
  o.start[o.n] = start;
  o.nblocks = start;
  return o;
}
// A 64 x 16 launch that mixes levels with lazy stencil loads (SEL_LAZY,
// 2^20 cells or more) and levels without runs the plain instantiation for
// all of them, so the large levels read every stencil.  When the lazy ones
// hold at least a quarter of the tiles (a 4096^2 instance batched with
// 512^2 ones) the launch is worth splitting in two.
bool split_lazy(const Batch& b) {
  int lazy = 0;
  for (int n = 0; n < b.n; ++n)
    if (b.sel[n] & SEL_LAZY) lazy += b.start[n + 1] - b.start[n];
  return lazy > 0 && lazy < b.start[b.n] && 4 * lazy >= b.start[b.n];
}

void launch_mg_down(hipStream_t st, const Batch& b, int pcg, int nsweeps, int first, float weight, int restrict_,
                    int cg_check, int th) {
  bool axpy = false;  // the fused CG update needs the AX instantiation (2 sweeps only)
  for (int n = 0; n < b.n; ++n) axpy = axpy || ((b.sel[n] & SEL_AXPY) && b.level[n] == 0);
  if (axpy && (nsweeps != 2 || th != MT_H)) {  // the host marks only 2-sweep 64 x 16 launches; anything else is a bug
    set_launch_error("launch_mg_down: fused CG update on a launch without 2 sweeps on 64 x 16 tiles");
    return;
  }
  if (th != MT_H && th != MT_H8) {
    set_launch_error("launch_mg_down: tile height neither 16 nor 8");
    return;
  }
  int nb = b.start[b.n];
  if (nb <= 0) return;
  if (th == MT_H && split_lazy(b)) {
    launch_mg_down(st, batch_subset(b, SEL_LAZY, true), pcg, nsweeps, first, weight, restrict_, cg_check, th);
    launch_mg_down(st, batch_subset(b, SEL_LAZY, false), pcg, nsweeps, first, weight, restrict_, cg_check, th);
    return;
  }
  bool lz = true;  // lazy stencil loads when every level of the launch streams from HBM
  for (int n = 0; n < b.n; ++n) lz = lz && (b.sel[n] & SEL_LAZY);
  if (th == MT_H8) {
    switch (nsweeps) {
      case 0: MG_LAUNCH((k_mg_down<0, 1, false, MT_H8>), nb, st, b, pcg, first, weight, restrict_, cg_check); break;
      case 1: MG_LAUNCH((k_mg_down<1, 3, false, MT_H8>), nb, st, b, pcg, first, weight, restrict_, cg_check); break;
      case 2: MG_LAUNCH((k_mg_down<2, 5, false, MT_H8>), nb, st, b, pcg, first, weight, restrict_, cg_check); break;
      case 3: MG_LAUNCH((k_mg_down<3, 7, false, MT_H8>), nb, st, b, pcg, first, weight, restrict_, cg_check); break;
      default: MG_LAUNCH((k_mg_down<4, 9, false, MT_H8>), nb, st, b, pcg, first, weight, restrict_, cg_check); break;
    }
    return;
  }
#define MG_DOWN16(NS, HALO, AX)                                                                            \
  do {                                                                                                   \
    if (lz) MG_LAUNCH((k_mg_down<NS, HALO, AX, MT_H, true>), nb, st, b, pcg, first, weight, restrict_, cg_check); \
    else MG_LAUNCH((k_mg_down<NS, HALO, AX, MT_H, false>), nb, st, b, pcg, first, weight, restrict_, cg_check);  \
  } while (0)
  switch (nsweeps) {
    case 0: MG_DOWN16(0, 1, false); break;
    case 1: MG_DOWN16(1, 3, false); break;
    case 2:
      if (axpy) MG_DOWN16(2, 5, true);
      else MG_DOWN16(2, 5, false);
      break;
    default: MG_DOWN16(3, 7, false); break;
  }
#undef MG_DOWN16
}

// prolongation + nsweeps 0..3 post-smoothing sweeps (+ r . z partials, 64 x 16 tiles only)
void launch_mg_up(hipStream_t st, const Batch& b, int pcg, int nsweeps, int first, float weight, int rz, int th) {
  if ((th != MT_H && th != MT_H8) || (rz && th != MT_H)) {
    set_launch_error("launch_mg_up: tile height neither 16 nor 8, or r . z partials on 64 x 8 tiles");
    return;
  }
  int nb = b.start[b.n];
  if (nb <= 0) return;
  if (th == MT_H8) {
    switch (nsweeps) {
      case 0: MG_LAUNCH((k_mg_up<0, 0, MT_H8>), nb, st, b, pcg, first, weight, rz); break;
      case 1: MG_LAUNCH((k_mg_up<1, 2, MT_H8>), nb, st, b, pcg, first, weight, rz); break;
      case 2: MG_LAUNCH((k_mg_up<2, 4, MT_H8>), nb, st, b, pcg, first, weight, rz); break;
      default: MG_LAUNCH((k_mg_up<3, 6, MT_H8>), nb, st, b, pcg, first, weight, rz); break;
    }
    return;
  }
  if (split_lazy(b)) {
    launch_mg_up(st, batch_subset(b, SEL_LAZY, true), pcg, nsweeps, first, weight, rz, th);
    launch_mg_up(st, batch_subset(b, SEL_LAZY, false), pcg, nsweeps, first, weight, rz, th);
    return;
  }
  bool lz = true;  // lazy stencil loads when every level of the launch streams from HBM
  for (int n = 0; n < b.n; ++n) lz = lz && (b.sel[n] & SEL_LAZY);
#define MG_UP16(NS, HALO)                                                                 \
  do {                                                                                    \
    if (lz) MG_LAUNCH((k_mg_up<NS, HALO, MT_H, true>), nb, st, b, pcg, first, weight, rz);  \
    else MG_LAUNCH((k_mg_up<NS, HALO, MT_H, false>), nb, st, b, pcg, first, weight, rz);    \
  } while (0)
  // with rz a halo one cell wider than the sweeps need (z . A z on the tile)
  switch (nsweeps) {
    case 0:
      if (rz) MG_UP16(0, 1);
      else MG_UP16(0, 0);
      break;
    case 1:
      if (rz) MG_UP16(1, 3);
      else MG_UP16(1, 2);
      break;
    case 2:
      if (rz) MG_UP16(2, 5);
      else MG_UP16(2, 4);
      break;
    default:
      if (rz) MG_UP16(3, 7);
      else MG_UP16(3, 6);
      break;
  }
#undef MG_UP16
}

// levels lvl and lvl + 1 of an up pass in one launch (2 post-smoothing
// sweeps on each, + r . z and z . A z partials with rz: 64 x 16 tiles only)
void launch_mg_up2(hipStream_t st, const Batch& b, int pcg, int first, float weight, int rz, int th) {
  if ((th != MT_H && th != MT_H8) || (rz && th != MT_H)) {
    set_launch_error("launch_mg_up2: tile height neither 16 nor 8, or r . z partials on 64 x 8 tiles");
    return;
  }
  int nb = b.start[b.n];
  if (nb <= 0) return;
  if (th == MT_H && split_lazy(b)) {
    launch_mg_up2(st, batch_subset(b, SEL_LAZY, true), pcg, first, weight, rz, th);
    launch_mg_up2(st, batch_subset(b, SEL_LAZY, false), pcg, first, weight, rz, th);
    return;
  }
  bool lz = true;
  for (int n = 0; n < b.n; ++n) lz = lz && (b.sel[n] & SEL_LAZY);
  if (th == MT_H8) {
    MG_LAUNCH((k_mg_up2<2, 4, MT_H8>), nb, st, b, pcg, first, weight, rz);
  } else if (rz) {
    if (lz) MG_LAUNCH((k_mg_up2<2, 5, MT_H, true>), nb, st, b, pcg, first, weight, rz);
    else MG_LAUNCH((k_mg_up2<2, 5, MT_H, false>), nb, st, b, pcg, first, weight, rz);
  } else {
    if (lz) MG_LAUNCH((k_mg_up2<2, 4, MT_H, true>), nb, st, b, pcg, first, weight, rz);
    else MG_LAUNCH((k_mg_up2<2, 4, MT_H, false>), nb, st, b, pcg, first, weight, rz);
  }
}

void launch_mg_coarse(hipStream_t st, const CoarseBatch& b, int pcg, int symmetric, int max_cells) {
  const int nb = b.start[b.n];
  if (nb <= 0) return;
  hipLaunchKernelGGL(k_mg_coarse, dim3(nb), dim3(MNT), bottom_lds_bytes(max_cells), st, b, pcg, symmetric);
}

// per cell: stencil (16) + x, b (8) + flags (4)
size_t bottom_lds_bytes(int cells) { return (size_t)cells * 28; }

// The bottom takes every level of at most 4096 cells (a 64 x 64 level and
// the ones below it) where a workgroup may hold the ~150 KB of LDS that needs,
// else of at most 1024 cells (EFL_BOTTOM_CELLS=1024 asks for the smaller one).
// With the register-block levels, at 1024^2 the larger bottom replaces two
// tiled launches per V-cycle: 654.7 against 644.7 steps/s on MI355X.
int bottom_cells_for_device(int device) {
  const char* e = getenv("EFL_BOTTOM_CELLS");
  if (e && atoi(e) == 1024) return 1024;
  return bottom_cells_enable(device, 4096);
}

// Let the bottom solve take levels of up to `cells` cells on this device
// (4096: the 64 x 64 level too), raising k_bottom's dynamic LDS limit.
// Returns the size granted: `cells`, or 1024 if the device refuses.
int bottom_cells_enable(int device, int cells) {
  if (cells <= 1024) return 1024;
  int lds = 0;
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess) lds = 0;
  const size_t need = bottom_lds_bytes(cells * 4 / 3 + 64);
  if ((size_t)lds >= need + 1024 &&
      hipFuncSetAttribute(reinterpret_cast<const void*>(k_bottom), hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)need) == hipSuccess)
    return cells;
  (void)hipGetLastError();
  return 1024;
}

void launch_bottom(hipStream_t st, const Batch& b, int pcg, int symmetric, int max_cells) {
  if (b.n <= 0) return;
  hipLaunchKernelGGL(k_bottom, dim3(b.n), dim3(BOT_NT), bottom_lds_bytes(max_cells), st, b, pcg, symmetric);
}

int mg_tile_w() { return MT_W; }
int mg_tile_h() { return MT_H; }
int mg_setup_tile_w() { return SU_W; }
int mg_setup_tile_h() { return SU_H; }

}  // namespace efl
