// efl_mg.hip - fused multigrid for gfx950: the V-cycle of
// projection/multi_grid.rs:390-475 (and its symmetric variant used as the
// CG preconditioner) in two launches per level plus one for the bottom.
//
//   mg_down    a block stages a 64 x 16 tile of one level plus a halo of
//              x, b, the air level set and the area fractions in LDS, runs
//              the pre-smoothing half-sweeps there (each one shrinks the
//              region whose values are exact by one cell), writes the tile
//              and, with `restrict`, forms the residual (residual.wgsl) of
//              its cells and restricts it (restriction.wgsl) to the 32 x 8
//              coarse cells under the tile; r is never stored.
//   mg_up      prolongation.wgsl applied while staging, then the
//              post-smoothing half-sweeps; at level 0 of the CG
//              preconditioner it also masks z and forms the block partial
//              of r . z.
//   bottom     one workgroup runs the V-cycle below the first level of at
//              most 1024 cells in LDS (reference order, no launches).
//
// Red-black Gauss-Seidel inside a colour is order independent, so these
// kernels produce exactly the values of the one-kernel-per-half-sweep
// reference dataflow (same operation order per cell; -ffp-contract=off).
//
// A block's halo belongs to other blocks' tiles, so a tiled launch never
// writes the x it reads: x and Level::xs form a pair, each launch reads one
// (Batch::sel) and writes every in-level cell of its tiles to the other.
// Writing in place would let a block that starts late stage neighbour
// values that were already smoothed.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "efl_internal.h"
#include "efl_launch.h"

namespace efl {

constexpr int MT_W = 64;   // tile width (one wavefront per tile row)
constexpr int MT_H = 16;   // tile height
constexpr int MNT = 256;   // threads per block
constexpr int BOTTOM_CELLS = 1024;

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
__device__ __forceinline__ float4 ldg(const float4* p, size_t i) {
  const f4v v = ((const __attribute__((address_space(1))) f4v*)p)[i];
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void stg(float* p, size_t i, float v) {
  ((__attribute__((address_space(1))) float*)p)[i] = v;
}
__device__ __forceinline__ void stg(float4* p, size_t i, float4 v) {
  const f4v w = {v.x, v.y, v.z, v.w};
  ((__attribute__((address_space(1))) f4v*)p)[i] = w;
}

__device__ __forceinline__ const Level& mg_level(const DevInst* I, int l, int pcg) {
  return (pcg && l == 0) ? I->lev_pcg0 : I->lev[l];
}

struct MSite {
  const DevInst* I;
  int tx0, ty0;  // tile origin (cells)
  int lvl;
  int ex, ey;    // update extent
  int slot;
  int sel;       // buffer selectors (Batch::sel)
};

__device__ __forceinline__ MSite mg_locate(const Batch& b) {
  int blk = blockIdx.x;
  int i = 0;
  while (i + 1 < b.n && blk >= b.start[i + 1]) ++i;
  int lb = blk - b.start[i];
  MSite s;
  s.I = b.inst[i];
  s.tx0 = (lb % b.tiles_x[i]) * MT_W;
  s.ty0 = (lb / b.tiles_x[i]) * MT_H;
  s.lvl = b.level[i];
  s.ex = b.ex[i];
  s.ey = b.ey[i];
  s.slot = i;
  s.sel = b.sel[i];
  return s;
}

// LDS indices of the four neighbours (-x, +x, -y, +y) of LDS cell q whose
// level cell is (gx, gy); a neighbour outside the level maps to q itself so
// every read stays inside the staged region.
struct Nbr {
  bool in[4];
  int j[4];
};
__device__ __forceinline__ Nbr mg_nbr(int q, int gx, int gy, int w, int h, int rowlen) {
  Nbr n;
  n.in[0] = gx > 0;
  n.in[1] = gx + 1 < w;
  n.in[2] = gy > 0;
  n.in[3] = gy + 1 < h;
  n.j[0] = n.in[0] ? q - 1 : q;
  n.j[1] = n.in[1] ? q + 1 : q;
  n.j[2] = n.in[2] ? q - rowlen : q;
  n.j[3] = n.in[3] ? q + rowlen : q;
  return n;
}

// The GS update of gauss_seidel.wgsl:44-88 on LDS-staged data, written
// without branches: every LDS read of the cell is issued up front and the
// early exits and the fluid/air/outside cases of each neighbour become
// selects.  A skipped term leaves its sum untouched (no "+ 0"), so the
// arithmetic is exactly that of the branching form.
__device__ __forceinline__ float gs_lds(const float* sx, const float* sb, const float* sl, const float4* sa, int q,
                                        int gx, int gy, int w, int h, int rowlen, float weight, float c0) {
  const Nbr n = mg_nbr(q, gx, gy, w, h, rowlen);
  const float lc = sl[q], p_old = sx[q], bq = sb[q];
  const float4 f = sa[q];
  float lv[4], xv[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    lv[k] = sl[n.j[k]];
    xv[k] = sx[n.j[k]];
  }
  const float fv[4] = {f.x, f.y, f.z, f.w};
  float denom = 0.0f;
  float nume = c0 * bq;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const bool fl = lv[k] < 0.0f;
    const float d_fl = denom + fv[k];
    const float n_fl = nume + fv[k] * xv[k];
    const float d_air = denom + fv[k] * (1.0f - lv[k] / lc);
    denom = n.in[k] ? (fl ? d_fl : d_air) : denom;
    nume = (n.in[k] && fl) ? n_fl : nume;
  }
  const float pn = nume / denom;
  const float v = weight * pn + (1.0f - weight) * p_old;
  const bool zero = lc >= 0.0f || closed4(f) || fabsf(denom) < 1.0e-6f;
  return zero ? 0.0f : v;
}

// residual.wgsl:12-57 on LDS-staged data, branch-free like gs_lds
__device__ __forceinline__ float resid_lds(const float* sx, const float* sb, const float* sl, const float4* sa, int q,
                                           int gx, int gy, int w, int h, int rowlen, float factor) {
  const Nbr n = mg_nbr(q, gx, gy, w, h, rowlen);
  const float phi = sl[q], xc = sx[q];
  const float4 f = sa[q];
  float res = sb[q];
  float lv[4], xv[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    lv[k] = sl[n.j[k]];
    xv[k] = sx[n.j[k]];
  }
  const float fv[4] = {f.x, f.y, f.z, f.w};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const float r_fl = res - fv[k] * (xc - xv[k]) * factor;
    const float r_air = res - fv[k] * (1.0f - lv[k] / phi) * xc * factor;
    res = n.in[k] ? (lv[k] < 0.0f ? r_fl : r_air) : res;
  }
  const bool zero = phi >= 0.0f || closed4(f);
  return zero ? 0.0f : res;
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

// Stage [tile + halo] of a level into LDS.  All global loads of a thread are
// issued before its LDS stores (latency overlap).  Cells outside the level
// read 0.  x comes from xin; with C, prolongation.wgsl adds the coarse x in
// cxin on the fly.
template <int HALO>
__device__ __forceinline__ void stage_tile(const Level& L, const float* xin, const Level* C, const float* cxin, int x0,
                                           int y0, float* sx, float* sb, float* sl, float4* sa) {
  constexpr int RW = MT_W + 2 * HALO, RH = MT_H + 2 * HALO, RN = RW * RH;
  constexpr int NIT = (RN + MNT - 1) / MNT;
  float rx[NIT], rb[NIT], rl[NIT], rc[NIT];
  float4 ra[NIT];
  bool par[NIT];
#pragma unroll
  for (int k = 0; k < NIT; ++k) {
    int q = threadIdx.x + k * MNT;
    int gx = x0 - HALO + q % RW, gy = y0 - HALO + q / RW;
    rx[k] = rb[k] = rl[k] = rc[k] = 0.0f;
    ra[k] = make_float4(0.f, 0.f, 0.f, 0.f);
    par[k] = false;
    if (q < RN && mg_inb(gx, gy, L.w, L.h)) {
      size_t id = (size_t)gy * L.w + gx;
      rx[k] = ldg(xin, id);
      rb[k] = ldg(L.b, id);
      rl[k] = ldg(L.ls, id);
      ra[k] = ldg(L.af, id);
      // the coarse value is fetched with the fine ones and used below only
      // on fluid cells
      int cx = gx >> 1, cy = gy >> 1;
      if (C && cx < C->w && cy < C->h) {
        par[k] = true;
        rc[k] = ldg(cxin, (size_t)cy * C->w + cx);
      }
    }
  }
  if (C) {  // prolongation: x_f += x_c on fluid fine cells with a parent
#pragma unroll
    for (int k = 0; k < NIT; ++k)
      if (par[k] && rl[k] < 0.0f) rx[k] = rc[k] + rx[k];
  }
#pragma unroll
  for (int k = 0; k < NIT; ++k) {
    int q = threadIdx.x + k * MNT;
    if (q < RN) {
      sx[q] = rx[k];
      sb[q] = rb[k];
      sl[q] = rl[k];
      sa[q] = ra[k];
    }
  }
}

// 2 * NS half-sweeps on the staged region (first colour of each sweep:
// 1 red, 0 black).  Threads walk the cells of one colour directly; a thread
// computes all of its cells of a half-sweep before it stores any (cells of
// one colour read only the other colour), so their LDS reads overlap.  Cells
// outside [0, ex) x [0, ey) are not updated (dispatch coverage); half-sweep
// hs updates only cells at least hs + 1 from the region border, whose
// neighbours are still exact.
template <int HALO, int NS>
__device__ __forceinline__ void sweep_tile(int w, int h, int x0, int y0, int ex, int ey, int first, float weight,
                                           float c0, float* sx, const float* sb, const float* sl, const float4* sa) {
  constexpr int RW = MT_W + 2 * HALO, RH = MT_H + 2 * HALO, HW = RW / 2;
  constexpr int NIT = (HW * RH + MNT - 1) / MNT;
  const int ox = x0 - HALO, oy = y0 - HALO;  // level coordinates of LDS (0, 0)
  const int base = (ox + oy) & 1;
  // local bounds from the level and the dispatch extent
  const int bx0 = -ox, by0 = -oy, bx1 = min(ex, w) - ox, by1 = min(ey, h) - oy;
  int ly[NIT], lx2[NIT];
#pragma unroll
  for (int k = 0; k < NIT; ++k) {
    const int t = threadIdx.x + k * MNT;
    ly[k] = t / HW;  // >= RH for t past the colour's cells: never active
    lx2[k] = 2 * (t - ly[k] * HW);
  }
#pragma unroll
  for (int hs = 0; hs < 2 * NS; ++hs) {
    const int parity = (hs & 1) == 0 ? first : 1 - first;
    const int m = hs + 1;
    const int lo_x = max(m, bx0), hi_x = min(RW - m, bx1), lo_y = max(m, by0), hi_y = min(RH - m, by1);
    float nv[NIT];
    int qq[NIT];
    bool act[NIT];
#pragma unroll
    for (int k = 0; k < NIT; ++k) {
      const int lx = lx2[k] + ((parity + ly[k] + base) & 1);
      act[k] = lx >= lo_x && lx < hi_x && ly[k] >= lo_y && ly[k] < hi_y;
      // an inactive slot evaluates an interior dummy cell and discards it
      qq[k] = act[k] ? ly[k] * RW + lx : RW + 1;
      nv[k] = gs_lds(sx, sb, sl, sa, qq[k], ox + lx, oy + ly[k], w, h, RW, weight, c0);
    }
#pragma unroll
    for (int k = 0; k < NIT; ++k)
      if (act[k]) sx[qq[k]] = nv[k];
    __syncthreads();
  }
}

// Every in-level cell of the tile goes to xout, updated or not, so xout is a
// complete copy of x after the launch.
template <int HALO>
__device__ __forceinline__ void write_tile(const Level& L, float* xout, int x0, int y0, const float* sx) {
  constexpr int RW = MT_W + 2 * HALO;
#pragma unroll
  for (int k = 0; k < (MT_W * MT_H) / MNT; ++k) {
    int t = threadIdx.x + k * MNT;
    int lx = t % MT_W, ly = t / MT_W;
    int gx = x0 + lx, gy = y0 + ly;
    if (gx < L.w && gy < L.h) stg(xout, (size_t)gy * L.w + gx, sx[(ly + HALO) * RW + lx + HALO]);
  }
}

// ---------------------------------------------------------------------------
// down: NS pre-smoothing sweeps (+ residual + restriction).  The halo is one
// cell wider than the sweeps need so that the residual of every tile cell
// sees exact neighbours.
// ---------------------------------------------------------------------------
template <int NS, int HALO>
__global__ __launch_bounds__(MNT) void k_mg_down(Batch b, int pcg, int first, float weight, int restrict_) {
  constexpr int RW = MT_W + 2 * HALO, RH = MT_H + 2 * HALO, RN = RW * RH;
  __shared__ float sx[RN], sb[RN], sl[RN];
  __shared__ float4 sa[RN];
  MSite s = mg_locate(b);
  const DevInst* I = s.I;
  const Level L = mg_level(I, s.lvl, pcg);
  float* xin = (s.sel & 1) ? L.xs : L.x;
  float* xout = (s.sel & 1) ? L.x : L.xs;
  stage_tile<HALO>(L, xin, nullptr, nullptr, s.tx0, s.ty0, sx, sb, sl, sa);
  __syncthreads();
  const float scale = (float)(1 << s.lvl);
  sweep_tile<HALO, NS>(L.w, L.h, s.tx0, s.ty0, s.ex, s.ey, first, weight, gs_c0(I, scale), sx, sb, sl, sa);
  write_tile<HALO>(L, xout, s.tx0, s.ty0, sx);
  if (!restrict_) return;
  // one coarse cell per thread: 32 x 8 coarse cells under the 64 x 16 tile
  const Level Cc = I->lev[s.lvl + 1];
  const int cx = s.tx0 / 2 + (int)(threadIdx.x % 32), cy = s.ty0 / 2 + (int)(threadIdx.x / 32);
  if (cx >= Cc.w || cy >= Cc.h) return;
  const float factor = resid_factor(I, scale);
  const int lx = 2 * (cx - s.tx0 / 2) + HALO, ly = 2 * (cy - s.ty0 / 2) + HALO;
  float r4[4];  // (0,0), (0,1), (1,0), (1,1) as restriction.wgsl sums them
  const int dxs[4] = {0, 0, 1, 1}, dys[4] = {0, 1, 0, 1};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    int gx = 2 * cx + dxs[k], gy = 2 * cy + dys[k];
    int q = (ly + dys[k]) * RW + lx + dxs[k];
    r4[k] = mg_inb(gx, gy, L.w, L.h) ? resid_lds(sx, sb, sl, sa, q, gx, gy, L.w, L.h, RW, factor) : 0.0f;
  }
  float br = 0.25f * (((r4[0] + r4[1]) + r4[2]) + r4[3]);
  const int q00 = ly * RW + lx, q01 = q00 + RW, q10 = q00 + 1, q11 = q00 + RW + 1;
  float p0 = sl[q00], p1 = sl[q01], p2 = sl[q10], p3 = sl[q11];
  float pr = p0;
  if (fabsf(p1) < fabsf(pr)) pr = p1;
  if (fabsf(p2) < fabsf(pr)) pr = p2;
  if (fabsf(p3) < fabsf(pr)) pr = p3;
  float4 a00 = sa[q00], a10 = sa[q10], a01 = sa[q01], a11 = sa[q11];
  size_t id = (size_t)cy * Cc.w + cx;
  stg(Cc.af, id,
      make_float4(0.5f * (a00.x + a01.x), 0.5f * (a10.y + a11.y), 0.5f * (a00.z + a10.z), 0.5f * (a01.w + a11.w)));
  stg(Cc.b, id, br);
  stg(Cc.ls, id, pr);
  stg(Cc.x, id, 0.0f);
}

// ---------------------------------------------------------------------------
// up: prolongation while staging, NS post-smoothing sweeps.  With rz (level 0
// of the CG preconditioner) z is masked to the pressure unknowns and the
// block partial of r . z goes to red[local block].
// ---------------------------------------------------------------------------
template <int NS, int HALO>
__global__ __launch_bounds__(MNT) void k_mg_up(Batch b, int pcg, int first, float weight, int rz) {
  constexpr int RW = MT_W + 2 * HALO, RH = MT_H + 2 * HALO, RN = RW * RH;
  __shared__ float sx[RN], sb[RN], sl[RN];
  __shared__ float4 sa[RN];
  __shared__ double wsum[MNT / 64];
  MSite s = mg_locate(b);
  const DevInst* I = s.I;
  const Level L = mg_level(I, s.lvl, pcg);
  const Level C = I->lev[s.lvl + 1];
  float* xin = (s.sel & 1) ? L.xs : L.x;
  float* xout = (s.sel & 1) ? L.x : L.xs;
  stage_tile<HALO>(L, xin, &C, (s.sel & 2) ? C.xs : C.x, s.tx0, s.ty0, sx, sb, sl, sa);
  __syncthreads();
  const float scale = (float)(1 << s.lvl);
  sweep_tile<HALO, NS>(L.w, L.h, s.tx0, s.ty0, s.ex, s.ey, first, weight, gs_c0(I, scale), sx, sb, sl, sa);
  if (!rz) {
    write_tile<HALO>(L, xout, s.tx0, s.ty0, sx);
    return;
  }
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k < (MT_W * MT_H) / MNT; ++k) {
    int t = threadIdx.x + k * MNT;
    int lx = t % MT_W, ly = t / MT_W;
    int gx = s.tx0 + lx, gy = s.ty0 + ly;
    if (gx < L.w && gy < L.h) {
      int q = (ly + HALO) * RW + lx + HALO;
      float z = sx[q];
      if (!(sl[q] < 0.0f) || closed4(sa[q])) z = 0.0f;
      stg(xout, (size_t)gy * L.w + gx, z);
      acc += (double)sb[q] * (double)z;
    }
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = wsum[0];
    for (int w = 1; w < MNT / 64; ++w) t += wsum[w];
    I->red[(size_t)blockIdx.x - b.start[s.slot]] = t;
  }
}

// ---------------------------------------------------------------------------
// Bottom of the V-cycle in one workgroup of 1024 threads (32 x 32): levels
// l0 .. nlev-1 staged in LDS; the level table lives in LDS too.  One
// workgroup owns the whole level, so x is read from and written back to the
// buffer Batch::sel names.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void b_sweep(int w, int h, int off, float* x, const float* bb, const float* ls,
                                        const float4* af, int parity, float c0) {
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int gy = ty; gy < h; gy += 32)
    for (int gx = tx; gx < w; gx += 32) {
      if (((gx + gy) & 1) != parity) continue;
      int q = off + gy * w + gx;
      x[q] = gs_lds(x, bb, ls, af, q, gx, gy, w, h, w, 1.0f, c0);
    }
  __syncthreads();
}

__global__ __launch_bounds__(1024) void k_bottom(Batch b, int pcg, int symmetric) {
  extern __shared__ float4 smem_f4[];
  __shared__ int lw[kMaxLevels + 1], lh[kMaxLevels + 1], loff[kMaxLevels + 1];
  const int slot = blockIdx.x;
  const DevInst* I = b.inst[slot];
  const int l0 = b.level[slot];
  const int nl = I->nlev;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  if (threadIdx.x == 0) {
    int total = 0;
    for (int l = l0; l < nl; ++l) {
      const Level& L = mg_level(I, l, pcg);
      lw[l] = L.w;
      lh[l] = L.h;
      loff[l] = total;
      total += L.w * L.h;
    }
    loff[nl] = total;
  }
  __syncthreads();
  const int total = loff[nl];
  float4* af = smem_f4;
  float* x = (float*)(af + total);
  float* bb = x + total;
  float* ls = bb + total;
  {
    const Level L = mg_level(I, l0, pcg);
    const float* xb = (b.sel[slot] & 1) ? L.xs : L.x;
    for (int t = threadIdx.x; t < L.w * L.h; t += 1024) {
      x[t] = ldg(xb, t);
      bb[t] = ldg(L.b, t);
      ls[t] = ldg(L.ls, t);
      af[t] = ldg(L.af, t);
    }
  }
  __syncthreads();
  // sweep counts ride in the batch: ex = pre, ey = post | coarsest << 16
  const int pre = b.ex[slot], post = b.ey[slot] & 0xffff, coarse = b.ey[slot] >> 16;
  for (int l = l0; l < nl; ++l) {
    const int w = lw[l], h = lh[l], off = loff[l];
    const float scale = (float)(1 << l);
    const float c0 = gs_c0(I, scale);
    if (l == nl - 1) {
      for (int k = 0; k < coarse; ++k) {
        b_sweep(w, h, off, x, bb, ls, af, 1, c0);
        b_sweep(w, h, off, x, bb, ls, af, 0, c0);
      }
      if (symmetric)
        for (int k = 0; k < coarse; ++k) {
          b_sweep(w, h, off, x, bb, ls, af, 0, c0);
          b_sweep(w, h, off, x, bb, ls, af, 1, c0);
        }
      break;
    }
    for (int k = 0; k < pre; ++k) {
      b_sweep(w, h, off, x, bb, ls, af, 1, c0);
      b_sweep(w, h, off, x, bb, ls, af, 0, c0);
    }
    // residual + restriction into level l + 1 (each coarse cell forms the
    // residuals of its four children)
    const int cw = lw[l + 1], ch = lh[l + 1], coff = loff[l + 1];
    const float factor = resid_factor(I, scale);
    for (int cy = ty; cy < ch; cy += 32)
      for (int cx = tx; cx < cw; cx += 32) {
        const int dxs[4] = {0, 0, 1, 1}, dys[4] = {0, 1, 0, 1};
        float r4[4], l4[4];
        float4 a4[4];
        for (int k = 0; k < 4; ++k) {
          int gx = 2 * cx + dxs[k], gy = 2 * cy + dys[k];
          bool in = mg_inb(gx, gy, w, h);
          int q = off + gy * w + gx;
          r4[k] = in ? resid_lds(x, bb, ls, af, q, gx, gy, w, h, w, factor) : 0.0f;
          l4[k] = in ? ls[q] : 0.0f;
          a4[k] = in ? af[q] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
        float br = 0.25f * (((r4[0] + r4[1]) + r4[2]) + r4[3]);
        float pr = l4[0];
        if (fabsf(l4[1]) < fabsf(pr)) pr = l4[1];
        if (fabsf(l4[2]) < fabsf(pr)) pr = l4[2];
        if (fabsf(l4[3]) < fabsf(pr)) pr = l4[3];
        // a4 order: (0,0) (0,1) (1,0) (1,1)
        int q = coff + cy * cw + cx;
        af[q] = make_float4(0.5f * (a4[0].x + a4[1].x), 0.5f * (a4[2].y + a4[3].y), 0.5f * (a4[0].z + a4[2].z),
                            0.5f * (a4[1].w + a4[3].w));
        bb[q] = br;
        ls[q] = pr;
        x[q] = 0.0f;
      }
    __syncthreads();
  }
  for (int l = nl - 2; l >= l0; --l) {
    const int w = lw[l], h = lh[l], off = loff[l];
    const int cw = lw[l + 1], ch = lh[l + 1], coff = loff[l + 1];
    for (int gy = ty; gy < h; gy += 32)
      for (int gx = tx; gx < w; gx += 32) {
        int cx = gx >> 1, cy = gy >> 1;
        int q = off + gy * w + gx;
        if (cx < cw && cy < ch && ls[q] < 0.0f) x[q] = x[coff + cy * cw + cx] + x[q];
      }
    __syncthreads();
    const float c0 = gs_c0(I, (float)(1 << l));
    for (int k = 0; k < post; ++k) {
      if (symmetric) {
        b_sweep(w, h, off, x, bb, ls, af, 0, c0);
        b_sweep(w, h, off, x, bb, ls, af, 1, c0);
      } else {
        b_sweep(w, h, off, x, bb, ls, af, 1, c0);
        b_sweep(w, h, off, x, bb, ls, af, 0, c0);
      }
    }
  }
  const Level L = mg_level(I, l0, pcg);
  float* xb = (b.sel[slot] & 1) ? L.xs : L.x;
  for (int t = threadIdx.x; t < L.w * L.h; t += 1024) stg(xb, t, x[t]);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
#define MG_LAUNCH(kern, nb, st, ...) hipLaunchKernelGGL(kern, dim3(nb), dim3(MNT), 0, st, __VA_ARGS__)

// nsweeps 0..3 pre-smoothing sweeps (+ residual/restriction)
void launch_mg_down(hipStream_t st, const Batch& b, int pcg, int nsweeps, int first, float weight, int restrict_) {
  int nb = b.start[b.n];
  if (nb <= 0) return;
  switch (nsweeps) {
    case 0: MG_LAUNCH((k_mg_down<0, 1>), nb, st, b, pcg, first, weight, restrict_); break;
    case 1: MG_LAUNCH((k_mg_down<1, 3>), nb, st, b, pcg, first, weight, restrict_); break;
    case 2: MG_LAUNCH((k_mg_down<2, 5>), nb, st, b, pcg, first, weight, restrict_); break;
    default: MG_LAUNCH((k_mg_down<3, 7>), nb, st, b, pcg, first, weight, restrict_); break;
  }
}

// prolongation + nsweeps 0..3 post-smoothing sweeps (+ r . z partials)
void launch_mg_up(hipStream_t st, const Batch& b, int pcg, int nsweeps, int first, float weight, int rz) {
  int nb = b.start[b.n];
  if (nb <= 0) return;
  switch (nsweeps) {
    case 0: MG_LAUNCH((k_mg_up<0, 0>), nb, st, b, pcg, first, weight, rz); break;
    case 1: MG_LAUNCH((k_mg_up<1, 2>), nb, st, b, pcg, first, weight, rz); break;
    case 2: MG_LAUNCH((k_mg_up<2, 4>), nb, st, b, pcg, first, weight, rz); break;
    default: MG_LAUNCH((k_mg_up<3, 6>), nb, st, b, pcg, first, weight, rz); break;
  }
}

size_t bottom_lds_bytes(int cells) { return (size_t)cells * (16 + 3 * 4); }

void launch_bottom(hipStream_t st, const Batch& b, int pcg, int symmetric, int max_cells) {
  if (b.n <= 0) return;
  hipLaunchKernelGGL(k_bottom, dim3(b.n), dim3(1024), bottom_lds_bytes(max_cells), st, b, pcg, symmetric);
}

int mg_tile_w() { return MT_W; }
int mg_tile_h() { return MT_H; }
int mg_bottom_cells() { return BOTTOM_CELLS; }

}  // namespace efl
