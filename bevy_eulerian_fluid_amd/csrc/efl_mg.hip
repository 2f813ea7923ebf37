// efl_mg.hip - fused multigrid for gfx950: the V-cycle of
// projection/multi_grid.rs:390-475 (and its symmetric variant used as the
// CG preconditioner) in a handful of launches per cycle.
//
//   smooth_tile     k red-black sweeps of one level in one launch: a block
//                   stages a 64 x 16 tile plus a 2k-cell halo of x, b, the
//                   air level set and the area fractions in LDS and runs the
//                   2k half-sweeps there (every half-sweep shrinks the region
//                   whose values are exact by one cell), then writes the tile.
//   resid_restrict  residual.wgsl + restriction.wgsl: each thread forms the
//                   residual of its 2 x 2 fine cells and writes the coarse
//                   b, level set, area fractions and x = 0; r is never stored.
//   prolong_smooth  prolongation.wgsl applied while staging the tile, then
//                   the post-smoothing sweeps as in smooth_tile.
//   bottom          one workgroup runs the whole V-cycle below a level of at
//                   most 1024 cells in LDS (reference order, no launches).
//
// Red-black Gauss-Seidel inside a colour is order independent, so these
// kernels produce exactly the values of the one-kernel-per-half-sweep
// reference dataflow (same operation order per cell; -ffp-contract=off).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "efl_internal.h"
#include "efl_launch.h"

namespace efl {

constexpr int MT_W = 64;        // tile width (one wavefront per tile row)
constexpr int MT_H = 16;        // tile height
constexpr int MNT = 256;        // threads per block
constexpr int BOTTOM_CELLS = 1024;

__device__ __forceinline__ bool mg_inb(int x, int y, int w, int h) {
  return (unsigned)x < (unsigned)w && (unsigned)y < (unsigned)h;
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
  return s;
}

// The GS update of gauss_seidel.wgsl:44-88 on LDS-staged data.  Neighbour
// (jx, jy) is read at LDS index q + offs[k]; in-bounds tests use global
// coordinates.
__device__ __forceinline__ float gs_lds(const float* sx, const float* sb, const float* sl, const float4* sa, int q,
                                        int gx, int gy, int w, int h, int rowlen, float weight, float c0) {
  float lc = sl[q];
  if (lc >= 0.0f) return 0.0f;
  float4 f = sa[q];
  float p_old = sx[q];
  if (f.x == 0.0f && f.y == 0.0f && f.z == 0.0f && f.w == 0.0f) return 0.0f;
  float denom = 0.0f;
  float nume = c0 * sb[q];
  const float fv[4] = {f.x, f.y, f.z, f.w};
  const int ox[4] = {-1, 1, 0, 0}, oy[4] = {0, 0, -1, 1};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (mg_inb(gx + ox[k], gy + oy[k], w, h)) {
      int j = q + ox[k] + oy[k] * rowlen;
      float lv = sl[j];
      if (lv < 0.0f) {
        denom += fv[k];
        nume += fv[k] * sx[j];
      } else {
        denom += fv[k] * (1.0f - lv / lc);
      }
    }
  }
  if (fabsf(denom) < 1.0e-6f) return 0.0f;
  float pn = nume / denom;
  return weight * pn + (1.0f - weight) * p_old;
}

// dx * dx * rho / dt, evaluated left to right as the WGSL does
__device__ __forceinline__ float gs_c0(const DevInst* I, float scale) {
  float dx = scale * I->dx;
  return dx * dx * I->rho / I->dt;
}

// Stage [tile + halo] of a level into LDS (x, b, level set, fractions); cells
// outside the level read 0 (they are never used as neighbours).  With
// `prolong`, prolongation.wgsl is applied to x on the fly.
template <int HALO>
__device__ void stage_tile(const Level& L, const Level* C, int x0, int y0, float* sx, float* sb, float* sl, float4* sa) {
  constexpr int RW = MT_W + 2 * HALO, RH = MT_H + 2 * HALO;
  for (int q = threadIdx.x; q < RW * RH; q += MNT) {
    int gx = x0 - HALO + q % RW, gy = y0 - HALO + q / RW;
    if (mg_inb(gx, gy, L.w, L.h)) {
      size_t id = (size_t)gy * L.w + gx;
      float xv = L.x[id];
      float lv = L.ls[id];
      if (C) {  // prolongation: x_f += x_c on fluid fine cells with a parent
        int cx = gx >> 1, cy = gy >> 1;
        if (cx < C->w && cy < C->h && lv < 0.0f) xv = C->x[(size_t)cy * C->w + cx] + xv;
      }
      sx[q] = xv;
      sb[q] = L.b[id];
      sl[q] = lv;
      sa[q] = L.af[id];
    } else {
      sx[q] = 0.0f;
      sb[q] = 0.0f;
      sl[q] = 0.0f;
      sa[q] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
  }
}

// 2 * nsweeps half-sweeps on the staged region; the first colour of each
// sweep is `first` (1 = red, 0 = black).  Cells outside [0, ex) x [0, ey)
// are not updated (dispatch coverage).
template <int HALO>
__device__ void sweep_tile(const Level& L, int x0, int y0, int ex, int ey, int nsweeps, int first, bool alternate,
                           float weight, float c0, float* sx, const float* sb, const float* sl, const float4* sa) {
  constexpr int RW = MT_W + 2 * HALO, RH = MT_H + 2 * HALO;
  int hs = 0;
  for (int s = 0; s < nsweeps; ++s) {
    int c1 = (alternate && (s & 1)) ? 1 - first : first;
    for (int half = 0; half < 2; ++half, ++hs) {
      int parity = half == 0 ? c1 : 1 - c1;
      int m = hs + 1;  // margin: values stay exact on the region shrunk by m
      int iw = RW - 2 * m, ih = RH - 2 * m;
      // cells of one colour: half of the inner region, two per row pair
      for (int t = threadIdx.x; t < iw * ih; t += MNT) {
        int lx = m + t % iw, ly = m + t / iw;
        int gx = x0 - HALO + lx, gy = y0 - HALO + ly;
        if (((gx + gy) & 1) != parity) continue;
        if (gx < 0 || gy < 0 || gx >= ex || gy >= ey || gx >= L.w || gy >= L.h) continue;
        int q = ly * RW + lx;
        sx[q] = gs_lds(sx, sb, sl, sa, q, gx, gy, L.w, L.h, RW, weight, c0);
      }
      __syncthreads();
    }
  }
}

template <int HALO>
__device__ void write_tile(const Level& L, int x0, int y0, int ex, int ey, const float* sx) {
  constexpr int RW = MT_W + 2 * HALO;
  for (int t = threadIdx.x; t < MT_W * MT_H; t += MNT) {
    int lx = t % MT_W, ly = t / MT_W;
    int gx = x0 + lx, gy = y0 + ly;
    if (gx < ex && gy < ey && gx < L.w && gy < L.h) L.x[(size_t)gy * L.w + gx] = sx[(ly + HALO) * RW + lx + HALO];
  }
}

template <int HALO>
__global__ __launch_bounds__(MNT) void k_smooth_tile(Batch b, int pcg, int nsweeps, int first, int alternate,
                                                      float weight, int prolong) {
  constexpr int RN = (MT_W + 2 * HALO) * (MT_H + 2 * HALO);
  __shared__ float sx[RN], sb[RN], sl[RN];
  __shared__ float4 sa[RN];
  MSite s = mg_locate(b);
  const DevInst* I = s.I;
  const Level& L = mg_level(I, s.lvl, pcg);
  const Level* C = prolong ? &I->lev[s.lvl + 1] : nullptr;
  stage_tile<HALO>(L, C, s.tx0, s.ty0, sx, sb, sl, sa);
  __syncthreads();
  float c0 = gs_c0(I, (float)(1 << s.lvl));
  sweep_tile<HALO>(L, s.tx0, s.ty0, s.ex, s.ey, nsweeps, first, alternate != 0, weight, c0, sx, sb, sl, sa);
  write_tile<HALO>(L, s.tx0, s.ty0, s.ex, s.ey, sx);
}

// residual.wgsl:12-57 at one fine cell, reading global memory
__device__ __forceinline__ float resid_at(const DevInst* I, const Level& F, int x, int y, float scale) {
  size_t id = (size_t)y * F.w + x;
  float phi = F.ls[id];
  if (phi >= 0.0f) return 0.0f;
  float4 f = F.af[id];
  if (f.x == 0.0f && f.y == 0.0f && f.z == 0.0f && f.w == 0.0f) return 0.0f;
  float xc = F.x[id];
  float dx = I->dx * scale;
  float factor = I->dt / I->rho / dx / dx;
  float res = F.b[id];
  const float fv[4] = {f.x, f.y, f.z, f.w};
  const int ox[4] = {-1, 1, 0, 0}, oy[4] = {0, 0, -1, 1};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    int jx = x + ox[k], jy = y + oy[k];
    if (mg_inb(jx, jy, F.w, F.h)) {
      size_t jd = (size_t)jy * F.w + jx;
      float pn = F.ls[jd];
      if (pn < 0.0f) res -= fv[k] * (xc - F.x[jd]) * factor;
      else res -= fv[k] * (1.0f - pn / phi) * xc * factor;
    }
  }
  return res;
}

// residual + restriction (restriction.wgsl:10-60), one coarse cell per thread;
// tiles are 64 x 4 coarse cells
__global__ __launch_bounds__(MNT) void k_resid_restrict(Batch b, int pcg) {
  int blk = blockIdx.x;
  int i = 0;
  while (i + 1 < b.n && blk >= b.start[i + 1]) ++i;
  int lb = blk - b.start[i];
  int cx = (lb % b.tiles_x[i]) * 64 + (int)(threadIdx.x % 64);
  int cy = (lb / b.tiles_x[i]) * 4 + (int)(threadIdx.x / 64);
  if (cx >= b.ex[i] || cy >= b.ey[i]) return;
  const DevInst* I = b.inst[i];
  int l = b.level[i];
  const Level& F = mg_level(I, l, pcg);
  const Level& C = I->lev[l + 1];
  float scale = (float)(1 << l);
  int fx = 2 * cx, fy = 2 * cy;
  auto r_at = [&](int x, int y) { return mg_inb(x, y, F.w, F.h) ? resid_at(I, F, x, y, scale) : 0.0f; };
  auto l_at = [&](int x, int y) { return mg_inb(x, y, F.w, F.h) ? F.ls[(size_t)y * F.w + x] : 0.0f; };
  auto a_at = [&](int x, int y) {
    return mg_inb(x, y, F.w, F.h) ? F.af[(size_t)y * F.w + x] : make_float4(0.f, 0.f, 0.f, 0.f);
  };
  float br = 0.25f * (((r_at(fx, fy) + r_at(fx, fy + 1)) + r_at(fx + 1, fy)) + r_at(fx + 1, fy + 1));
  float q0 = l_at(fx, fy), q1 = l_at(fx, fy + 1), q2 = l_at(fx + 1, fy), q3 = l_at(fx + 1, fy + 1);
  float pr = q0;
  if (fabsf(q1) < fabsf(pr)) pr = q1;
  if (fabsf(q2) < fabsf(pr)) pr = q2;
  if (fabsf(q3) < fabsf(pr)) pr = q3;
  float4 a00 = a_at(fx, fy), a10 = a_at(fx + 1, fy), a01 = a_at(fx, fy + 1), a11 = a_at(fx + 1, fy + 1);
  size_t id = (size_t)cy * C.w + cx;
  C.af[id] = make_float4(0.5f * (a00.x + a01.x), 0.5f * (a10.y + a11.y), 0.5f * (a00.z + a10.z), 0.5f * (a01.w + a11.w));
  C.b[id] = br;
  C.ls[id] = pr;
  C.x[id] = 0.0f;
}

// ---------------------------------------------------------------------------
// Bottom of the V-cycle in one workgroup: levels l0 .. nlev-1 staged in LDS.
// ---------------------------------------------------------------------------
struct BLev {
  int w, h, off;
};

__device__ void b_sweep(const BLev& L, float* x, const float* bb, const float* ls, const float4* af, int parity,
                        float c0) {
  for (int t = threadIdx.x; t < L.w * L.h; t += blockDim.x) {
    int gx = t % L.w, gy = t / L.w;
    if (((gx + gy) & 1) != parity) continue;
    int q = L.off + t;
    x[q] = gs_lds(x, bb, ls, af, q, gx, gy, L.w, L.h, L.w, 1.0f, c0);
  }
  __syncthreads();
}

__global__ __launch_bounds__(1024) void k_bottom(Batch b, int pcg, int symmetric) {
  extern __shared__ unsigned char smem[];
  const int slot = blockIdx.x;
  const DevInst* I = b.inst[slot];
  const int l0 = b.level[slot];
  const int nl = I->nlev;
  BLev lv[kMaxLevels];
  int total = 0;
  for (int l = l0; l < nl; ++l) {
    const Level& L = mg_level(I, l, pcg);
    lv[l] = {L.w, L.h, total};
    total += L.w * L.h;
  }
  float4* af = (float4*)smem;
  float* x = (float*)(af + total);
  float* bb = x + total;
  float* ls = bb + total;
  float* rr = ls + total;
  // stage level l0
  {
    const Level& L = mg_level(I, l0, pcg);
    for (int t = threadIdx.x; t < L.w * L.h; t += blockDim.x) {
      x[t] = L.x[t];
      bb[t] = L.b[t];
      ls[t] = L.ls[t];
      af[t] = L.af[t];
    }
  }
  __syncthreads();
  // sweep counts ride in the batch: ex = pre, ey = post | coarsest << 16
  const int pre = b.ex[slot], post = b.ey[slot] & 0xffff, coarse = b.ey[slot] >> 16;
  // down
  for (int l = l0; l < nl; ++l) {
    const BLev& L = lv[l];
    float scale = (float)(1 << l);
    float c0 = gs_c0(I, scale);
    if (l == nl - 1) {
      for (int k = 0; k < coarse; ++k) {
        b_sweep(L, x, bb, ls, af, 1, c0);
        b_sweep(L, x, bb, ls, af, 0, c0);
      }
      if (symmetric)
        for (int k = 0; k < coarse; ++k) {
          b_sweep(L, x, bb, ls, af, 0, c0);
          b_sweep(L, x, bb, ls, af, 1, c0);
        }
      break;
    }
    for (int k = 0; k < pre; ++k) {
      b_sweep(L, x, bb, ls, af, 1, c0);
      b_sweep(L, x, bb, ls, af, 0, c0);
    }
    // residual into rr
    float dxs = I->dx * scale;
    float factor = I->dt / I->rho / dxs / dxs;
    for (int t = threadIdx.x; t < L.w * L.h; t += blockDim.x) {
      int gx = t % L.w, gy = t / L.w;
      int q = L.off + t;
      float phi = ls[q];
      float4 f = af[q];
      float res = 0.0f;
      if (phi < 0.0f && !(f.x == 0.0f && f.y == 0.0f && f.z == 0.0f && f.w == 0.0f)) {
        float xc = x[q];
        res = bb[q];
        const float fv[4] = {f.x, f.y, f.z, f.w};
        const int ox[4] = {-1, 1, 0, 0}, oy[4] = {0, 0, -1, 1};
        for (int k = 0; k < 4; ++k) {
          if (mg_inb(gx + ox[k], gy + oy[k], L.w, L.h)) {
            int j = q + ox[k] + oy[k] * L.w;
            float pn = ls[j];
            if (pn < 0.0f) res -= fv[k] * (xc - x[j]) * factor;
            else res -= fv[k] * (1.0f - pn / phi) * xc * factor;
          }
        }
      }
      rr[q] = res;
    }
    __syncthreads();
    // restriction into level l + 1
    const BLev& Cc = lv[l + 1];
    for (int t = threadIdx.x; t < Cc.w * Cc.h; t += blockDim.x) {
      int cx = t % Cc.w, cy = t / Cc.w, fx = 2 * cx, fy = 2 * cy;
      auto R = [&](int xx, int yy) { return mg_inb(xx, yy, L.w, L.h) ? rr[L.off + yy * L.w + xx] : 0.0f; };
      auto Ls = [&](int xx, int yy) { return mg_inb(xx, yy, L.w, L.h) ? ls[L.off + yy * L.w + xx] : 0.0f; };
      auto A = [&](int xx, int yy) {
        return mg_inb(xx, yy, L.w, L.h) ? af[L.off + yy * L.w + xx] : make_float4(0.f, 0.f, 0.f, 0.f);
      };
      float br = 0.25f * (((R(fx, fy) + R(fx, fy + 1)) + R(fx + 1, fy)) + R(fx + 1, fy + 1));
      float q0 = Ls(fx, fy), q1 = Ls(fx, fy + 1), q2 = Ls(fx + 1, fy), q3 = Ls(fx + 1, fy + 1);
      float pr = q0;
      if (fabsf(q1) < fabsf(pr)) pr = q1;
      if (fabsf(q2) < fabsf(pr)) pr = q2;
      if (fabsf(q3) < fabsf(pr)) pr = q3;
      float4 a00 = A(fx, fy), a10 = A(fx + 1, fy), a01 = A(fx, fy + 1), a11 = A(fx + 1, fy + 1);
      int q = Cc.off + t;
      af[q] = make_float4(0.5f * (a00.x + a01.x), 0.5f * (a10.y + a11.y), 0.5f * (a00.z + a10.z), 0.5f * (a01.w + a11.w));
      bb[q] = br;
      ls[q] = pr;
      x[q] = 0.0f;
    }
    __syncthreads();
  }
  // up
  for (int l = nl - 2; l >= l0; --l) {
    const BLev& L = lv[l];
    const BLev& Cc = lv[l + 1];
    for (int t = threadIdx.x; t < Cc.w * Cc.h; t += blockDim.x) {
      int cx = t % Cc.w, cy = t / Cc.w;
      float corr = x[Cc.off + t];
      const int ox[4] = {0, 1, 0, 1}, oy[4] = {0, 0, 1, 1};
      for (int k = 0; k < 4; ++k) {
        int fx = 2 * cx + ox[k], fy = 2 * cy + oy[k];
        if (fx < L.w && fy < L.h) {
          int q = L.off + fy * L.w + fx;
          if (ls[q] < 0.0f) x[q] = corr + x[q];
        }
      }
    }
    __syncthreads();
    float c0 = gs_c0(I, (float)(1 << l));
    for (int k = 0; k < post; ++k) {
      if (symmetric) {
        b_sweep(L, x, bb, ls, af, 0, c0);
        b_sweep(L, x, bb, ls, af, 1, c0);
      } else {
        b_sweep(L, x, bb, ls, af, 1, c0);
        b_sweep(L, x, bb, ls, af, 0, c0);
      }
    }
  }
  // write level l0's x back (the only output used above)
  const Level& L = mg_level(I, l0, pcg);
  for (int t = threadIdx.x; t < L.w * L.h; t += blockDim.x) L.x[t] = x[t];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
// nsweeps in 0..3 per launch (the halo, 2 * nsweeps cells, sets the LDS size)
void launch_smooth_tile(hipStream_t st, const Batch& b, int pcg, int nsweeps, int first, int alternate, float weight,
                        int prolong) {
  int nb = b.start[b.n];
  if (nb <= 0) return;
  switch (nsweeps) {
    case 1:
      hipLaunchKernelGGL(k_smooth_tile<2>, dim3(nb), dim3(MNT), 0, st, b, pcg, 1, first, alternate, weight, prolong);
      break;
    case 2:
      hipLaunchKernelGGL(k_smooth_tile<4>, dim3(nb), dim3(MNT), 0, st, b, pcg, 2, first, alternate, weight, prolong);
      break;
    case 3:
      hipLaunchKernelGGL(k_smooth_tile<6>, dim3(nb), dim3(MNT), 0, st, b, pcg, 3, first, alternate, weight, prolong);
      break;
    default:  // 0 sweeps: prolongation only
      hipLaunchKernelGGL(k_smooth_tile<2>, dim3(nb), dim3(MNT), 0, st, b, pcg, 0, first, alternate, weight, prolong);
      break;
  }
}

void launch_resid_restrict(hipStream_t st, const Batch& b, int pcg) {
  int nb = b.start[b.n];
  if (nb > 0) hipLaunchKernelGGL(k_resid_restrict, dim3(nb), dim3(MNT), 0, st, b, pcg);
}

size_t bottom_lds_bytes(int cells) { return (size_t)cells * (16 + 4 * 4); }

void launch_bottom(hipStream_t st, const Batch& b, int pcg, int symmetric, int max_cells) {
  if (b.n <= 0) return;
  hipLaunchKernelGGL(k_bottom, dim3(b.n), dim3(1024), bottom_lds_bytes(max_cells), st, b, pcg, symmetric);
}

int mg_tile_w() { return MT_W; }
int mg_tile_h() { return MT_H; }
int mg_bottom_cells() { return BOTTOM_CELLS; }

}  // namespace efl
