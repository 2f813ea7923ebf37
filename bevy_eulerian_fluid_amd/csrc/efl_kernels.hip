// efl_kernels.hip - HIP kernels for gfx950 of the per-step fluid pipeline.
//
// Each kernel computes what one WGSL entry point of bevy_eulerian_fluid
// computes (the file:line it answers to is given at the kernel), with the
// same floating-point operation order so that results match the reference
// semantics bit for bit on IEEE hardware (the library is built with
// -ffp-contract=off; division and sqrt are correctly rounded).
//
// Execution model: every kernel is batched.  A launch covers a list of
// instances (efl::Batch); blockIdx.x is split into per-instance ranges and,
// inside a range, 64 x 4 tiles (one wavefront per row, so every global access
// of a row is one coalesced 256-byte transaction).  Extents per instance are
// the dispatch coverage of the reference (pipeline.rs:175-187), so grids that
// are not multiples of 8 / 64 leave exactly the cells the reference leaves.
//
// Out-of-bounds texture loads read 0 and out-of-bounds stores are dropped,
// as under Vulkan robustImageAccess2.

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "efl_internal.h"
#include "efl_launch.h"
#include "efl_reduce.h"

namespace efl {

constexpr int TX = 64;
constexpr int TY = 4;
constexpr int NT = TX * TY;
constexpr int SF_TH = 32;  // k_sample_forces tile height
static_assert(NT == RED_NT, "reduction kernels use RED_NT threads per block");

// ---------------------------------------------------------------------------
// indexing helpers
// ---------------------------------------------------------------------------
struct Site {
  const DevInst* I;
  int x, y;
  int ex, ey;
  int lvl;
  int slot;
  bool ok;
};

// The batch slot of this block.  Slot 0's entries are read first, in one
// round of scalar loads with b.n, and replaced only when the batch holds
// more than one instance (the search over Batch::start is a chain of
// dependent loads).
__device__ __forceinline__ int batch_slot(const Batch& b, int blk) {
  return b.n > 1 ? slot_of(b.start, b.n, blk) : 0;
}

__device__ __forceinline__ Site locate(const Batch& b) {
  const int blk = blockIdx.x;
  int st = b.start[0], tw = b.tiles_x[0], ex = b.ex[0], ey = b.ey[0], lvl = b.level[0];
  const DevInst* I = b.inst[0];
  const int i = batch_slot(b, blk);
  if (i > 0) {
    st = b.start[i];
    tw = b.tiles_x[i];
    ex = b.ex[i];
    ey = b.ey[i];
    lvl = b.level[i];
    I = b.inst[i];
  }
  const int lb = blk - st;
  const int tx = lb % tw, ty = lb / tw;
  Site s;
  s.I = I;
  s.x = tx * TX + (int)(threadIdx.x % TX);
  s.y = ty * TY + (int)(threadIdx.x / TX);
  s.ex = ex;
  s.ey = ey;
  s.lvl = lvl;
  s.slot = i;
  s.ok = s.x < s.ex && s.y < s.ey;
  return s;
}

__device__ __forceinline__ bool inb(int x, int y, int w, int h) {
  return (unsigned)x < (unsigned)w && (unsigned)y < (unsigned)h;
}
template <class T>
__device__ __forceinline__ T ldz(const T* p, int w, int h, int x, int y) {
  return inb(x, y, w, h) ? p[(size_t)y * w + x] : T(0);
}
__device__ __forceinline__ float4 ldz4(const float4* p, int w, int h, int x, int y) {
  return inb(x, y, w, h) ? p[(size_t)y * w + x] : make_float4(0.f, 0.f, 0.f, 0.f);
}
__device__ __forceinline__ float2 ldz2(const float2* p, int w, int h, int x, int y) {
  return inb(x, y, w, h) ? p[(size_t)y * w + x] : make_float2(0.f, 0.f);
}
template <class T>
__device__ __forceinline__ void stz(T* p, int w, int h, int x, int y, T v) {
  if (inb(x, y, w, h)) p[(size_t)y * w + x] = v;
}

// WGSL built-ins with the conventions of the reference pipeline
__device__ __forceinline__ float sat01(float v) { return fminf(fmaxf(v, 0.0f), 1.0f); }
__device__ __forceinline__ float lerp_w(float a, float b, float t) { return a * (1.0f - t) + b * t; }
__device__ __forceinline__ float sgn(float v) { return v > 0.0f ? 1.0f : (v < 0.0f ? -1.0f : 0.0f); }
__device__ __forceinline__ float hyp(float a, float b) { return sqrtf(a * a + b * b); }
__device__ __forceinline__ int to_i32(float v) {
  if (isnan(v)) return 0;
  if (v >= 2147483647.0f) return 2147483647;
  if (v <= -2147483648.0f) return (-2147483647 - 1);
  return (int)v;
}
// mat4 (column-major) * (x, y, 0, 1)
template <class P>
__device__ __forceinline__ float2 xform(P m, float x, float y) {
  float ox = ((m[0] * x + m[4] * y) + m[8] * 0.0f) + m[12] * 1.0f;
  float oy = ((m[1] * x + m[5] * y) + m[9] * 0.0f) + m[13] * 1.0f;
  return make_float2(ox, oy);
}

// area_fraction / area_fractions, utils/area_fraction.wgsl:10-56
__device__ __forceinline__ float face_open(float a, float b) { return sat01(fmaxf(a, b) / fabsf(a - b)); }
__device__ __forceinline__ float4 fractions_at(const float* ls, int w, int h, int x, int y) {
  float a0 = ldz(ls, w, h, x - 1, y - 1), a1 = ldz(ls, w, h, x, y - 1), a2 = ldz(ls, w, h, x + 1, y - 1);
  float a3 = ldz(ls, w, h, x - 1, y), a4 = ldz(ls, w, h, x, y), a5 = ldz(ls, w, h, x + 1, y);
  float a6 = ldz(ls, w, h, x - 1, y + 1), a7 = ldz(ls, w, h, x, y + 1), a8 = ldz(ls, w, h, x + 1, y + 1);
  float mm = 0.25f * (((a0 + a1) + a3) + a4);
  float pm = 0.25f * (((a1 + a2) + a4) + a5);
  float mp = 0.25f * (((a3 + a4) + a6) + a7);
  float pp = 0.25f * (((a4 + a5) + a7) + a8);
  return make_float4(face_open(mm, mp), face_open(pm, pp), face_open(mm, pm), face_open(mp, pp));
}
__device__ __forceinline__ bool all_closed(float4 f) { return f.x == 0.0f && f.y == 0.0f && f.z == 0.0f && f.w == 0.0f; }

// utils/coordinate.wgsl:51-94
__device__ __forceinline__ float sample_u(const float* u, int w, int h, float px, float py) {
  int i = to_i32(rintf(px)), j = to_i32(floorf(py));
  float fi = (px + 0.5f) - (float)i, fj = py - (float)j;
  float a = lerp_w(ldz(u, w, h, i, j), ldz(u, w, h, i + 1, j), fi);
  float b = lerp_w(ldz(u, w, h, i, j + 1), ldz(u, w, h, i + 1, j + 1), fi);
  return lerp_w(a, b, fj);
}
__device__ __forceinline__ float sample_v(const float* v, int w, int h, float px, float py) {
  int i = to_i32(floorf(px)), j = to_i32(rintf(py));
  float fi = px - (float)i, fj = (py + 0.5f) - (float)j;
  float a = lerp_w(ldz(v, w, h, i, j), ldz(v, w, h, i + 1, j), fi);
  float b = lerp_w(ldz(v, w, h, i, j + 1), ldz(v, w, h, i + 1, j + 1), fi);
  return lerp_w(a, b, fj);
}
template <class DI>
__device__ __forceinline__ float2 backtrace(DI I, float px, float py, float dt) {
  int W = I->W, H = I->H;
  float vx = sample_u(I->u0, W + 1, H, px, py), vy = sample_v(I->v0, W, H + 1, px, py);
  float hdt = 0.5f * dt;
  float mx = px - hdt * vx, my = py - hdt * vy;
  float wx = sample_u(I->u0, W + 1, H, mx, my), wy = sample_v(I->v0, W, H + 1, mx, my);
  return make_float2(px - dt * wx, py - dt * wy);
}

// ---------------------------------------------------------------------------
// K1 initialize_grid_center (initialize_grid_center.wgsl:5-15)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(NT) void k_init_center(Batch b) {
  Site s = locate(b);
  if (!s.ok) return;
  CDevInst* I = cdev(s.I);
  size_t id = (size_t)s.y * I->W + s.x;
  float level = (float)(uint32_t)(I->H - s.y);
  I->ls0[id] = level;
  I->ls1[id] = level;
  I->grad[id] = make_float2(0.0f, -1.0f);
}

// K2 initialize_grid_edge (initialize_grid_edge.wgsl:6-17): u at (x, y), v at (y, x)
__global__ __launch_bounds__(NT) void k_init_edge(Batch b) {
  Site s = locate(b);
  if (!s.ok) return;
  CDevInst* I = cdev(s.I);
  int W = I->W, H = I->H;
  stz(I->u0, W + 1, H, s.x, s.y, 0.0f);
  stz(I->u1, W + 1, H, s.x, s.y, 0.0f);
  stz(I->v0, W, H + 1, s.y, s.x, 0.0f);
  stz(I->v1, W, H + 1, s.y, s.x, 0.0f);
}

// ---------------------------------------------------------------------------
// K3 update_fluid_source (update_fluid_source.wgsl:30-107).  Two cells write
// each shared face in the reference; here each face is gathered from its two
// cells with the later (raster-order) cell taking precedence.
// ---------------------------------------------------------------------------
struct SrcOut { float level; float vx, vy; bool upd; };

__device__ __forceinline__ float box_level(float hx, float hy, float cx, float cy, float x, float y) {
  float dx = fabsf(cx - x) - hx, dy = fabsf(cy - y) - hy;
  if (dx < 0.0f) return dy < 0.0f ? fmaxf(dx, dy) : dy;
  return dy < 0.0f ? dx : hyp(dx, dy);
}

template <class DI>
__device__ SrcOut eval_sources(DI I, const efl_source* src, uint32_t n, int x, int y, float level_in) {
  float dimx = (float)I->W, dimy = (float)I->H;
  float ux = (float)x / dimx, uy = (float)y / dimy;
  float nx = ux * 2.0f + (-1.0f), ny = uy * (-2.0f) + 1.0f;
  float px = (0.5f * dimx) * nx, py = (0.5f * dimy) * ny;
  SrcOut o{level_in, 0.0f, 0.0f, false};
  bool has_source = false;
  for (uint32_t i = 0; i < n; ++i) {
    const efl_source& q = src[i];
    float sdf;
    if (q.shape == 0u) sdf = hyp(q.center[0] - px, q.center[1] - py) - q.data[0];
    else if (q.shape == 1u) sdf = box_level(q.data[0], q.data[1], q.center[0], q.center[1], px, py);
    else sdf = 0.0f;
    if (q.mode == 0u) {
      o.level = fminf(o.level, sdf);
      has_source = true;
      if (sdf < 0.0f) {
        o.vx = o.vx + q.velocity[0];
        o.vy = o.vy + q.velocity[1];
        o.upd = true;
      }
    } else if (!has_source) {
      o.level = fmaxf(o.level, -sdf);
    }
  }
  return o;
}

__global__ __launch_bounds__(NT) void k_fluid_source(Batch b, int oneshot) {
  Site s = locate(b);
  if (!s.ok) return;
  CDevInst* I = cdev(s.I);
  uint32_t n = oneshot ? I->nsrc_init : I->nsrc;
  const efl_source* src = oneshot ? I->src_init : I->src;
  int W = I->W;
  int x = s.x, y = s.y;
  size_t id = (size_t)y * W + x;
  SrcOut me = eval_sources(I, src, n, x, y, I->ls1[id]);
  I->ls1[id] = me.level;
  if (n == 0) return;
  // u face (x, y): cell (x, y) wins over cell (x - 1, y)
  if (me.upd) {
    I->u0[(size_t)y * (W + 1) + x] = me.vx;
  } else if (x > 0) {
    SrcOut l = eval_sources(I, src, n, x - 1, y, 0.0f);
    if (l.upd) I->u0[(size_t)y * (W + 1) + x] = l.vx;
  }
  if (x == s.ex - 1 && me.upd) I->u0[(size_t)y * (W + 1) + x + 1] = me.vx;
  // v face (x, y): cell (x, y) wins over cell (x, y - 1)
  if (me.upd) {
    I->v0[(size_t)y * W + x] = me.vy;
  } else if (y > 0) {
    SrcOut t = eval_sources(I, src, n, x, y - 1, 0.0f);
    if (t.upd) I->v0[(size_t)y * W + x] = t.vy;
  }
  if (y == s.ey - 1 && me.upd) I->v0[(size_t)(y + 1) * W + x] = me.vy;
}

// ---------------------------------------------------------------------------
// K4 update_solid (update_solid.wgsl:15-175, solid_obstacle.wgsl:43-89)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float seg_dist(float x, float y, float ax, float ay, float bx, float by) {
  float abx = bx - ax, aby = by - ay;
  float t = sat01(((x - ax) * abx + (y - ay) * aby) / (abx * abx + aby * aby));
  float qx = ax + t * abx, qy = ay + t * aby;
  return hyp(x - qx, y - qy);
}
__device__ __forceinline__ float side(float x, float y, float ax, float ay, float bx, float by) {
  float px = x - ax, py = y - ay, abx = bx - ax, aby = by - ay;
  return sgn(py * abx - px * aby);
}
__device__ __forceinline__ float det3(float a, float b, float c, float d, float e, float f, float g, float h, float i) {
  return a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
}
__device__ float det_mat4(const float* m) {  // Laplace expansion along the first row
  auto A = [&](int r, int c) { return m[c * 4 + r]; };
  float d0 = det3(A(1, 1), A(1, 2), A(1, 3), A(2, 1), A(2, 2), A(2, 3), A(3, 1), A(3, 2), A(3, 3));
  float d1 = det3(A(1, 0), A(1, 2), A(1, 3), A(2, 0), A(2, 2), A(2, 3), A(3, 0), A(3, 2), A(3, 3));
  float d2 = det3(A(1, 0), A(1, 1), A(1, 3), A(2, 0), A(2, 1), A(2, 3), A(3, 0), A(3, 1), A(3, 3));
  float d3 = det3(A(1, 0), A(1, 1), A(1, 2), A(2, 0), A(2, 1), A(2, 2), A(3, 0), A(3, 1), A(3, 2));
  return A(0, 0) * d0 - A(0, 1) * d1 + A(0, 2) * d2 - A(0, 3) * d3;
}
__device__ float obstacle_level(const efl_obstacle& o, float x, float y) {
  switch (o.shape) {
    case 0u:
      return hyp(x - o.transform[12], y - o.transform[13]) - o.values[0];
    case 1u: {
      if (det_mat4(o.transform) == 0.0f) return 1.0e6f;
      float2 q = xform(o.inverse_transform, x, y);
      float hx = o.values[0], hy = o.values[1];
      float ax = fabsf(q.x) - hx, ay = fabsf(q.y) - hy;
      bool inx = fabsf(q.x) < hx, iny = fabsf(q.y) < hy;
      if (inx) return iny ? fmaxf(ax, ay) : ay;
      return iny ? ax : hyp(ax, ay);
    }
    case 2u: {
      float2 a = xform(o.transform, o.values[0], o.values[1]);
      float2 c = xform(o.transform, o.values[2], o.values[3]);
      return seg_dist(x, y, a.x, a.y, c.x, c.y) - o.values[4];
    }
    case 4u: {
      float2 p0 = xform(o.transform, o.values[0], o.values[1]);
      float2 p1 = xform(o.transform, o.values[2], o.values[3]);
      float2 p2 = xform(o.transform, o.values[4], o.values[5]);
      float d0 = seg_dist(x, y, p0.x, p0.y, p1.x, p1.y);
      float d1 = seg_dist(x, y, p1.x, p1.y, p2.x, p2.y);
      float d2 = seg_dist(x, y, p2.x, p2.y, p0.x, p0.y);
      float m = fminf(fminf(d0, d1), d2);
      bool inside = side(x, y, p0.x, p0.y, p1.x, p1.y) > 0.0f && side(x, y, p1.x, p1.y, p2.x, p2.y) > 0.0f &&
                    side(x, y, p2.x, p2.y, p0.x, p0.y) > 0.0f;
      return inside ? -m : m;
    }
    default:
      return 1.0e6f;
  }
}

// A disk (cx, cy, R) in world units that holds the whole shape: the signed
// distance of every point p is then >= |p - c| - R (outside: the distance to
// the shape is at least that to the disk; inside: the boundary is no farther
// than the disk's).  R = -1 marks a shape whose level is the 1e6 sentinel.
__device__ float4 obstacle_bound(const efl_obstacle& o) {
  switch (o.shape) {
    case 0u: return make_float4(o.transform[12], o.transform[13], o.values[0], 0.0f);
    case 1u: {
      if (det_mat4(o.transform) == 0.0f) return make_float4(0.0f, 0.0f, -1.0f, 0.0f);
      // local q = A p + t (inverse transform): centre c = -A^-1 t, and
      // |q| >= s_min(A) |p - c|, so the radius in world units is
      // |half extents| / s_min(A)
      const float* m = o.inverse_transform;
      const double a = m[0], bb = m[4], c = m[1], d = m[5], tx = m[12], ty = m[13];
      const double det = a * d - bb * c;
      if (det == 0.0) return make_float4(0.0f, 0.0f, 3.0e38f, 0.0f);
      const double cx = -(d * tx - bb * ty) / det, cy = -(-c * tx + a * ty) / det;
      const double S = a * a + bb * bb + c * c + d * d;
      const double smin = sqrt(fmax(0.0, 0.5 * (S - sqrt(fmax(0.0, S * S - 4.0 * det * det)))));
      const double hr = sqrt((double)o.values[0] * o.values[0] + (double)o.values[1] * o.values[1]);
      return make_float4((float)cx, (float)cy, smin > 0.0 ? (float)(hr / smin) : 3.0e38f, 0.0f);
    }
    case 2u: {
      float2 a = xform(o.transform, o.values[0], o.values[1]);
      float2 c = xform(o.transform, o.values[2], o.values[3]);
      return make_float4(0.5f * (a.x + c.x), 0.5f * (a.y + c.y), 0.5f * hyp(c.x - a.x, c.y - a.y) + o.values[4], 0.0f);
    }
    case 4u: {
      float2 p0 = xform(o.transform, o.values[0], o.values[1]);
      float2 p1 = xform(o.transform, o.values[2], o.values[3]);
      float2 p2 = xform(o.transform, o.values[4], o.values[5]);
      float cx = (p0.x + p1.x + p2.x) / 3.0f, cy = (p0.y + p1.y + p2.y) / 3.0f;
      float r = fmaxf(fmaxf(hyp(p0.x - cx, p0.y - cy), hyp(p1.x - cx, p1.y - cy)), hyp(p2.x - cx, p2.y - cy));
      return make_float4(cx, cy, r, 0.0f);
    }
    default: return make_float4(0.0f, 0.0f, -1.0f, 0.0f);
  }
}

// An obstacle can change a cell's results only where one of its three
// levels (cell centre, u face, v face) is below the current level or below
// 0.5; obstacles whose bounding disk proves otherwise (with a one-cell margin
// for rounding) are skipped, the others are evaluated in index order as the
// reference loops (the minimum, and the last obstacle below 0.5 winning the
// id and the face velocities, are unchanged).
__global__ __launch_bounds__(NT) void k_update_solid(Batch b) {
  __shared__ float4 bnd[EFL_MAX_SOLIDS];
  __shared__ int cand[EFL_MAX_SOLIDS];
  __shared__ float wmin[NT / 64];
  __shared__ int wcnt[NT / 64];
  static_assert(EFL_MAX_SOLIDS <= NT, "one thread per obstacle");
  Site s = locate(b);
  CDevInst* I = cdev(s.I);
  const int W = I->W, H = I->H, x = s.x, y = s.y;
  auto world = [&](float fx, float fy) {
    float ux = fx / (float)W, uy = fy / (float)H;
    return xform(I->T, (ux - 0.5f) * I->sizex, (-uy + 0.5f) * I->sizey);
  };
  // The tile's candidates: the obstacles that can change any of its cells'
  // results, in index order.  The cells' sample points (centres and the
  // faces half a cell below and left) lie in a parallelogram, the world
  // image of [x0 - 0.5, x0 + TX - 1] x [y0 - 0.5, y0 + TY - 1], inside the
  // disk (tc, rt) through its corners.  Per obstacle with bounding disk
  // (c, R) and d = |tc - c|: every cell's bound lb = |pc - c| - R - 1 is at
  // least L = d - rt - R - 2 (one cell of slack for rounding), and its level
  // anywhere in the tile at most U = d + rt + R + 1.  The border distance
  // bounds a cell's level too, so every cell ends with level <= Umin (the
  // least of those bounds).  An obstacle with L > Umin and L - fd >= 0.5 is
  // then skipped by the per-cell tests at every cell of the tile (lb >=
  // level, lb >= 0.5, lb - fd >= 0.5) and is not the cell's smallest bound
  // either (the smallest bound is at most Umin), so leaving it out of both
  // loops changes no result.
  const int x0 = x - (int)(threadIdx.x % TX), y0 = y - (int)(threadIdx.x / TX);
  const float2 q0 = world((float)x0 - 0.5f, (float)y0 - 0.5f), q3 = world((float)(x0 + TX - 1), (float)(y0 + TY - 1));
  const float2 q1 = world((float)(x0 + TX - 1), (float)y0 - 0.5f), q2 = world((float)x0 - 0.5f, (float)(y0 + TY - 1));
  const float tcx = 0.5f * (q0.x + q3.x), tcy = 0.5f * (q0.y + q3.y);
  const float rt = fmaxf(fmaxf(hyp(q0.x - tcx, q0.y - tcy), hyp(q3.x - tcx, q3.y - tcy)),
                         fmaxf(hyp(q1.x - tcx, q1.y - tcy), hyp(q2.x - tcx, q2.y - tcy)));
  const float2 c00 = world((float)x0, (float)y0), u00 = world((float)x0 - 0.5f, (float)y0),
               v00 = world((float)x0, (float)y0 - 0.5f);
  const float fdt = fmaxf(hyp(u00.x - c00.x, u00.y - c00.y), hyp(v00.x - c00.x, v00.y - c00.y));
  // the largest border distance min(tx, ty) over the tile's cells
  const int xe = min(x0 + TX - 1, W - 1), ye = min(y0 + TY - 1, H - 1);
  const int xm = min(max((W - 1) / 2, x0), xe), ym = min(max((H - 1) / 2, y0), ye);
  const float ub = (float)min(min(xm, abs(W - 1 - xm)), min(ym, abs(H - 1 - ym)));
  const uint32_t nobs = I->nobs;
  const int k = (int)threadIdx.x;
  float L = 0.0f, U = 3.0e38f;
  bool use = false;
  if ((uint32_t)k < nobs) {
    const float4 bd = obstacle_bound(I->obs[k]);
    bnd[k] = bd;
    use = bd.z >= 0.0f;  // R = -1: level 1e6 everywhere, changes nothing
    const float d = hyp(tcx - bd.x, tcy - bd.y);
    L = d - rt - bd.z - 2.0f;
    if (use) U = d + rt + bd.z + 1.0f;
  }
  float um = U;
  for (int off = 32; off > 0; off >>= 1) um = fminf(um, __shfl_xor(um, off, 64));
  if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = um;
  __syncthreads();
  float umin = ub;
  for (int w = 0; w < NT / 64; ++w) umin = fminf(umin, wmin[w]);
  use = use && (L <= umin || L - fdt < 0.5f);
  const unsigned long long m = __ballot(use);
  if ((threadIdx.x & 63) == 0) wcnt[threadIdx.x >> 6] = __popcll(m);
  __syncthreads();
  int base = 0, ncand = 0;
  for (int w = 0; w < NT / 64; ++w) {
    base += w < (int)(threadIdx.x >> 6) ? wcnt[w] : 0;
    ncand += wcnt[w];
  }
  if (use) cand[base + __popcll(m & ((1ull << (threadIdx.x & 63)) - 1ull))] = k;
  __syncthreads();
  if (!s.ok) return;
  size_t id = (size_t)y * W + x;
  if (x == 0 || y == 0 || x == W - 1 || y == H - 1) {
    I->lss[id] = 0.0f;
    I->us[(size_t)y * (W + 1) + x] = 0.0f;
    I->vs[id] = 0.0f;
    I->sid[id] = -1;
    return;
  }
  float2 pc = world((float)x, (float)y);
  float2 pu = world((float)x - 0.5f, (float)y - 0.0f);
  float2 pv = world((float)x - 0.0f, (float)y - 0.5f);
  int tx = min(x, abs(W - 1 - x)), ty = min(y, abs(H - 1 - y));
  float level = (float)min(tx, ty);
  float uu = 0.0f, vv = 0.0f;
  int32_t id_s = -1;
  // the faces' sample points lie within fd of the centre
  const float fd = fmaxf(hyp(pu.x - pc.x, pu.y - pc.y), hyp(pv.x - pc.x, pv.y - pc.y));
  // lb = |pc - bounding centre| - radius - 1 bounds a body's level at the
  // centre from below, and lb - fd its level at the faces.  The body with
  // the smallest bound is evaluated first, so that `level` (the minimum,
  // order independent) is tight before the pass in index order, which keeps
  // the reference's rule that the last body in index order sets the id and
  // the face velocities; a body is evaluated at the centre only if it could
  // lower the level or set the id, at the faces only if it could reach 0.5.
  // Both passes run over the tile's candidates (index order).
  int kb = -1;
  float lbest = 3.0e38f;
  for (int j = 0; j < ncand; ++j) {
    const int kk = __builtin_amdgcn_readfirstlane(cand[j]);
    const float4 bd = bnd[kk];
    const float lb = hyp(pc.x - bd.x, pc.y - bd.y) - bd.z - 1.0f;
    if (lb < lbest) {
      lbest = lb;
      kb = kk;
    }
  }
  float lc_best = 0.0f;
  if (kb >= 0) {
    lc_best = obstacle_level(I->obs[kb], pc.x, pc.y);
    if (lc_best < level) level = lc_best;
  }
  for (int j = 0; j < ncand; ++j) {
    const int kk = __builtin_amdgcn_readfirstlane(cand[j]);
    const float4 bd = bnd[kk];
    const float lb = hyp(pc.x - bd.x, pc.y - bd.y) - bd.z - 1.0f;
    const bool centre = kk == kb || lb < level || lb < 0.5f, faces = lb - fd < 0.5f;
    if (!centre && !faces) continue;
    const efl_obstacle& o = I->obs[kk];
    if (centre) {
      const float lc = kk == kb ? lc_best : obstacle_level(o, pc.x, pc.y);
      if (lc < level) level = lc;
      if (lc < 0.5f) id_s = (int32_t)o.entity_id;
    }
    if (!faces) continue;
    if (obstacle_level(o, pu.x, pu.y) < 0.5f) {
      float ry = pu.y - o.transform[13];
      uu = o.linear_velocity[0] + (-o.angular_velocity * ry);
    }
    if (obstacle_level(o, pv.x, pv.y) < 0.5f) {
      float rx = pv.x - o.transform[12];
      vv = -(o.linear_velocity[1] + o.angular_velocity * rx);
    }
  }
  I->us[(size_t)y * (W + 1) + x] = uu;
  I->vs[id] = vv;
  I->lss[id] = level;
  I->sid[id] = id_s;
}

// K5 update_area_fraction (update_area_fraction.wgsl:6-14)
__global__ __launch_bounds__(NT) void k_area_fraction(Batch b) {
  Site s = locate(b);
  if (!s.ok) return;
  CDevInst* I = cdev(s.I);
  I->afs[(size_t)s.y * I->W + s.x] = fractions_at(I->lss, I->W, I->H, s.x, s.y);
}

// ---------------------------------------------------------------------------
// K6 / K7 advect_u, advect_v (advect_velocity.wgsl:11-52)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(NT) void k_advect_u(Batch b) {
  Site s = locate(b);
  if (!s.ok) return;
  CDevInst* I = cdev(s.I);
  int W = I->W, H = I->H;
  float2 q = backtrace(I, (float)s.x + (-0.5f), (float)s.y + 0.0f, I->dt);
  bool inside = q.x >= 0.0f && q.y >= 0.0f && q.x <= (float)(W + 1) - 1.0f && q.y <= (float)H - 1.0f;
  size_t id = (size_t)s.y * (W + 1) + s.x;
  I->u1[id] = inside ? sample_u(I->u0, W + 1, H, q.x, q.y) : I->u0[id];
}
__global__ __launch_bounds__(NT) void k_advect_v(Batch b) {
  Site s = locate(b);
  if (!s.ok) return;
  CDevInst* I = cdev(s.I);
  int W = I->W, H = I->H;
  float2 q = backtrace(I, (float)s.x + 0.0f, (float)s.y + (-0.5f), I->dt);
  bool inside = q.x >= 0.0f && q.y >= 0.0f && q.x <= (float)W - 1.0f && q.y <= (float)(H + 1) - 1.0f;
  size_t id = (size_t)s.y * W + s.x;
  I->v1[id] = inside ? sample_v(I->v0, W, H + 1, q.x, q.y) : I->v0[id];
}

// ---------------------------------------------------------------------------
// K8 / K9 apply_forces_u, apply_forces_v (apply_forces.wgsl:17-76)
// ---------------------------------------------------------------------------
template <class DI>
__device__ __forceinline__ float local_force(DI I, int comp, int x, int y) {
  const float b = -1.0f / ((2.0f * 10.0f) * 10.0f);
  float net = comp == 0 ? I->gx : I->gy;
  for (int n = (int)I->nforces - 1; n >= 0; --n) {
    const efl_force& f = I->forces[n];
    float dx = (float)x - f.position[0], dy = (float)y - f.position[1];
    net = net + f.force[comp] * expf(b * (dx * dx + dy * dy));
  }
  return net;
}
__global__ __launch_bounds__(NT) void k_forces_u(Batch b) {
  Site s = locate(b);
  if (!s.ok) return;
  CDevInst* I = cdev(s.I);
  int W = I->W, H = I->H;
  size_t id = (size_t)s.y * (W + 1) + s.x;
  float4 fa = fractions_at(I->ls0, W, H, s.x, s.y);
  float fs = ldz4(I->afs, W, H, s.x, s.y).x;
  if (fa.x == 1.0f || fs == 0.0f) { I->u1[id] = 0.0f; return; }
  float net = local_force(I, 0, s.x, s.y);
  I->u1[id] = I->u1[id] + (net * I->dt) / I->dx;
}
__global__ __launch_bounds__(NT) void k_forces_v(Batch b) {
  Site s = locate(b);
  if (!s.ok) return;
  CDevInst* I = cdev(s.I);
  int W = I->W, H = I->H;
  size_t id = (size_t)s.y * W + s.x;
  float4 fa = fractions_at(I->ls0, W, H, s.x, s.y);
  float fs = ldz4(I->afs, W, H, s.x, s.y).z;
  if (fa.z == 1.0f || fs == 0.0f) { I->v1[id] = 0.0f; return; }
  float net = local_force(I, 1, s.x, s.y);
  I->v1[id] = I->v1[id] + (net * I->dt) / I->dx;
}

// K6 + K8 (and K7 + K9) in one launch for the step: the advected face value
// stays in a register and the forces are applied to it, with the same
// expressions as k_advect_u / k_forces_u (k_advect_v / k_forces_v).  The two
// dispatches cover the same faces and the forces read only the air level
// set and the area fractions at the face, so the results are the same bits
// as the two launches'; the stage-by-stage entry points keep those.
template <int COMP>
__device__ __forceinline__ void advect_forces_at(const Site& s) {
  CDevInst* I = cdev(s.I);
  const int W = I->W, H = I->H;
  float val;
  size_t id;
  if (COMP == 0) {
    float2 q = backtrace(I, (float)s.x + (-0.5f), (float)s.y + 0.0f, I->dt);
    bool inside = q.x >= 0.0f && q.y >= 0.0f && q.x <= (float)(W + 1) - 1.0f && q.y <= (float)H - 1.0f;
    id = (size_t)s.y * (W + 1) + s.x;
    val = inside ? sample_u(I->u0, W + 1, H, q.x, q.y) : I->u0[id];
  } else {
    float2 q = backtrace(I, (float)s.x + 0.0f, (float)s.y + (-0.5f), I->dt);
    bool inside = q.x >= 0.0f && q.y >= 0.0f && q.x <= (float)W - 1.0f && q.y <= (float)(H + 1) - 1.0f;
    id = (size_t)s.y * W + s.x;
    val = inside ? sample_v(I->v0, W, H + 1, q.x, q.y) : I->v0[id];
  }
  const float4 fa = fractions_at(I->ls0, W, H, s.x, s.y);
  const float4 fs4 = ldz4(I->afs, W, H, s.x, s.y);
  const float fo = COMP == 0 ? fa.x : fa.z, fs = COMP == 0 ? fs4.x : fs4.z;
  if (fo == 1.0f || fs == 0.0f) {
    val = 0.0f;
  } else {
    const float net = local_force(I, COMP, s.x, s.y);
    val = val + (net * I->dt) / I->dx;
  }
  (COMP == 0 ? I->u1 : I->v1)[id] = val;
}
template <int COMP>
__global__ __launch_bounds__(NT) void k_advect_forces(Batch b) {
  Site s = locate(b);
  if (s.ok) advect_forces_at<COMP>(s);
}
// both components in one launch (Batch::level 0: u faces, 1: v faces)
__global__ __launch_bounds__(NT) void k_advect_forces_uv(Batch b) {
  Site s = locate(b);
  if (!s.ok) return;
  if (s.lvl == 0) advect_forces_at<0>(s);
  else advect_forces_at<1>(s);
}

// ---------------------------------------------------------------------------
// K10 divergence (divergence.wgsl:16-42)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(NT) void k_divergence(Batch b) {
  Site s = locate(b);
  if (!s.ok) return;
  CDevInst* I = cdev(s.I);
  int W = I->W, H = I->H, x = s.x, y = s.y;
  float4 f = fractions_at(I->lss, W, H, x, y);
  float du_f = f.y * ldz(I->u1, W + 1, H, x + 1, y) - f.x * ldz(I->u1, W + 1, H, x, y);
  float dv_f = f.w * ldz(I->v1, W, H + 1, x, y + 1) - f.z * ldz(I->v1, W, H + 1, x, y);
  float du_s = (1.0f - f.y) * ldz(I->us, W + 1, H, x + 1, y) - (1.0f - f.x) * ldz(I->us, W + 1, H, x, y);
  float dv_s = (1.0f - f.w) * ldz(I->vs, W, H + 1, x, y + 1) - (1.0f - f.z) * ldz(I->vs, W, H + 1, x, y);
  I->div[(size_t)y * W + x] = -(((du_f + dv_f) + du_s) + dv_s) / I->dx;
}

// ---------------------------------------------------------------------------
// Pressure operator (gauss_seidel.wgsl:44-88, residual.wgsl:12-57)
// ---------------------------------------------------------------------------
__device__ __forceinline__ const Level& level_of(const DevInst* I, int l, int pcg) {
  return (pcg && l == 0) ? I->lev_pcg0 : I->lev[l];
}
__device__ __forceinline__ Level level_of(CDevInst* I, int l, int pcg) {
  return (pcg && l == 0) ? lev_copy(I->lev_pcg0) : lev_copy(I->lev[l]);
}

template <class DI>
__device__ __forceinline__ float gs_value(DI I, const Level& L, int x, int y, float weight, float scale) {
  size_t id = (size_t)y * L.w + x;
  float lc = L.ls[id];
  if (lc >= 0.0f) return 0.0f;
  float4 f = L.af[id];
  float p_old = L.x[id];
  if (all_closed(f)) return 0.0f;
  float dx = scale * I->dx;
  float denom = 0.0f;
  float nume = dx * dx * I->rho / I->dt * L.b[id];
  const float fv[4] = {f.x, f.y, f.z, f.w};
  const int ox[4] = {-1, 1, 0, 0}, oy[4] = {0, 0, -1, 1};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    int jx = x + ox[k], jy = y + oy[k];
    if (inb(jx, jy, L.w, L.h)) {
      size_t jd = (size_t)jy * L.w + jx;
      float lv = L.ls[jd];
      if (lv < 0.0f) {
        denom += fv[k];
        nume += fv[k] * L.x[jd];
      } else {
        denom += fv[k] * (1.0f - lv / lc);
      }
    }
  }
  if (fabsf(denom) < 1.0e-6f) return 0.0f;
  float pn = nume / denom;
  return weight * pn + (1.0f - weight) * p_old;
}

// gauss_seidel_red (parity 1) / gauss_seidel_black (parity 0), :12-42
__global__ __launch_bounds__(NT) void k_gs_color(Batch b, int parity, float weight, int pcg) {
  Site s = locate(b);
  if (!s.ok) return;
  CDevInst* I = cdev(s.I);
  const Level& L = level_of(I, s.lvl, pcg);
  if (s.x >= L.w || s.y >= L.h) return;
  if (((s.x + s.y) & 1) != parity) return;
  float v = gs_value(I, L, s.x, s.y, weight, (float)(1 << s.lvl));
  L.x[(size_t)s.y * L.w + s.x] = v;
}

template <class DI>
__device__ __forceinline__ float residual_value(DI I, const Level& L, int x, int y, float scale) {
  size_t id = (size_t)y * L.w + x;
  float phi = L.ls[id];
  if (phi >= 0.0f) return 0.0f;
  float4 f = L.af[id];
  if (all_closed(f)) return 0.0f;
  float xc = L.x[id];
  float dx = I->dx * scale;
  float factor = I->dt / I->rho / dx / dx;
  float res = L.b[id];
  const float fv[4] = {f.x, f.y, f.z, f.w};
  const int ox[4] = {-1, 1, 0, 0}, oy[4] = {0, 0, -1, 1};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    int jx = x + ox[k], jy = y + oy[k];
    if (inb(jx, jy, L.w, L.h)) {
      size_t jd = (size_t)jy * L.w + jx;
      float pn = L.ls[jd];
      if (pn < 0.0f) res -= fv[k] * (xc - L.x[jd]) * factor;
      else res -= fv[k] * (1.0f - pn / phi) * xc * factor;
    }
  }
  return res;
}

// residual.wgsl:12-57
__global__ __launch_bounds__(NT) void k_residual(Batch b, int pcg) {
  Site s = locate(b);
  if (!s.ok) return;
  CDevInst* I = cdev(s.I);
  const Level& L = level_of(I, s.lvl, pcg);
  L.r[(size_t)s.y * L.w + s.x] = residual_value(I, L, s.x, s.y, (float)(1 << s.lvl));
}

// restriction.wgsl:10-60, level lvl -> lvl + 1 (extent = coarse grid)
__global__ __launch_bounds__(NT) void k_restrict(Batch b, int pcg) {
  Site s = locate(b);
  if (!s.ok) return;
  CDevInst* I = cdev(s.I);
  const Level& F = level_of(I, s.lvl, pcg);
  const Level Cc = lev_copy(I->lev[s.lvl + 1]);
  int x = s.x, y = s.y, fx = 2 * x, fy = 2 * y;
  float br = 0.25f * (((ldz(F.r, F.w, F.h, fx, fy) + ldz(F.r, F.w, F.h, fx, fy + 1)) + ldz(F.r, F.w, F.h, fx + 1, fy)) +
                      ldz(F.r, F.w, F.h, fx + 1, fy + 1));
  float q0 = ldz(F.ls, F.w, F.h, fx, fy), q1 = ldz(F.ls, F.w, F.h, fx, fy + 1);
  float q2 = ldz(F.ls, F.w, F.h, fx + 1, fy), q3 = ldz(F.ls, F.w, F.h, fx + 1, fy + 1);
  float pr = q0;
  if (fabsf(q1) < fabsf(pr)) pr = q1;
  if (fabsf(q2) < fabsf(pr)) pr = q2;
  if (fabsf(q3) < fabsf(pr)) pr = q3;
  float4 a00 = ldz4(F.af, F.w, F.h, fx, fy), a10 = ldz4(F.af, F.w, F.h, fx + 1, fy);
  float4 a01 = ldz4(F.af, F.w, F.h, fx, fy + 1), a11 = ldz4(F.af, F.w, F.h, fx + 1, fy + 1);
  size_t id = (size_t)y * Cc.w + x;
  Cc.af[id] = make_float4(0.5f * (a00.x + a01.x), 0.5f * (a10.y + a11.y), 0.5f * (a00.z + a10.z), 0.5f * (a01.w + a11.w));
  Cc.b[id] = br;
  Cc.ls[id] = pr;
  Cc.x[id] = 0.0f;
}

// prolongation.wgsl:5-34, level lvl + 1 -> lvl (extent = coarse grid)
__global__ __launch_bounds__(NT) void k_prolong(Batch b, int pcg) {
  Site s = locate(b);
  if (!s.ok) return;
  CDevInst* I = cdev(s.I);
  const Level& F = level_of(I, s.lvl, pcg);
  const Level Cc = lev_copy(I->lev[s.lvl + 1]);
  float corr = Cc.x[(size_t)s.y * Cc.w + s.x];
  const int ox[4] = {0, 1, 0, 1}, oy[4] = {0, 0, 1, 1};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    int fx = 2 * s.x + ox[k], fy = 2 * s.y + oy[k];
    if (fx < F.w && fy < F.h) {
      size_t fd = (size_t)fy * F.w + fx;
      if (F.ls[fd] < 0.0f) F.x[fd] = corr + F.x[fd];
    }
  }
}

// K16 / K17 jacobi_iteration (+ REVERSE), jacobi_iteration.wgsl:18-91
__global__ __launch_bounds__(NT) void k_jacobi(Batch b, int reverse) {
  Site s = locate(b);
  if (!s.ok) return;
  CDevInst* I = cdev(s.I);
  int W = I->W, H = I->H, x = s.x, y = s.y;
  const float* src = reverse ? I->p1 : I->p0;
  float* dst = reverse ? I->p0 : I->p1;
  size_t id = (size_t)y * W + x;
  float lc = I->ls0[id];
  if (lc >= 0.0f) { dst[id] = 0.0f; return; }
  float4 f = fractions_at(I->lss, W, H, x, y);
  if (all_closed(f)) { dst[id] = src[id]; return; }
  float pm = ldz(src, W, H, x - 1, y), pp = ldz(src, W, H, x + 1, y);
  float qm = ldz(src, W, H, x, y - 1), qp = ldz(src, W, H, x, y + 1);
  float ffm = fminf(0.0f, ldz(I->ls0, W, H, x - 1, y) / lc);
  float ffp = fminf(0.0f, ldz(I->ls0, W, H, x + 1, y) / lc);
  float fgm = fminf(0.0f, ldz(I->ls0, W, H, x, y - 1) / lc);
  float fgp = fminf(0.0f, ldz(I->ls0, W, H, x, y + 1) / lc);
  float coef = ((f.x * (1.0f - ffm) + f.y * (1.0f - ffp)) + f.z * (1.0f - fgm)) + f.w * (1.0f - fgp);
  if (fabsf(coef) < 1.0e-6f) { dst[id] = 0.0f; return; }
  float factor = I->dx * I->rho / I->dt;
  auto stepf = [](float v) { return v >= 0.0f ? 1.0f : 0.0f; };
  float a = f.x * stepf(ffm) * pm, bb = f.y * stepf(ffp) * pp;
  float c = f.z * stepf(fgm) * qm, d = f.w * stepf(fgp) * qp;
  dst[id] = ((((a + bb) + c) + d) - factor * I->div[id]) / coef;
}

// ---------------------------------------------------------------------------
// K18 / K19 solve_velocity_u, solve_velocity_v
// ---------------------------------------------------------------------------
__device__ __forceinline__ void solve_u_at(const Site& s) {
  CDevInst* I = cdev(s.I);
  int W = I->W, H = I->H, x = s.x, y = s.y;
  size_t id = (size_t)y * (W + 1) + x;
  float factor = I->dt / (I->dx * I->rho);
  if (ldz4(I->afs, W, H, x, y).x == 0.0f) { I->u0[id] = I->us[id]; return; }
  float pij = ldz(I->p0, W, H, x, y), pim = ldz(I->p0, W, H, x - 1, y);
  float lp = ldz(I->ls0, W, H, x, y), lm = ldz(I->ls0, W, H, x - 1, y);
  if (lm >= 0.0f && lp < 0.0f) pim = lm / lp * pij;
  else if (lm < 0.0f && lp >= 0.0f) pij = lp / lm * pim;
  else if (lm >= 0.0f && lp >= 0.0f) { I->u0[id] = 0.0f; return; }
  I->u0[id] = I->u1[id] - factor * (pij - pim);
}
__device__ __forceinline__ void solve_v_at(const Site& s) {
  CDevInst* I = cdev(s.I);
  int W = I->W, H = I->H, x = s.x, y = s.y;
  size_t id = (size_t)y * W + x;
  float factor = I->dt / (I->dx * I->rho);
  if (ldz4(I->afs, W, H, x, y).z == 0.0f) { I->v0[id] = I->vs[id]; return; }
  float pij = ldz(I->p0, W, H, x, y), pjm = ldz(I->p0, W, H, x, y - 1);
  float lp = ldz(I->ls0, W, H, x, y), lm = ldz(I->ls0, W, H, x, y - 1);
  if (lm >= 0.0f && lp < 0.0f) pjm = lm / lp * pij;
  else if (lm < 0.0f && lp >= 0.0f) pij = lp / lm * pjm;
  I->v0[id] = I->v1[id] - factor * (pij - pjm);
}

__global__ __launch_bounds__(NT) void k_solve_u(Batch b) {
  Site s = locate(b);
  if (s.ok) solve_u_at(s);
}
__global__ __launch_bounds__(NT) void k_solve_v(Batch b) {
  Site s = locate(b);
  if (s.ok) solve_v_at(s);
}
// both components in one launch: slots with Batch::level 0 are u faces,
// level 1 v faces (uniform per block)
__global__ __launch_bounds__(NT) void k_solve_uv(Batch b) {
  Site s = locate(b);
  if (!s.ok) return;
  if (s.lvl == 0) solve_u_at(s);
  else solve_v_at(s);
}

// ---------------------------------------------------------------------------
// K20-K23 velocity extrapolation (extrapolate/*.wgsl), fused: the validity
// mask of one velocity component (initialize_u_valid.wgsl /
// initialize_v_valid.wgsl: a face is valid (1) where the open fraction of the
// air level set averaged to its two end corners is 0) and its 10
// extrapolation passes (extrapolate_u.wgsl / extrapolate_v.wgsl, loop
// render_node.rs:540-608) in one launch.  A 64 x 16 tile of faces (1024
// threads) stages its faces plus a 10-face halo (one per pass) in LDS: the
// level set cells the mask needs, the validity, and - only if its passes can
// change anything - the velocities; pass p is exact on faces at least p + 1
// from the region border, so the tile's faces are exact after 10.  Each pass updates
// only faces inside the dispatch coverage [0, ex) x [0, ey), so faces in the
// reference's coverage holes keep their values, as there.  Within a pass the
// face values change in place: a pass writes only faces that were invalid
// and reads only faces that were valid.
//
// The launch writes its fields in place while later tiles stage halos, and
// what a late tile may stage is never consumed: the velocity of a face that
// some tile fills was invalid at the start, so no tile reads its staged
// value before computing it itself; valid faces are never written; the
// first validity plane is recomputed from the level set on covered faces and
// written unchanged elsewhere; the second is overwritten on covered faces by
// the first pass before any pass reads it.
// ---------------------------------------------------------------------------
constexpr int XP = 10;  // passes = halo
constexpr int XNT = 1024;  // threads per tile: a tile's passes are its critical path
constexpr int XT_W = 64, XT_H = 16;
constexpr int XR_W = XT_W + 2 * XP, XR_H = XT_H + 2 * XP, XR_N = XR_W * XR_H;
constexpr int XL_W = XR_W + 2, XL_N = XL_W * (XR_H + 2);

template <int COMP>
__device__ __forceinline__ void extrapolate_tile(const Batch& b, int i, int lb, float* sf, uint32_t* sv0, uint32_t* sv1,
                                                 float* sl) {
  const DevInst* I = b.inst[i];
  const int W = I->W, H = I->H;
  const int w = COMP == 0 ? W + 1 : W, h = COMP == 0 ? H : H + 1;
  const int ex = b.ex[i], ey = b.ey[i];
  const int ox = (lb % b.tiles_x[i]) * XT_W - XP, oy = (lb / b.tiles_x[i]) * XT_H - XP;
  float* F = COMP == 0 ? I->u0 : I->v0;
  int32_t* V0 = COMP == 0 ? I->uval0 : I->vval0;
  int32_t* V1 = COMP == 0 ? I->uval1 : I->vval1;
  // level-set cells [ox - 1, ox + XR_W + 1) x [oy - 1, oy + XR_H + 1), every
  // load issued before any is stored (a cell outside the grid reads index 0
  // and is stored as 0)
  {
    constexpr int LIT = (XL_N + XNT - 1) / XNT;
    float lv[LIT];
#pragma unroll
    for (int k = 0; k < LIT; ++k) {
      const int q = (int)threadIdx.x + k * XNT, x = ox - 1 + q % XL_W, y = oy - 1 + q / XL_W;
      lv[k] = I->ls0[q < XL_N && inb(x, y, W, H) ? (size_t)y * W + x : 0];
    }
#pragma unroll
    for (int k = 0; k < LIT; ++k) {
      const int q = (int)threadIdx.x + k * XNT, x = ox - 1 + q % XL_W, y = oy - 1 + q / XL_W;
      if (q < XL_N) sl[q] = inb(x, y, W, H) ? lv[k] : 0.0f;
    }
  }
  __syncthreads();
  // Validity of the region: recomputed on covered faces, staged (both
  // planes) elsewhere.  A covered face's second plane is written by pass 0
  // before any pass reads it, so it is not staged; velocities are staged only
  // by tiles whose passes run.
  bool any_valid = false, any_open = false;
  for (int q = threadIdx.x; q < XR_N; q += XNT) {
    const int lx = q % XR_W, ly = q / XR_W, x = ox + lx, y = oy + ly;
    uint32_t v0 = 0, v1 = 0;
    if (inb(x, y, w, h)) {
      const size_t id = (size_t)y * w + x;
      if (x < ex && y < ey) {  // the validity kernels cover the same faces
        const int c = (ly + 1) * XL_W + lx + 1;  // cell (x, y)
        float c0, c1, c2, c3, c4, c5;
        if (COMP == 0) {
          c0 = sl[c - XL_W - 1]; c1 = sl[c - XL_W]; c2 = sl[c - 1]; c3 = sl[c]; c4 = sl[c + XL_W - 1]; c5 = sl[c + XL_W];
        } else {
          c0 = sl[c - XL_W - 1]; c1 = sl[c - 1]; c2 = sl[c - XL_W]; c3 = sl[c]; c4 = sl[c - XL_W + 1]; c5 = sl[c + 1];
        }
        const float vm = 0.25f * (((c0 + c1) + c2) + c3), vp = 0.25f * (((c2 + c3) + c4) + c5);
        v0 = face_open(vm, vp) > 0.0f ? 0 : 1;
      } else {
        v0 = (uint32_t)V0[id];
        v1 = (uint32_t)V1[id];
      }
    }
    sv0[q] = v0;
    sv1[q] = v1;
    const bool covered = inb(x, y, w, h) && x < ex && y < ey;
    any_valid |= v0 == 1 || (v1 == 1 && !covered);  // uncovered faces keep both planes
    any_open |= v0 == 0 && covered;
  }
  const int cx1 = min(ex, w) - ox, cy1 = min(ey, h) - oy;  // coverage in local coordinates
  // The passes change nothing unless the region holds a valid face (in the
  // recomputed plane, or in either staged plane of a face outside the
  // coverage, which even and odd passes read) and a covered invalid one:
  // with no valid face no face can become valid, with no covered invalid
  // face none is written.  Then every covered face of the tile ends with
  // both planes equal to its recomputed mask and keeps its velocity.
  const int has_valid = __syncthreads_or(any_valid ? 1 : 0);
  const int has_open = __syncthreads_or(any_open ? 1 : 0);
  if (!(has_valid && has_open)) {
    for (int t = threadIdx.x; t < XT_W * XT_H; t += XNT) {
      const int lx = XP + t % XT_W, ly = XP + t / XT_W, x = ox + lx, y = oy + ly;
      if (!inb(x, y, w, h)) continue;
      const size_t id = (size_t)y * w + x;
      const int q = ly * XR_W + lx;
      V0[id] = sv0[q];
      if (lx < cx1 && ly < cy1) V1[id] = sv0[q];
    }
    return;
  }
  {
    constexpr int FIT = (XR_N + XNT - 1) / XNT;
    float fv[FIT];
#pragma unroll
    for (int k = 0; k < FIT; ++k) {
      const int q = (int)threadIdx.x + k * XNT, x = ox + q % XR_W, y = oy + q / XR_W;
      fv[k] = F[q < XR_N && inb(x, y, w, h) ? (size_t)y * w + x : 0];
    }
#pragma unroll
    for (int k = 0; k < FIT; ++k) {
      const int q = (int)threadIdx.x + k * XNT, x = ox + q % XR_W, y = oy + q / XR_W;
      if (q < XR_N) sf[q] = inb(x, y, w, h) ? fv[k] : 0.0f;
    }
  }
  __syncthreads();
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;   // 64 x 16 threads walk the window
  for (int p = 0; p < XP; ++p) {
    const uint32_t* vin = (p & 1) ? sv1 : sv0;  // pass 0: valid0 -> valid1, then back
    uint32_t* vout = (p & 1) ? sv0 : sv1;
    const int m = p + 1;
    const int lo_x = max(m, -ox), hi_x = min(XR_W - m, cx1), lo_y = max(m, -oy), hi_y = min(XR_H - m, cy1);
    for (int ly = lo_y + ty; ly < hi_y; ly += 16)
      for (int lx = lo_x + tx; lx < hi_x; lx += 64) {
        const int q = ly * XR_W + lx;
        if (vin[q] == 1) {
          vout[q] = 1;
          continue;
        }
        const int x = ox + lx, y = oy + ly;
        int count = 0;
        float nu = 0.0f;
        if (x + 1 < w && y < h && vin[q + 1] == 1) { nu += sf[q + 1]; count += 1; }
        if (x - 1 > 0 && y > 0 && vin[q - 1] == 1) { nu += sf[q - 1]; count += 1; }
        if (x < w && y + 1 < h && vin[q + XR_W] == 1) { nu += sf[q + XR_W]; count += 1; }
        if (x > 0 && y - 1 > 0 && vin[q - XR_W] == 1) { nu += sf[q - XR_W]; count += 1; }
        if (count > 0) {
          nu /= (float)count;
          sf[q] = nu;
          vout[q] = 1;
        } else {
          vout[q] = 0;
        }
      }
    __syncthreads();
  }
  for (int t = threadIdx.x; t < XT_W * XT_H; t += XNT) {
    const int lx = XP + t % XT_W, ly = XP + t / XT_W, x = ox + lx, y = oy + ly;
    if (!inb(x, y, w, h)) continue;
    const size_t id = (size_t)y * w + x;
    const int q = ly * XR_W + lx;
    F[id] = sf[q];
    V0[id] = sv0[q];
    V1[id] = sv1[q];
  }
}

// Batch::level of a slot names the component (0 = u, 1 = v): both run in
// one launch
__global__ __launch_bounds__(XNT) void k_extrapolate_fused(Batch b) {
  __shared__ float sf[XR_N];
  __shared__ uint32_t sv[2][XR_N];  // 32-bit: no two lanes write one LDS word
  __shared__ float sl[XL_N];
  const int blk = xcd_tile((int)blockIdx.x, b.nblocks);
  const int i = slot_of(b.start, b.n, blk);
  const int lb = blk - b.start[i];
  if (b.level[i] == 0) extrapolate_tile<0>(b, i, lb, sf, sv[0], sv[1], sl);
  else extrapolate_tile<1>(b, i, lb, sf, sv[0], sv[1], sl);
}

// ---------------------------------------------------------------------------
// K24 advect_levelset with CUBIC (advect_levelset.wgsl:11-74)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float catmull(float y0, float y1, float y2, float y3, float t) {
  float d1 = 0.5f * (y2 - y0), d2 = 0.5f * (y3 - y1);
  float a2 = ((-2.0f * d1) - d2) + 3.0f * (y2 - y1);
  float a3 = (d1 + d2) - 2.0f * (y2 - y1);
  return (((a3 * t * t * t) + (a2 * t * t)) + (d1 * t)) + y1;
}
__global__ __launch_bounds__(NT) void k_advect_levelset(Batch b) {
  Site s = locate(b);
  if (!s.ok) return;
  CDevInst* I = cdev(s.I);
  int W = I->W, H = I->H;
  float2 q = backtrace(I, (float)s.x, (float)s.y, I->dt);
  float nx = fminf(fmaxf(q.x, 0.0f), (float)W - 1.0f);
  float ny = fminf(fmaxf(q.y, 0.0f), (float)H - 1.0f);
  int bx = to_i32(nx), by = to_i32(ny);
  float tx = nx - (float)bx, ty = ny - (float)by;
  float r[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    int yy = by - 1 + k;
    r[k] = catmull(ldz(I->ls0, W, H, bx - 1, yy), ldz(I->ls0, W, H, bx, yy), ldz(I->ls0, W, H, bx + 1, yy),
                   ldz(I->ls0, W, H, bx + 2, yy), tx);
  }
  float nl = catmull(r[0], r[1], r[2], r[3], ty);
  if (fabsf(nl) < 1000.0f) I->ls1[(size_t)s.y * W + s.x] = nl;
}

// ---------------------------------------------------------------------------
// K25 - K27 jump flooding (reinitialize_levelset/shaders/jump_flooding/*.wgsl)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(NT) void k_jfa_init(Batch b) {
  Site s = locate(b);
  if (!s.ok) return;
  CDevInst* I = cdev(s.I);
  int W = I->W, H = I->H, x = s.x, y = s.y;
  float level = I->ls1[(size_t)y * W + x];
  float md = 10.0f;
  float2 seed = make_float2(-1.0f, -1.0f);
  for (int i = -1; i <= 1; ++i)
    for (int j = -1; j <= 1; ++j) {
      if (i == 0 && j == 0) continue;
      int nx = x + i, ny = y + j;
      if (!inb(nx, ny, W, H)) continue;
      float nl = I->ls1[(size_t)ny * W + nx];
      if (nl >= 0.0f && !(level >= 0.0f)) {
        float d = level / (level - nl);
        if (fabsf(d) < md) {
          md = fabsf(d);
          seed = make_float2((float)x + d * (float)i, (float)y + d * (float)j);
        }
      }
    }
  I->seed0[(size_t)y * W + x] = seed;
}
__device__ __forceinline__ bool seed_ok(float2 q) { return q.x != -1.0f && q.y != -1.0f; }
__global__ __launch_bounds__(NT) void k_jfa_step(Batch b, int step) {
  Site s = locate(b);
  if (!s.ok) return;
  CDevInst* I = cdev(s.I);
  int W = I->W, H = I->H, x = s.x, y = s.y;
  const int src_idx = s.lvl;  // per-instance ping-pong index
  const float2* si = src_idx ? I->seed1 : I->seed0;
  float2* so = src_idx ? I->seed0 : I->seed1;
  float2 cur = si[(size_t)y * W + x];
  float2 best = cur;
  float bd = 1e10f;
  if (seed_ok(cur)) bd = hyp(cur.x - (float)x, cur.y - (float)y);
  for (int i = -1; i <= 1; ++i)
    for (int j = -1; j <= 1; ++j) {
      if (i == 0 && j == 0) continue;
      int nx = x + i * step, ny = y + j * step;
      if (!inb(nx, ny, W, H)) continue;
      float2 q = si[(size_t)ny * W + nx];
      if (!seed_ok(q)) continue;
      float d = hyp(q.x - (float)x, q.y - (float)y);
      if (d < bd) { best = q; bd = d; }
    }
  so[(size_t)y * W + x] = best;
}
__global__ __launch_bounds__(NT) void k_jfa_sdf(Batch b) {
  Site s = locate(b);
  if (!s.ok) return;
  CDevInst* I = cdev(s.I);
  size_t id = (size_t)s.y * I->W + s.x;
  float2 q = (s.lvl ? I->seed1 : I->seed0)[id];
  float sdf = hyp(q.x - (float)s.x, q.y - (float)s.y);
  I->ls0[id] = sdf * (I->ls1[id] < 0.0f ? -1.0f : 1.0f);
}

// ---------------------------------------------------------------------------
// K28 - K30 fast iterative method (reinitialize_levelset/shaders/fast_iterative_method)
// Labels: 0 NONE, 1 SOURCE, 2 ACTIVE.  The update kernel of the reference
// races on neighbours; it runs here as two phases per iteration with every
// read in a phase seeing the state at the start of the phase:
//   A: ACTIVE cells solve their quadratic; converged ones are marked (3).
//   B: NONE cells next to a cell marked in A (first in -x, +x, -y, +y order)
//      solve on the phase-A values and become ACTIVE if that lowers |phi|.
// ---------------------------------------------------------------------------
constexpr uint32_t L_NONE = 0u, L_SOURCE = 1u, L_ACTIVE = 2u, L_CONVERGED = 3u;
constexpr float FIM_FAR = 1e30f;

// K28 initialisation (source cells keep their level, the others start at
// +-FIM_FAR) and the first active labels (the neighbours of source cells) in
// one launch.  A cell is a source when an in-grid neighbour's level set has
// the other sign or is zero.  A covered neighbour's label is its own source
// test, recomputed here from the level set with the same expression; an
// uncovered neighbour (outside the dispatch cover 8 floor(W / 8) x
// 8 floor(H / 8)) keeps the label it had, which no thread of this launch
// writes.  So the results are those of the initialisation launch followed
// by the labelling launch.
__global__ __launch_bounds__(NT) void k_fim_init_active(Batch b) {
  // the block's 64 x 4 tile: the level set over it and a two-cell ring in
  // LDS, then the SOURCE label of every cell of the tile and a one-cell ring
  // (computed for covered cells, the stored one for uncovered cells), then
  // each cell's own initialisation and label from LDS
  constexpr int SW = TX + 4, SH = TY + 4, SN = SW * SH, RW1 = TX + 2, RN1 = RW1 * (TY + 2);
  __shared__ float sls[SN];
  __shared__ uint32_t ssrc[RN1];
  Site s = locate(b);
  CDevInst* I = cdev(s.I);
  const int W = I->W, H = I->H;
  const int x0 = s.x - (int)(threadIdx.x % TX), y0 = s.y - (int)(threadIdx.x / TX);
  constexpr int SIT = (SN + NT - 1) / NT;
  float lv[SIT];
#pragma unroll
  for (int k = 0; k < SIT; ++k) {  // every load first (outside the grid: index 0, unused)
    const int q = (int)threadIdx.x + k * NT, x = x0 - 2 + q % SW, y = y0 - 2 + q / SW;
    lv[k] = I->ls1[q < SN && inb(x, y, W, H) ? (size_t)y * W + x : 0];
  }
#pragma unroll
  for (int k = 0; k < SIT; ++k) {
    const int q = (int)threadIdx.x + k * NT;
    if (q < SN) sls[q] = lv[k];
  }
  __syncthreads();
  const int ox[4] = {-1, 1, 0, 0}, oy[4] = {0, 0, -1, 1};
  for (int q = (int)threadIdx.x; q < RN1; q += NT) {
    const int rx = q % RW1, ry = q / RW1, x = x0 - 1 + rx, y = y0 - 1 + ry;
    uint32_t src = 0;
    if (inb(x, y, W, H)) {
      if (x < s.ex && y < s.ey) {  // covered: the source test
        const int c = (ry + 1) * SW + rx + 1;
        const float level = sls[c];
        bool t = false;
#pragma unroll
        for (int k = 0; k < 4; ++k)
          if (!t && inb(x + ox[k], y + oy[k], W, H) && level * sls[c + ox[k] + oy[k] * SW] <= 0.0f) t = true;
        src = t ? 1u : 0u;
      } else {  // outside the cover: the label it had
        src = I->lab0[(size_t)y * W + x] == L_SOURCE ? 1u : 0u;
      }
    }
    ssrc[q] = src;
  }
  __syncthreads();
  if (!s.ok) return;
  const int x = s.x, y = s.y;
  const size_t id = (size_t)y * W + x;
  const int r = (y - y0 + 1) * RW1 + (x - x0 + 1);
  const float level = sls[(y - y0 + 2) * SW + (x - x0 + 2)];
  const bool src = ssrc[r] != 0u;
  I->ls0[id] = src ? level : sgn(level) * FIM_FAR;
  I->lab0[id] = src ? L_SOURCE : L_NONE;
  if (src) {
    I->lab[id] = L_SOURCE;
    return;
  }
  uint32_t out = L_NONE;
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (inb(x + ox[k], y + oy[k], W, H) && ssrc[r + ox[k] + oy[k] * RW1] != 0u) out = L_ACTIVE;
  I->lab[id] = out;
}

// The iterations, temporally blocked: a 64 x 16 tile with a halo of 2 FK
// cells runs up to FK iterations (phases A and B, A converging a cell when
// it moved by at most 1e-6) in LDS.  Each phase reads only the previous
// phase's values, and a phase's exact region shrinks by one cell, so after
// FK iterations the tile itself holds what FK grid-wide rounds give.  Tiles
// read one buffer pair and write the other (ls0/lab <-> phi_tmp/lab_tmp,
// `src` selects), so no tile reads a halo another tile has already advanced;
// an even number of launches ends in ls0/lab.  Batch::level = this
// instance's iterations in this launch.  Tiles cover the whole grid so every
// cell reaches the output buffer; only cells inside the reference's dispatch
// cover (8 floor(W/8) x 8 floor(H/8)) are updated, the others pass their
// values through (phase A's copy of them is their current value and they
// never trigger).
constexpr int FK = 5, FH = 2 * FK;
constexpr int FT_W = 64, FT_H = 16;
constexpr int FNT = 1024;  // threads per tile: a tile's iterations are its critical path
constexpr int FR_W = FT_W + 2 * FH, FR_H = FT_H + 2 * FH, FR_N = FR_W * FR_H;

__device__ __forceinline__ float eikonal_lds(const float* sp, int c, bool l, bool r, bool u, bool d) {
  const float R2 = 1.41421354f;
  float xm = fminf(l ? fabsf(sp[c - 1]) : FIM_FAR, r ? fabsf(sp[c + 1]) : FIM_FAR);
  float ym = fminf(u ? fabsf(sp[c - FR_W]) : FIM_FAR, d ? fabsf(sp[c + FR_W]) : FIM_FAR);
  float dd = xm - ym;
  if (dd > R2) return ym + 1.0f;
  if (dd < -R2) return xm + 1.0f;
  return 0.5f * ((xm + ym) + sqrtf(2.0f - dd * dd));
}

__global__ __launch_bounds__(FNT) void k_fim_fused(Batch b, int src) {
  __shared__ float sp[2][FR_N];
  __shared__ uint32_t sl[2][FR_N];  // 32-bit: no two lanes write one LDS word
  const int blk = xcd_tile((int)blockIdx.x, b.nblocks);
  const int i = slot_of(b.start, b.n, blk);
  const int lb = blk - b.start[i];
  const DevInst* I = b.inst[i];
  const int iters = b.level[i];
  const int W = I->W, H = I->H, cw = (W / 8) * 8, ch = (H / 8) * 8;
  const int ox = (lb % b.tiles_x[i]) * FT_W - FH, oy = (lb / b.tiles_x[i]) * FT_H - FH;
  const float* pin = src ? I->phi_tmp : I->ls0;
  const uint32_t* lin = src ? I->lab_tmp : I->lab;
  float* pout = src ? I->ls0 : I->phi_tmp;
  uint32_t* lout = src ? I->lab : I->lab_tmp;
  // staging: every load of the thread's cells issued before any is stored,
  // so their latencies overlap (a cell outside the grid reads index 0 and
  // is stored as 0 / NONE)
  constexpr int FIT = (FR_N + FNT - 1) / FNT;
  bool active = false;
  float pv[FIT];
  uint32_t lv[FIT];
#pragma unroll
  for (int k = 0; k < FIT; ++k) {
    const int c = (int)threadIdx.x + k * FNT;
    const int x = ox + c % FR_W, y = oy + c / FR_W;
    const size_t id = c < FR_N && inb(x, y, W, H) ? (size_t)y * W + x : 0;
    pv[k] = pin[id];
    lv[k] = lin[id];
  }
#pragma unroll
  for (int k = 0; k < FIT; ++k) {
    const int c = (int)threadIdx.x + k * FNT;
    if (c >= FR_N) continue;
    const int x = ox + c % FR_W, y = oy + c / FR_W;
    const bool in = inb(x, y, W, H);
    const float p = in ? pv[k] : 0.0f;
    const uint32_t l = in ? lv[k] : L_NONE;
    sp[0][c] = p;
    sl[0][c] = l;
    active |= l == L_ACTIVE && x < cw && y < ch;
  }
  // Only an ACTIVE cell inside the cover starts a change (stored labels are
  // never CONVERGED): a region without one passes through unchanged, and the
  // tile is copied to the output pair as it is.
  const int n_it = __syncthreads_or(active ? 1 : 0) ? iters : 0;
  const int nb[4] = {-1, 1, -FR_W, FR_W};
  for (int it = 0; it < n_it; ++it) {
    const int ma = 2 * it + 1, mb = 2 * it + 2;
    for (int c = threadIdx.x; c < FR_N; c += FNT) {  // phase A
      const int cx = c % FR_W, cy = c / FR_W, x = ox + cx, y = oy + cy;
      if (cx < ma || cx >= FR_W - ma || cy < ma || cy >= FR_H - ma || !inb(x, y, W, H)) continue;
      const uint32_t l = sl[0][c];
      const float p = sp[0][c];
      if (l != L_ACTIVE || x >= cw || y >= ch) {
        sp[1][c] = p;
        sl[1][c] = l;
      } else {
        const float q = eikonal_lds(sp[0], c, x > 0, x < W - 1, y > 0, y < H - 1) * sgn(p);
        sp[1][c] = q;
        sl[1][c] = (fabsf(p - q) > 1e-6f) ? L_ACTIVE : L_CONVERGED;
      }
    }
    __syncthreads();
    for (int c = threadIdx.x; c < FR_N; c += FNT) {  // phase B
      const int cx = c % FR_W, cy = c / FR_W, x = ox + cx, y = oy + cy;
      if (cx < mb || cx >= FR_W - mb || cy < mb || cy >= FR_H - mb || !inb(x, y, W, H)) continue;
      if (x >= cw || y >= ch) {  // outside the dispatch cover: unchanged
        sp[0][c] = sp[1][c];
        continue;
      }
      const uint32_t l = sl[1][c];
      if (l == L_CONVERGED) {
        sl[0][c] = L_NONE;
        sp[0][c] = sp[1][c];
        continue;
      }
      if (l != L_NONE) {
        sp[0][c] = sp[1][c];
        continue;
      }
      const bool inn[4] = {x > 0, x < W - 1, y > 0, y < H - 1};
      int trig = -1;
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (trig < 0 && inn[k] && sl[1][c + nb[k]] == L_CONVERGED) trig = k;
      const float pa = sp[1][c];
      if (trig < 0) {
        sp[0][c] = pa;
        continue;
      }
      const float qn = eikonal_lds(sp[1], c, inn[0], inn[1], inn[2], inn[3]);
      if (fabsf(pa) > qn) {
        sp[0][c] = qn * sgn(sp[1][c + nb[trig]]);
        sl[0][c] = L_ACTIVE;
      } else {
        sp[0][c] = pa;
      }
    }
    __syncthreads();
  }
  for (int t = threadIdx.x; t < FT_W * FT_H; t += FNT) {
    const int c = (FH + t / FT_W) * FR_W + FH + t % FT_W;
    const int x = ox + FH + t % FT_W, y = oy + FH + t / FT_W;
    if (!inb(x, y, W, H)) continue;
    pout[(size_t)y * W + x] = sp[0][c];
    lout[(size_t)y * W + x] = sl[0][c];
  }
}

// ---------------------------------------------------------------------------
// K31 levelset_gradient (levelset_gradient.wgsl:4-21)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(NT) void k_ls_gradient(Batch b) {
  Site s = locate(b);
  if (!s.ok) return;
  CDevInst* I = cdev(s.I);
  int W = I->W, H = I->H, x = s.x, y = s.y;
  if (x == 0 || y == 0) return;
  float gx = 0.5f * (ldz(I->ls0, W, H, x + 1, y) - ldz(I->ls0, W, H, x - 1, y));
  float gy = 0.5f * (ldz(I->ls0, W, H, x, y + 1) - ldz(I->ls0, W, H, x, y - 1));
  I->grad[(size_t)y * W + x] = make_float2(gx, gy);
}

// ---------------------------------------------------------------------------
// K32 / K33 fluid -> solid forces (fluid_to_solid/*.wgsl)
// ---------------------------------------------------------------------------
// Bins are planar i32 fixed point: force x, force y, torque (the reference's
// three), plus a fourth, the torque about the centre of mass in the world
// frame (-rx * fy - ry * fx with f in grid axes), which has no counterpart and
// feeds the rigid-body integrator.
__device__ __forceinline__ int fixed_point(float v) { return to_i32(v * (1.0f / 0.1f)); }
// Centre of body k about which its torques are taken: the triangle's
// centroid, a circle's, rectangle's or capsule's translation, else the origin.
// The obstacle's first 96 bytes (ids, values, transform) are loaded as six
// 16-byte words before the shape selects the result, in one round.
static_assert(offsetof(efl_obstacle, values) == 8 && offsetof(efl_obstacle, transform) == 32 &&
                  sizeof(efl_obstacle) % 16 == 0,
              "efl_obstacle layout");
__device__ __forceinline__ float2 body_centre(CDevInst* I, int k) {
  float2 c = make_float2(0.0f, 0.0f);
  if ((uint32_t)k < I->nobs) {
    const float4* q = reinterpret_cast<const float4*>(&I->obs[k]);
    const float4 q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3], q4 = q[4], q5 = q[5];
    // keeps the compiler from sinking the triangle-only loads behind the test
    asm volatile("" ::"v"(q2.x), "v"(q2.y), "v"(q3.x), "v"(q3.y), "v"(q4.x), "v"(q4.y));
    const uint32_t shape = __float_as_uint(q0.y);
    // values: q0.z q0.w q1.x q1.y q1.z q1.w; transform: q2 (0-3), q3 (4-7), q4 (8-11), q5 (12-15)
    const float lx = ((q0.z + q1.x) + q1.z) / 3.0f, ly = ((q0.w + q1.y) + q1.w) / 3.0f;
    // xform(transform, lx, ly)
    const float tx = ((q2.x * lx + q3.x * ly) + q4.x * 0.0f) + q5.x * 1.0f;
    const float ty = ((q2.y * lx + q3.y * ly) + q4.y * 0.0f) + q5.y * 1.0f;
    c.x = shape == 4u ? tx : (shape <= 2u ? q5.x : 0.0f);
    c.y = shape == 4u ? ty : (shape <= 2u ? q5.y : 0.0f);
  }
  return c;
}
// A block sums a TX x SF_TH tile (batch tiles of that size).  Every load is
// issued before any result is needed: the tile's solid level set with a
// one-cell ring (zero outside the grid) into LDS, each thread's column of
// solid ids and pressures into registers, every body's centre into LDS.
// Then each thread sums its column's cells (runs of one body in registers),
// each wavefront its lanes, the block in LDS bins, and the block adds its
// nonzero bins to the instance's, one global atomic each.  The sums wrap as
// i32 (as atomicAdd) and are exact in any order.
__global__ __launch_bounds__(NT) void k_sample_forces(Batch b) {
  constexpr int RW = TX + 2, RH = SF_TH + 2, ROWS = SF_TH / TY;
  __shared__ int sbin[4 * EFL_MAX_SOLIDS];
  __shared__ float sls[RW * RH];
  __shared__ float2 scen[EFL_MAX_SOLIDS];
  const int blk = blockIdx.x;
  const int i = batch_slot(b, blk);
  CDevInst* I = cdev(b.inst[i]);
  const int lb = blk - b.start[i], tw = b.tiles_x[i], ex = b.ex[i], ey = b.ey[i];
  const int W = I->W, H = I->H;
  const int ox = (lb % tw) * TX, oy = (lb / tw) * SF_TH;
  const int tx = (int)(threadIdx.x % TX), ty = (int)(threadIdx.x / TX);
  const int x = ox + tx;
  const int32_t* sidp = I->sid;
  const float *p0p = I->p0, *lssp = I->lss;
  int sid[ROWS];
  float pr[ROWS];
#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
    const int y = oy + ty + r * TY;
    const bool in = x < ex && y < ey;
    const size_t id = (size_t)y * W + x;
    sid[r] = in ? sidp[id] : -1;
    pr[r] = in ? p0p[id] : 0.0f;
  }
  float stage[(RW * RH + NT - 1) / NT];
#pragma unroll
  for (int k = 0; k < (RW * RH + NT - 1) / NT; ++k) {
    const int t = (int)threadIdx.x + k * NT;
    stage[k] = t < RW * RH ? ldz(lssp, W, H, ox - 1 + t % RW, oy - 1 + t / RW) : 0.0f;
  }
  // the instance's scalars, loaded with the rest rather than behind the
  // barrier where they are first used
  const float dxv = I->dx, sx = I->sizex, sy = I->sizey;
  const float T[16] = {I->T[0], I->T[1], 0.f, 0.f, I->T[4], I->T[5], 0.f, 0.f,
                       I->T[8], I->T[9], 0.f, 0.f, I->T[12], I->T[13], 0.f, 0.f};
  asm volatile("" ::"s"(dxv), "s"(sx), "s"(sy), "s"(T[0]), "s"(T[1]), "s"(T[4]), "s"(T[5]), "s"(T[8]),
               "s"(T[9]), "s"(T[12]), "s"(T[13]));
  static_assert(EFL_MAX_SOLIDS == NT, "one body centre per thread");
  const float2 cen = body_centre(I, (int)threadIdx.x);
  for (int j = threadIdx.x; j < 4 * EFL_MAX_SOLIDS; j += NT) sbin[j] = 0;
  scen[threadIdx.x] = cen;
#pragma unroll
  for (int k = 0; k < (RW * RH + NT - 1) / NT; ++k) {
    const int t = (int)threadIdx.x + k * NT;
    if (t < RW * RH) sls[t] = stage[k];
  }
  __syncthreads();
  int cs = -1;
  unsigned acc[4] = {0u, 0u, 0u, 0u};
  // one copy of the cell body, the row's values picked from registers (an
  // eighth of the unrolled code)
#pragma unroll 1
  for (int r = 0; r < ROWS; ++r) {
    int cand = sid[0];
    float prr = pr[0];
#pragma unroll
    for (int k = 1; k < ROWS; ++k) {
      cand = r == k ? sid[k] : cand;
      prr = r == k ? pr[k] : prr;
    }
    // a cell without a solid id contributes nothing whatever its fractions
    if (cand < 0 || cand >= EFL_MAX_SOLIDS) continue;
    const int ly = ty + r * TY + 1, lx = tx + 1;
    const float* c = &sls[ly * RW + lx];
    // fractions_at from the staged level set (same sums, same order)
    float a0 = c[-RW - 1], a1 = c[-RW], a2 = c[-RW + 1];
    float a3 = c[-1], a4 = c[0], a5 = c[1];
    float a6 = c[RW - 1], a7 = c[RW], a8 = c[RW + 1];
    float mm = 0.25f * (((a0 + a1) + a3) + a4);
    float pm = 0.25f * (((a1 + a2) + a4) + a5);
    float mp = 0.25f * (((a3 + a4) + a6) + a7);
    float pp = 0.25f * (((a4 + a5) + a7) + a8);
    // face_open(a, b) is 0 when max(a, b) <= 0 (a negative or zero quotient,
    // or 0 / 0), so a cell inside a body is closed without the divisions
    if (mm <= 0.0f && pm <= 0.0f && mp <= 0.0f && pp <= 0.0f) continue;
    const float4 f = make_float4(face_open(mm, mp), face_open(pm, pp), face_open(mm, pm), face_open(mp, pp));
    if (all_closed(f)) continue;
    const int y = oy + ty + r * TY;
    float fx = ((-(f.y - f.x)) * prr) * dxv;
    float fy = ((-(f.w - f.z)) * prr) * dxv;
    const float2 cc = scen[cand];
    float ux = (float)x / sx, uy = (float)y / sy;
    float2 w = xform(T, (ux - 0.5f) * sx, (-uy + 0.5f) * sy);
    float rx = w.x - cc.x, ry = w.y - cc.y;
    if (cand != cs) {
      if (cs >= 0)
#pragma unroll
        for (int k = 0; k < 4; ++k) atomicAdd(&sbin[k * EFL_MAX_SOLIDS + cs], (int)acc[k]);
      cs = cand;
#pragma unroll
      for (int k = 0; k < 4; ++k) acc[k] = 0u;
    }
    acc[0] += (unsigned)fixed_point(fx);
    acc[1] += (unsigned)fixed_point(fy);
    acc[2] += (unsigned)fixed_point(rx * fy - ry * fx);
    acc[3] += (unsigned)fixed_point(-rx * fy - ry * fx);
  }
  // when every holding lane of the wavefront holds the same body, fold the
  // wave first and add once
  const int lane = threadIdx.x & 63;
  const unsigned long long act = __ballot(cs >= 0);
  if (act != 0ull) {
    const int first = __ffsll((long long)act) - 1;
    const int cs0 = __shfl(cs, first, 64);
    if (__all(cs < 0 || cs == cs0)) {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        unsigned v = acc[k];
        for (int off = 32; off > 0; off >>= 1) v += (unsigned)__shfl_xor((int)v, off, 64);
        acc[k] = v;
      }
      if (lane == first)
#pragma unroll
        for (int k = 0; k < 4; ++k) atomicAdd(&sbin[k * EFL_MAX_SOLIDS + cs0], (int)acc[k]);
    } else if (cs >= 0) {
#pragma unroll
      for (int k = 0; k < 4; ++k) atomicAdd(&sbin[k * EFL_MAX_SOLIDS + cs], (int)acc[k]);
    }
  }
  __syncthreads();
  for (int j = threadIdx.x; j < 4 * EFL_MAX_SOLIDS; j += NT) {
    const int v = sbin[j];
    if (v != 0) atomicAdd(&I->bins[j], v);
  }
}
// The sums become forces and torques (fts) and the last pass's bins, and the
// accumulators return to zero for the next pass (instances start zeroed).
__global__ void k_accumulate_forces(Batch b) {
  const DevInst* I = b.inst[blockIdx.x];
  for (int i = threadIdx.x; i < EFL_MAX_SOLIDS; i += blockDim.x) {
    int v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      v[k] = I->bins[k * EFL_MAX_SOLIDS + i];
      I->bins_last[k * EFL_MAX_SOLIDS + i] = v[k];
      I->bins[k * EFL_MAX_SOLIDS + i] = 0;
    }
    I->fts[i] = make_float4((float)v[0] * 0.1f, (float)v[1] * 0.1f, (float)v[2] * 0.1f, (float)v[3] * 0.1f);
  }
}

// ---------------------------------------------------------------------------
// Reductions (deterministic: a fixed tree per block, partials folded in
// index order by fold_partials, efl_reduce.h).
// ---------------------------------------------------------------------------
template <class DI>
__device__ __forceinline__ bool active_at(DI I, size_t id) {
  return I->ls0[id] < 0.0f && !all_closed(I->afs[id]);
}

// Block partial sums: wavefront shuffles, then the 4 wave results in a
// fixed order; thread 0 stores component k at red[(c0 + k) * red_stride + lb].
template <int N, class DI>
__device__ __forceinline__ void block_partials(DI I, int c0, int lb, double (&v)[N], bool is_max = false) {
  __shared__ double sh[N][NT / 64];
#pragma unroll
  for (int k = 0; k < N; ++k) {
    double x = v[k];
    for (int off = 32; off > 0; off >>= 1) {
      double o = __shfl_xor(x, off, 64);
      x = is_max ? fmax(x, o) : x + o;
    }
    if ((threadIdx.x & 63) == 0) sh[k][threadIdx.x >> 6] = x;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < N; ++k) {
      double acc = sh[k][0];
      for (int w = 1; w < NT / 64; ++w) acc = is_max ? fmax(acc, sh[k][w]) : acc + sh[k][w];
      I->red[(size_t)(c0 + k) * I->red_stride + lb] = acc;
    }
  }
}

__device__ __forceinline__ int inst_blocks(const Batch& b, int slot) { return b.start[slot + 1] - b.start[slot]; }
__device__ __forceinline__ int local_block(const Batch& b, int slot) { return (int)blockIdx.x - b.start[slot]; }

// The CG kernels work on 64 x 16 tiles (one partial per tile): thread t
// takes column t % 64 of rows t / 64 + 4 r, r = 0..3.
constexpr int CG_TH = 16, CG_ROWS = CG_TH / TY;
static_assert(TX == kRedTileW && CG_TH == kRedTileH, "CG partials are counted in DevInst::cg_tiles");
struct CSite {
  const DevInst* I;
  int x, y0, ex, ey, slot, lb;
};
__device__ __forceinline__ CSite locate_cg(const Batch& b) {
  const int blk = blockIdx.x;
  int st = b.start[0], tw = b.tiles_x[0], ex = b.ex[0], ey = b.ey[0];
  const DevInst* I = b.inst[0];
  const int i = batch_slot(b, blk);
  if (i > 0) {
    st = b.start[i];
    tw = b.tiles_x[i];
    ex = b.ex[i];
    ey = b.ey[i];
    I = b.inst[i];
  }
  CSite s;
  s.lb = blk - st;
  s.I = I;
  s.x = (s.lb % tw) * TX + (int)(threadIdx.x % TX);
  s.y0 = (s.lb / tw) * CG_TH + (int)(threadIdx.x / TX);
  s.ex = ex;
  s.ey = ey;
  s.slot = i;
  return s;
}

// rr = ||r||^2 and bb = ||b||^2 over active cells (partials RED_RR, RED_BB),
// with r = b - A x for the finest level (x = p0, b = div); with pcg_init, r
// also goes to pcg_r.
__global__ __launch_bounds__(NT) void k_residual_norm(Batch b, int pcg_init) {
  const CSite s = locate_cg(b);
  CDevInst* I = cdev(s.I);
  double v[2] = {0.0, 0.0};
  // Level 0 is x = p0, b = div, ls = ls0, af = afs.  Every load of the
  // thread's rows (the cell, its four neighbours' level set and x) is issued
  // straight-line before any is used (a cell outside reads index 0); then
  // residual_value's expression, in its order, on the loaded values.
  const Level L = lev_copy(I->lev[0]);
  float* const rout = I->pcg_r;
  const float dx = I->dx * 1.0f;
  const float factor = I->dt / I->rho / dx / dx;
  bool in[CG_ROWS], nin[CG_ROWS][4];
  uint32_t id[CG_ROWS];
  float phi[CG_ROWS], xc[CG_ROWS], bv[CG_ROWS], pn[CG_ROWS][4], xn[CG_ROWS][4];
  float4 af[CG_ROWS];
  const int ox[4] = {-1, 1, 0, 0}, oy[4] = {0, 0, -1, 1};
#pragma unroll
  for (int r = 0; r < CG_ROWS; ++r) {
    const int y = s.y0 + TY * r;
    in[r] = s.x < s.ex && y < s.ey;
    id[r] = in[r] ? (uint32_t)(y * L.w + s.x) : 0u;
    phi[r] = L.ls[id[r]];
    af[r] = L.af[id[r]];
    xc[r] = L.x[id[r]];
    bv[r] = L.b[id[r]];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int jx = s.x + ox[k], jy = y + oy[k];
      nin[r][k] = in[r] && inb(jx, jy, L.w, L.h);
      const uint32_t jd = nin[r][k] ? (uint32_t)(jy * L.w + jx) : 0u;
      pn[r][k] = L.ls[jd];
      xn[r][k] = L.x[jd];
    }
  }
#pragma unroll
  for (int r = 0; r < CG_ROWS; ++r) {
    if (!in[r]) continue;
    const bool act = phi[r] < 0.0f && !all_closed(af[r]);  // active_at()
    float rv = 0.0f;
    if (act) {  // residual_value()
      float res = bv[r];
      const float fv[4] = {af[r].x, af[r].y, af[r].z, af[r].w};
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        if (!nin[r][k]) continue;
        if (pn[r][k] < 0.0f) res -= fv[k] * (xc[r] - xn[r][k]) * factor;
        else res -= fv[k] * (1.0f - pn[r][k] / phi[r]) * xc[r] * factor;
      }
      rv = res;
    }
    if (pcg_init) rout[id[r]] = rv;
    const float b2 = act ? bv[r] : 0.0f;
    v[0] += (double)rv * (double)rv;
    v[1] += (double)b2 * (double)b2;
  }
  block_partials<2>(I, RED_RR, s.lb, v);
}

// One block per instance folds the partials of the preceding launch (same
// Batch):  kind 0: rr, bb -> scal[SC_RR], scal[SC_BB];  kind 4: max |D| ->
// diag[0].
__global__ __launch_bounds__(NT) void k_finalize(Batch b, int kind) {
  const int slot = blockIdx.x;
  const DevInst* I = b.inst[slot];
  const int nblk = inst_blocks(b, slot);
  if (kind == 0) {
    double t[2];
    fold_partials<2>(I, RED_RR, nblk, false, t);
    if (threadIdx.x == 0) {
      I->scal[SC_RR] = t[0];
      I->scal[SC_BB] = t[1];
    }
  } else {
    double t[1];
    fold_partials<1>(I, RED_PQ, nblk, true, t);
    if (threadIdx.x == 0) I->diag[0] = (float)t[0];
  }
}

// The CG stopping test (init: of the initial residual, whose launch also
// left ||b||^2; else cg_stop() of the iteration k_pcg_axpy just ran): folds
// ||r||^2 and records the outcome.  One block per instance; the tiled
// preconditioner runs the same test in its first launch instead.
__global__ __launch_bounds__(NT) void k_cg_check(Batch b, int init) {
  const int slot = blockIdx.x;
  const DevInst* I = b.inst[slot];
  if (!init && cg_done(I)) return;
  const int nblk = inst_blocks(b, slot);
  double* sc = I->scal;
  if (init) {
    double t[2];
    fold_partials<2>(I, RED_RR, nblk, false, t);
    if (threadIdx.x != 0) return;
    sc[SC_RR] = t[0];
    sc[SC_BB] = t[1];
    sc[SC_ITER] = 0.0;
    const double rel = t[1] > 0.0 ? sqrt(t[0] / t[1]) : sqrt(t[0]);
    sc[SC_REL] = rel;
    const bool ok = I->cg_tol > 0.0f && rel <= (double)I->cg_tol;
    sc[SC_CONV] = (I->cg_tol > 0.0f && !ok) ? 0.0 : 1.0;  // until a later test records one
    sc[SC_DONE] = ok ? 1.0 : 0.0;
    return;
  }
  double t[1];
  fold_partials<1>(I, RED_RR, nblk, false, t);
  if (threadIdx.x == 0) cg_stop(I, t[0], true);
}

// After an asynchronous solve's iteration budget (one block per instance):
// count the solve in scal[SC_MISSED] when it has not met its tolerance, i.e.
// it has not stopped (the budget ran out) or stopped without converging.
__global__ void k_cg_tally(Batch b) {
  const DevInst* I = b.inst[blockIdx.x];
  if (threadIdx.x != 0) return;
  double* sc = I->scal;
  if (!(sc[SC_DONE] != 0.0 && sc[SC_CONV] != 0.0)) sc[SC_MISSED] = sc[SC_MISSED] + 1.0;
}

// One block per instance with DevInst::prefold folds the partials its CG
// consumers would otherwise each fold for themselves (the same fold_partials,
// so the scalar is the same either way): kind 0, after a preconditioner:
// r . z (and z . q_prev for flexible CG) -> scal[SC_RZF..];  kind 1, after
// k_pcg_apply: p . q over the Batch's tiles -> scal[SC_PQF].
// The host's view of a batch's scalars: the first 16 of each instance,
// packed, so that one copy brings them all back.
__global__ void k_gather_scalars(Batch b, double* dst) {
  const DevInst* I = b.inst[blockIdx.x];
  if (threadIdx.x < 16) dst[16 * blockIdx.x + threadIdx.x] = I->scal[threadIdx.x];
}

__global__ __launch_bounds__(NT) void k_prefold(Batch b, int kind) {
  const int slot = blockIdx.x;
  const DevInst* I = b.inst[slot];
  if (!I->prefold || cg_done(I)) return;
  if (kind == 0) {
    // r . z and the next component: z . q_prev of the flexible
    // preconditioner, or z . A z of the multigrid one (the fused CG
    // iteration's delta; unused by an unfused solve)
    double t[2] = {0.0, 0.0};
    fold_partials<2>(I, RED_RZ, (int)I->scal[SC_NRZ], false, t);
    if (threadIdx.x == 0) {
      I->scal[SC_RZF] = t[0];
      I->scal[SC_RZF + 1] = t[1];
    }
  } else {
    double t[1];
    fold_partials<1>(I, RED_PQ, inst_blocks(b, slot), false, t);
    if (threadIdx.x == 0) I->scal[SC_PQF] = t[0];
  }
}

// z masked to active cells, partials of r . z (RED_RZ)
__global__ __launch_bounds__(NT) void k_pcg_rz(Batch b) {
  const CSite s = locate_cg(b);
  CDevInst* I = cdev(s.I);
  if (cg_done(I)) return;
  double v[1] = {0.0};
  for (int r = 0; r < CG_ROWS; ++r) {
    const int y = s.y0 + TY * r;
    if (s.x >= s.ex || y >= s.ey) continue;
    size_t id = (size_t)y * I->W + s.x;
    float z = I->pcg_z[id];
    if (!active_at(I, id)) { z = 0.0f; I->pcg_z[id] = 0.0f; }
    v[0] += (double)I->pcg_r[id] * (double)z;
  }
  block_partials<1>(I, RED_RZ, s.lb, v);
  if (s.lb == 0 && threadIdx.x == 0) I->scal[SC_NRZ] = (double)inst_blocks(b, s.slot);
}

// Iteration iter = scal[SC_ITER] + 1, first half: every block folds the
// r . z of the preconditioner that just ran, beta = rz / rz_prev (0 for
// iter 1), and forms the search direction p = z + beta p_prev on the fly
// (written to pcg_p[iter & 1], recomputed at the neighbours it needs);
// q = A p (the operator of residual.wgsl with b = 0, negated) and partials
// of p . q.
__global__ __launch_bounds__(NT) void k_pcg_apply(Batch b) {
  // p = z + beta p_old over the tile and a one-cell ring, formed once per
  // cell; the operator reads its neighbours here
  constexpr int RW = TX + 2, RH = CG_TH + 2, RN = RW * RH, RIT = (RN + NT - 1) / NT;
  __shared__ float sp[RN];
  const CSite s = locate_cg(b);
  const DevInst* I = s.I;
  CDevInst* CI = cdev(I);  // descriptor fields in scalar loads, read once
  double* const scal = CI->scal;
  // Every load this block needs is issued before the r . z fold and the
  // done test, so their latencies overlap: the done flag and the iteration
  // (scalar), the flags of the thread's cells, z and p_old over the region,
  // then the stencils of the active cells (which the flags select).
  const double done_flag = sld(scal + SC_DONE);
  const int iter = (int)sld(scal + SC_ITER) + 1;
  const Level L = lev_copy(CI->lev[0]);
  // the first iteration takes p = z without reading the previous direction,
  // which a solve that diverged before may have left non-finite (0 * inf)
  const float* pold = iter > 1 ? ((iter - 1) & 1 ? CI->pcg_p[1] : CI->pcg_p[0]) : nullptr;
  float* pnew = iter & 1 ? CI->pcg_p[1] : CI->pcg_p[0];
  const float* z = CI->pcg_z;
  const int lx = (int)threadIdx.x % TX, ly0 = (int)threadIdx.x / TX;
  const int tx0 = s.x - lx, ty0 = s.y0 - ly0;
  int fl[CG_ROWS];
  size_t id[CG_ROWS];
  bool in[CG_ROWS];
#pragma unroll
  for (int r = 0; r < CG_ROWS; ++r) {
    const int y = s.y0 + TY * r;
    in[r] = s.x < s.ex && y < s.ey;
    id[r] = in[r] ? (size_t)y * L.w + s.x : 0;
    fl[r] = in[r] ? L.fl[id[r]] : 0;
  }
  float zr[RIT], pr[RIT];
#pragma unroll
  for (int k = 0; k < RIT; ++k) {
    const int u = (int)threadIdx.x + k * NT;
    const int gx = tx0 - 1 + u % RW, gy = ty0 - 1 + u / RW;
    const bool inside = u < RN && gx >= 0 && gy >= 0 && gx < L.w && gy < L.h;
    const size_t g = inside ? (size_t)gy * L.w + gx : 0;
    zr[k] = inside ? z[g] : 0.0f;
    pr[k] = inside && pold ? pold[g] : 0.0f;
  }
  float4 cf[CG_ROWS];
#pragma unroll
  for (int r = 0; r < CG_ROWS; ++r) {
    const bool act = (fl[r] & MG_LSNEG) && !(fl[r] & MG_CLOSED);  // active_at()
    cf[r] = act && !(fl[r] & MG_UNIT) ? L.cf[id[r]] : mg_cf_implied(fl[r]);
  }
  // earlier launches' scalars this launch does not write
  const bool flex = CI->cg_flex != 0;
  const double pq_prev = sld(scal + SC_PQ);
  const double rz_prev = sld(scal + SC_RZ0 + ((iter - 2) & 1));
  double rz[2] = {0.0, 0.0};
  double beta = 0.0;
  if (CI->prefold) {
    rz[0] = sld(scal + SC_RZF);
    rz[1] = sld(scal + SC_RZF + 1);
  } else if (flex) {
    fold_partials<2>(I, RED_RZ, (int)sld(scal + SC_NRZ), false, rz);
  } else {  // r . z of k_mg_up or k_pcg_rz: one partial per 64 x 16 tile
    double t[1];
    fold_partials<1>(I, RED_RZ, CI->cg_tiles, false, t);
    rz[0] = t[0];
  }
  if (done_flag != 0.0) return;  // uniform for the block
  if (flex) {  // flexible CG: beta = -(z . q_prev) / (p . q)_prev
    if (iter > 1 && pq_prev != 0.0) beta = -rz[1] / pq_prev;
  } else {
    if (iter > 1) beta = rz[0] / rz_prev;
  }
  if (s.lb == 0 && threadIdx.x == 0) {
    scal[SC_RZ0 + ((iter - 1) & 1)] = rz[0];
    scal[SC_BETA] = beta;
  }
  const float fb = (float)beta;
#pragma unroll
  for (int k = 0; k < RIT; ++k) {
    const int u = (int)threadIdx.x + k * NT;
    if (u < RN) sp[u] = pold ? zr[k] + fb * pr[k] : zr[k];
  }
  __syncthreads();
  const float factor = CI->dt / CI->rho / CI->dx / CI->dx;
  float* const qout = CI->pcg_q;
  double v[1] = {0.0};
#pragma unroll
  for (int r = 0; r < CG_ROWS; ++r) {
    if (!in[r]) continue;
    const int q0 = (ly0 + TY * r + 1) * RW + lx + 1;
    const float pc = sp[q0];
    pnew[id[r]] = pc;
    // residual.wgsl's operator from the precomputed stencil (k_mg_setup):
    // fluid neighbours c_k (p - p_k), air neighbours c_k p, times factor
    float q = 0.0f;
    if ((fl[r] & MG_LSNEG) && !(fl[r] & MG_CLOSED)) {
      const float cv[4] = {cf[r].x, cf[r].y, cf[r].z, cf[r].w};
      const int off[4] = {-1, 1, -RW, RW};
      float res = 0.0f;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        if (!((fl[r] >> k) & 1)) continue;
        if ((fl[r] >> (4 + k)) & 1) {
          res = res - cv[k] * (pc - sp[q0 + off[k]]) * factor;
        } else {
          res = res - cv[k] * pc * factor;
        }
      }
      q = -res;
    }
    qout[id[r]] = q;
    v[0] += (double)pc * (double)q;
  }
  block_partials<1>(I, RED_PQ, s.lb, v);
}

// Second half of the same iteration: every block folds p . q itself,
// alpha = rz / pq; x += alpha p, r -= alpha q, partials of ||r||^2 (RED_RR).
// A breakdown (p . q not positive, e.g. a non-finite state) ends the solve
// with the iterate and iteration count of the previous iteration; every
// block sees the same p . q, so all of them return.
__global__ __launch_bounds__(NT) void k_pcg_axpy(Batch b) {
  const CSite s = locate_cg(b);
  CDevInst* I = cdev(s.I);
  if (cg_done(I)) return;
  const int iter = (int)I->scal[SC_ITER] + 1;
  double t[1];
  if (I->prefold) t[0] = I->scal[SC_PQF];
  else fold_partials<1>(I, RED_PQ, inst_blocks(b, s.slot), false, t);
  if (!(t[0] > 0.0)) {  // breakdown (p . q <= 0 or not a number): stop before the update
    if (s.lb == 0 && threadIdx.x == 0) I->scal[SC_DONE] = 1.0;
    return;
  }
  const double alpha = I->scal[SC_RZ0 + ((iter - 1) & 1)] / t[0];
  if (s.lb == 0 && threadIdx.x == 0) {
    I->scal[SC_PQ] = t[0];
    I->scal[SC_ALPHA] = alpha;
    I->scal[SC_AL0 + ((iter - 1) & 1)] = alpha;  // what the fused form of the next iteration reads
  }
  if (s.lb == 0 && threadIdx.x == 0) {
    I->scal[SC_KCUR] = (double)iter;
    I->scal[SC_NRR] = (double)inst_blocks(b, s.slot);
  }
  const float fa = (float)alpha;
  const float* p = I->pcg_p[iter & 1];
  double v[1] = {0.0};
  for (int r = 0; r < CG_ROWS; ++r) {
    const int y = s.y0 + TY * r;
    if (s.x >= s.ex || y >= s.ey) continue;
    size_t id = (size_t)y * I->W + s.x;
    I->p0[id] = I->p0[id] + fa * p[id];
    float rv = I->pcg_r[id] - fa * I->pcg_q[id];
    I->pcg_r[id] = rv;
    v[0] += (double)rv * (double)rv;
  }
  block_partials<1>(I, RED_RR, s.lb, v);
}

// the CG initial guess (the previous pressure) restricted to the unknowns
__global__ __launch_bounds__(NT) void k_mask_p(Batch b) {
  const CSite s = locate_cg(b);
  CDevInst* I = cdev(s.I);
  for (int r = 0; r < CG_ROWS; ++r) {
    const int y = s.y0 + TY * r;
    if (s.x >= s.ex || y >= s.ey) continue;
    size_t id = (size_t)y * I->W + s.x;
    if (!active_at(I, id)) I->p0[id] = 0.0f;
  }
}

// partial max over active cells of |D| (the K10 numerator) on u0/v0
__global__ __launch_bounds__(NT) void k_max_divergence(Batch b) {
  Site s = locate(b);
  CDevInst* I = cdev(s.I);
  double v[1] = {0.0};
  if (s.ok) {
    int W = I->W, H = I->H, x = s.x, y = s.y;
    size_t id = (size_t)y * W + x;
    if (I->ls0[id] < 0.0f) {
      float4 f = fractions_at(I->lss, W, H, x, y);
      if (!all_closed(f)) {
        float du_f = f.y * ldz(I->u0, W + 1, H, x + 1, y) - f.x * ldz(I->u0, W + 1, H, x, y);
        float dv_f = f.w * ldz(I->v0, W, H + 1, x, y + 1) - f.z * ldz(I->v0, W, H + 1, x, y);
        float du_s = (1.0f - f.y) * ldz(I->us, W + 1, H, x + 1, y) - (1.0f - f.x) * ldz(I->us, W + 1, H, x, y);
        float dv_s = (1.0f - f.w) * ldz(I->vs, W, H + 1, x, y + 1) - (1.0f - f.z) * ldz(I->vs, W, H + 1, x, y);
        v[0] = fabsf(((du_f + dv_f) + du_s) + dv_s);
      }
    }
  }
  block_partials<1>(I, RED_PQ, local_block(b, s.slot), v, true);
}

// ---------------------------------------------------------------------------
// Fluid diagnostics (diagnostics/calculate_volume.wgsl:12-75,
// max_velocity.wgsl:13-81, min_velocity.wgsl:13-87): the volume (cells with
// ls0 < 0) and the largest and smallest cell-centred speed (the smallest over
// fluid cells), one 16 x 16 block per workgroup over the floor(W/16) x
// floor(H/16) cover as the reference dispatches them.  Counting, max and min
// are order independent, so the block folds use wavefront reductions.  The
// final pass keeps the reference's second stage: lane l of 512 adds the
// block values l, l + 512, ... in order (the speed passes add their block
// maxima and minima too), then the lanes fold by halving strides 256..1.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_fluid_diag_blocks(FluidDiagArgs a) {
  const int W = a.W, gx = W / 16;
  const int x = (int)blockIdx.x * 16 + (int)threadIdx.x % 16, y = (int)blockIdx.y * 16 + (int)threadIdx.x / 16;
  const float level = a.ls[(size_t)y * W + x];
  const float uc = 0.5f * (a.u[(size_t)y * (W + 1) + x] + a.u[(size_t)y * (W + 1) + x + 1]);
  const float vc = 0.5f * (a.v[(size_t)y * W + x] + a.v[(size_t)(y + 1) * W + x]);
  const float mag = sqrtf(uc * uc + vc * vc);
  float mx = mag, mn = level < 0.0f ? mag : 1e30f;
  const int cnt = __syncthreads_count(level < 0.0f);
  for (int off = 32; off > 0; off >>= 1) {
    mx = fmaxf(mx, __shfl_xor(mx, off, 64));
    mn = fminf(mn, __shfl_xor(mn, off, 64));
  }
  __shared__ float wmx[4], wmn[4];
  if ((threadIdx.x & 63) == 0) {
    wmx[threadIdx.x >> 6] = mx;
    wmn[threadIdx.x >> 6] = mn;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  const size_t np = (size_t)gx * (a.H / 16), wid = blockIdx.x + (size_t)blockIdx.y * gx;
  a.part[wid] = (float)cnt;
  a.part[np + wid] = fmaxf(fmaxf(wmx[0], wmx[1]), fmaxf(wmx[2], wmx[3]));
  a.part[2 * np + wid] = fminf(fminf(wmn[0], wmn[1]), fminf(wmn[2], wmn[3]));
}
__global__ __launch_bounds__(512) void k_fluid_diag_final(FluidDiagArgs a) {
  const int np = (a.W / 16) * (a.H / 16), l = (int)threadIdx.x;
  __shared__ float tv[512], tx[512], tn[512];
  float sv = 0.0f, sx = 0.0f, sn = 0.0f;
  for (int i = l; i < np; i += 512) {
    sv += a.part[i];
    sx += a.part[(size_t)np + i];
    sn += a.part[2 * (size_t)np + i];
  }
  tv[l] = sv;
  tx[l] = sx;
  tn[l] = sn;
  __syncthreads();
  for (int stride = 256; stride > 0; stride >>= 1) {
    if (l < stride) {
      tv[l] += tv[l + stride];
      tx[l] = fmaxf(tx[l], tx[l + stride]);
      tn[l] = fminf(tn[l], tn[l + stride]);
    }
    __syncthreads();
  }
  if (l == 0) {
    a.out[0] = tv[0];
    a.out[1] = tn[0];
    a.out[2] = tx[0];
  }
}

// ---------------------------------------------------------------------------
// Rigid bodies: semi-implicit Euler of the dynamic obstacles, one thread per
// body (the role avian2d plays for the reference).  The fluid forces of the
// contributing instances are summed; gravity, walls, fluid density and dx
// come from the first contributing instance.
// ---------------------------------------------------------------------------
__device__ float body_radius(const efl_obstacle& o) {
  switch (o.shape) {
    case 0u: return o.values[0];
    case 1u: return hyp(o.values[0], o.values[1]);
    case 2u: return fmaxf(hyp(o.values[0], o.values[1]), hyp(o.values[2], o.values[3])) + o.values[4];
    case 4u:
      return fmaxf(fmaxf(hyp(o.values[0], o.values[1]), hyp(o.values[2], o.values[3])), hyp(o.values[4], o.values[5]));
    default: return 0.0f;
  }
}

// One semi-implicit Euler step of body k of a list, given the fluid force
// (Fx, Fy) and torque Tq on it in grid units.  F = m a with the fluid force
// in cells (F_true = F dx, m_true = m dx^2) and an added-mass estimate
// (m_a = rho_fluid * area, previous acceleration) that keeps explicit
// coupling of light bodies stable.
__device__ void integrate_body(efl_obstacle* obs, const efl_body& bd, float4* state, int k, float Fx, float Fy,
                               float Tq, float dt, float gwx, float gwy, float dx0, float half_w, float half_h,
                               float rho_fluid) {
  efl_obstacle o = obs[k];
  float m = bd.density * bd.area;
  float J = bd.density * bd.inertia;
  float ma = 4.0f * rho_fluid * bd.area, Ja = 2.0f * rho_fluid * bd.inertia;
  float4 prev = state[k];
  float ax = (Fx + ma * prev.x + m * gwx / dx0) / (m + ma);
  float ay = (Fy + ma * prev.y + m * gwy / dx0) / (m + ma);
  float al = (J + Ja) > 0.0f ? (Tq + Ja * prev.z) / (J + Ja) : 0.0f;
  state[k] = make_float4(ax, ay, al, 0.0f);
  float vx = o.linear_velocity[0] + ax * dt;
  float vy = o.linear_velocity[1] + ay * dt;
  float w = o.angular_velocity + al * dt;
  // a body moves at most one cell per step (translation, and rotation at its
  // bounding radius): the area fractions and solid face velocities of the
  // next step assume it, and faster bodies trap and squeeze fluid
  float R = body_radius(o);
  const float vmax = 1.0f / dt;
  const float sp = sqrtf(vx * vx + vy * vy);
  if (sp > vmax) {
    const float sc = vmax / sp;
    vx = vx * sc;
    vy = vy * sc;
  }
  if (R > 0.0f && fabsf(w) * R > vmax) w = copysignf(vmax / R, w);
  float tx = o.transform[12] + vx * dt;
  float ty = o.transform[13] + vy * dt;
  float ang = atan2f(o.transform[1], o.transform[0]) + w * dt;
  float lx = fmaxf(half_w - 1.0f - R, 0.0f), ly = fmaxf(half_h - 1.0f - R, 0.0f);
  if (tx > lx) { tx = lx; vx = fminf(vx, 0.0f); }
  if (tx < -lx) { tx = -lx; vx = fmaxf(vx, 0.0f); }
  if (ty > ly) { ty = ly; vy = fminf(vy, 0.0f); }
  if (ty < -ly) { ty = -ly; vy = fmaxf(vy, 0.0f); }
  float c = cosf(ang), sn = sinf(ang);
  float T[16] = {c, sn, 0.f, 0.f, -sn, c, 0.f, 0.f, 0.f, 0.f, 1.f, 0.f, tx, ty, 0.f, 1.f};
  float Ti[16] = {c, -sn, 0.f, 0.f, sn, c, 0.f, 0.f, 0.f, 0.f, 1.f, 0.f, -(c * tx + sn * ty), -(-sn * tx + c * ty), 0.f, 1.f};
  for (int q = 0; q < 16; ++q) {
    obs[k].transform[q] = T[q];
    obs[k].inverse_transform[q] = Ti[q];
  }
  obs[k].linear_velocity[0] = vx;
  obs[k].linear_velocity[1] = vy;
  obs[k].angular_velocity = w;
}

// The context's body list, driven by the forces of every instance that
// shares it.
__global__ void k_integrate_bodies(BodyStep a) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.n) return;
  const efl_body bd = a.bodies[k];
  if (!bd.dynamic) return;
  float Fx = 0.0f, Fy = 0.0f, Tq = 0.0f;
  uint32_t id = a.obs[k].entity_id;
  if (id < (uint32_t)EFL_MAX_SOLIDS) {
    for (int i = 0; i < a.ninst; ++i) {
      float4 f = a.fts[i][id];
      float s = 1.0f / (a.dx[i] * a.dx[i]);
      Fx += f.x * s;
      Fy += -f.y * s;  // grid y points down
      Tq += f.w * s;   // world-frame torque bin
    }
  }
  integrate_body(a.obs, bd, a.state, k, Fx, Fy, Tq, a.dt, a.gwx, a.gwy, a.dx[0], a.half_w, a.half_h, a.rho_fluid);
}

// Instances with body lists of their own, one block per instance: the same
// step as k_integrate_bodies with a single instance's forces.
__global__ void k_integrate_own_bodies(BodyLists a) {
  const BodyList& L = a.list[blockIdx.x];
  for (int k = threadIdx.x; k < L.n; k += blockDim.x) {
    const efl_body bd = L.bodies[k];
    if (!bd.dynamic) continue;
    float Fx = 0.0f, Fy = 0.0f, Tq = 0.0f;
    uint32_t id = L.obs[k].entity_id;
    if (id < (uint32_t)EFL_MAX_SOLIDS) {
      float4 f = L.fts[id];
      float s = 1.0f / (L.dx * L.dx);
      Fx += f.x * s;
      Fy += -f.y * s;
      Tq += f.w * s;
    }
    integrate_body(L.obs, bd, L.state, k, Fx, Fy, Tq, a.dt, L.gwx, L.gwy, L.dx, L.half_w, L.half_h, L.rho_fluid);
  }
}

// ---------------------------------------------------------------------------
// launch wrappers
// ---------------------------------------------------------------------------
static inline int nblocks(const Batch& b) { return b.start[b.n]; }
#define EFL_LAUNCH(kern, b, st, ...)                                                   \
  do {                                                                                 \
    int nb_ = nblocks(b);                                                              \
    if (nb_ > 0) hipLaunchKernelGGL(kern, dim3(nb_), dim3(NT), 0, st, b, ##__VA_ARGS__); \
  } while (0)
#define EFL_LAUNCH_PER_INST(kern, b, st, ...)                                          \
  do {                                                                                 \
    if ((b).n > 0) hipLaunchKernelGGL(kern, dim3((b).n), dim3(256), 0, st, b, ##__VA_ARGS__); \
  } while (0)

void launch_init_center(hipStream_t st, const Batch& b) { EFL_LAUNCH(k_init_center, b, st); }
void launch_init_edge(hipStream_t st, const Batch& b) { EFL_LAUNCH(k_init_edge, b, st); }
void launch_fluid_source(hipStream_t st, const Batch& b, int oneshot) { EFL_LAUNCH(k_fluid_source, b, st, oneshot); }
void launch_update_solid(hipStream_t st, const Batch& b) { EFL_LAUNCH(k_update_solid, b, st); }
void launch_area_fraction(hipStream_t st, const Batch& b) { EFL_LAUNCH(k_area_fraction, b, st); }
void launch_advect_u(hipStream_t st, const Batch& b) { EFL_LAUNCH(k_advect_u, b, st); }
void launch_advect_v(hipStream_t st, const Batch& b) { EFL_LAUNCH(k_advect_v, b, st); }
void launch_forces_u(hipStream_t st, const Batch& b) { EFL_LAUNCH(k_forces_u, b, st); }
void launch_forces_v(hipStream_t st, const Batch& b) { EFL_LAUNCH(k_forces_v, b, st); }
void launch_divergence(hipStream_t st, const Batch& b) { EFL_LAUNCH(k_divergence, b, st); }
void launch_gs_color(hipStream_t st, const Batch& b, int parity, float weight, int pcg) {
  EFL_LAUNCH(k_gs_color, b, st, parity, weight, pcg);
}
void launch_residual(hipStream_t st, const Batch& b, int pcg) { EFL_LAUNCH(k_residual, b, st, pcg); }
void launch_restrict(hipStream_t st, const Batch& b, int pcg) { EFL_LAUNCH(k_restrict, b, st, pcg); }
void launch_prolong(hipStream_t st, const Batch& b, int pcg) { EFL_LAUNCH(k_prolong, b, st, pcg); }
void launch_jacobi(hipStream_t st, const Batch& b, int reverse) { EFL_LAUNCH(k_jacobi, b, st, reverse); }
void launch_solve_u(hipStream_t st, const Batch& b) { EFL_LAUNCH(k_solve_u, b, st); }
void launch_solve_v(hipStream_t st, const Batch& b) { EFL_LAUNCH(k_solve_v, b, st); }
// validity mask + 10 passes; b tiles the coverage in 64 x 16 tiles, level
// = component of each slot
void launch_extrapolate_fused(hipStream_t st, const Batch& b) {
  const int nb = b.start[b.n];
  if (nb <= 0) return;
  hipLaunchKernelGGL(k_extrapolate_fused, dim3(nb), dim3(XNT), 0, st, b);
}
int extrapolate_tile_w() { return XT_W; }
int extrapolate_tile_h() { return XT_H; }
int extrapolate_passes() { return XP; }
void launch_advect_levelset(hipStream_t st, const Batch& b) { EFL_LAUNCH(k_advect_levelset, b, st); }
void launch_jfa_init(hipStream_t st, const Batch& b) { EFL_LAUNCH(k_jfa_init, b, st); }
void launch_jfa_step(hipStream_t st, const Batch& b, int step) { EFL_LAUNCH(k_jfa_step, b, st, step); }
void launch_jfa_sdf(hipStream_t st, const Batch& b) { EFL_LAUNCH(k_jfa_sdf, b, st); }
void launch_fim_init_active(hipStream_t st, const Batch& b) { EFL_LAUNCH(k_fim_init_active, b, st); }
void launch_fim_fused(hipStream_t st, const Batch& b, int src) {
  const int nb = b.start[b.n];
  if (nb <= 0) return;
  hipLaunchKernelGGL(k_fim_fused, dim3(nb), dim3(FNT), 0, st, b, src);
}
int fim_tile_w() { return FT_W; }
int fim_tile_h() { return FT_H; }
int fim_iterations_per_launch() { return FK; }
void launch_ls_gradient(hipStream_t st, const Batch& b) { EFL_LAUNCH(k_ls_gradient, b, st); }
void launch_sample_forces(hipStream_t st, const Batch& b) { EFL_LAUNCH(k_sample_forces, b, st); }
void launch_accumulate_forces(hipStream_t st, const Batch& b) { EFL_LAUNCH_PER_INST(k_accumulate_forces, b, st); }
void launch_residual_norm(hipStream_t st, const Batch& b) {
  EFL_LAUNCH(k_residual_norm, b, st, 0);
  EFL_LAUNCH_PER_INST(k_finalize, b, st, 0);
}
void launch_pcg_init(hipStream_t st, const Batch& b) {
  EFL_LAUNCH(k_residual_norm, b, st, 1);
  EFL_LAUNCH_PER_INST(k_cg_check, b, st, 1);
}
void launch_pcg_rz(hipStream_t st, const Batch& b) { EFL_LAUNCH(k_pcg_rz, b, st); }
void launch_pcg_apply(hipStream_t st, const Batch& b) { EFL_LAUNCH(k_pcg_apply, b, st); }
void launch_pcg_axpy(hipStream_t st, const Batch& b) { EFL_LAUNCH(k_pcg_axpy, b, st); }
void launch_cg_check(hipStream_t st, const Batch& b) { EFL_LAUNCH_PER_INST(k_cg_check, b, st, 0); }
void launch_prefold(hipStream_t st, const Batch& b, int kind) { EFL_LAUNCH_PER_INST(k_prefold, b, st, kind); }
void launch_gather_scalars(hipStream_t st, const Batch& b, double* dst) {
  if (b.n > 0) hipLaunchKernelGGL(k_gather_scalars, dim3(b.n), dim3(64), 0, st, b, dst);
}
void launch_cg_tally(hipStream_t st, const Batch& b) {
  if (b.n > 0) hipLaunchKernelGGL(k_cg_tally, dim3(b.n), dim3(64), 0, st, b);
}
void launch_mask_p(hipStream_t st, const Batch& b) { EFL_LAUNCH(k_mask_p, b, st); }
void launch_advect_forces_u(hipStream_t st, const Batch& b) { EFL_LAUNCH(k_advect_forces<0>, b, st); }
void launch_advect_forces_v(hipStream_t st, const Batch& b) { EFL_LAUNCH(k_advect_forces<1>, b, st); }
void launch_advect_forces_uv(hipStream_t st, const Batch& b) { EFL_LAUNCH(k_advect_forces_uv, b, st); }
void launch_solve_uv(hipStream_t st, const Batch& b) { EFL_LAUNCH(k_solve_uv, b, st); }
void launch_integrate_bodies(hipStream_t st, const BodyStep& a) {
  if (a.n > 0) hipLaunchKernelGGL(k_integrate_bodies, dim3((a.n + 63) / 64), dim3(64), 0, st, a);
}
void launch_integrate_own_bodies(hipStream_t st, const BodyLists& a) {
  if (a.n > 0) hipLaunchKernelGGL(k_integrate_own_bodies, dim3(a.n), dim3(64), 0, st, a);
}
void launch_fluid_diagnostics(hipStream_t st, const FluidDiagArgs& a) {
  if (a.W >= 16 && a.H >= 16) hipLaunchKernelGGL(k_fluid_diag_blocks, dim3(a.W / 16, a.H / 16), dim3(256), 0, st, a);
  hipLaunchKernelGGL(k_fluid_diag_final, dim3(1), dim3(512), 0, st, a);
}
void launch_max_divergence(hipStream_t st, const Batch& b) {
  EFL_LAUNCH(k_max_divergence, b, st);
  EFL_LAUNCH_PER_INST(k_finalize, b, st, 4);
}

int tile_w() { return TX; }
int tile_h() { return TY; }
int cg_tile_h() { return CG_TH; }
int sample_forces_tile_h() { return SF_TH; }

}  // namespace efl
