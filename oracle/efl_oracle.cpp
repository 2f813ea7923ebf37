// efl_oracle.cpp - CPU oracle of the per-step fluid pipeline (TEST INFRASTRUCTURE).
//
// This file restates, kernel by kernel, the WGSL compute shaders that
// bevy_eulerian_fluid's render node dispatches each step, and the dispatch
// order itself.  It is the checker for the HIP library (tests/), the source
// of the committed golden fixtures (tests/golden/) and the CPU baseline that
// bench.py times.  The product never links it.
//
// Conventions fixed here (WGSL leaves them to the adapter; they are not
// pinned by any reference test):
//   - out-of-bounds textureLoad returns 0, out-of-bounds textureStore is
//     dropped (Vulkan robustImageAccess2 / D3D12 typed UAV behaviour);
//   - clamp(NaN, 0, 1) = 0, i.e. min(max(x, 0), 1) with IEEE maxNum;
//   - round() is round-half-to-even, i32(f) truncates and saturates (NaN -> 0);
//   - mix(a, b, t) = a * (1 - t) + b * t; no FMA contraction (built with
//     -ffp-contract=off); mat4 * vec4 sums the four column products left to
//     right;
//   - dispatch coverage follows pipeline.rs:175-187 exactly, so sizes that are
//     not multiples of 8 / 64 leave the same cells untouched as the reference;
//   - the fast iterative method's update kernel (fast_iterative_method/
//     update.wgsl:13-83) races in the reference; the oracle runs the
//     deterministic two-phase schedule documented at fim_update() below.
//   - update_fluid_source.wgsl:64-69 writes shared faces from two cells; the
//     oracle runs invocations in raster order, so the later cell wins.
//
// Every function cites the reference file:line it follows (paths relative
// to the reference repository root).

#include "efl_oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

int g_threads = 0;  // 0 = OpenMP default

// ------------------------------------------------------------------------
// WGSL built-in semantics
// ------------------------------------------------------------------------
inline float wclamp(float x, float lo, float hi) { return std::fmin(std::fmax(x, lo), hi); }
inline float wmix(float a, float b, float t) { return a * (1.0f - t) + b * t; }
inline float wsign(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }
inline float wlen(float x, float y) { return std::sqrt(x * x + y * y); }
inline long f2i(float x) {  // i32(f): truncate, saturate, NaN -> 0
  if (std::isnan(x)) return 0;
  if (x >= 2147483647.0f) return 2147483647L;
  if (x <= -2147483648.0f) return -2147483648L;
  return (long)std::trunc(x);
}
inline float wround(float x) { return std::nearbyint(x); }  // default mode: half to even

// mat4 (column-major) * vec4(x, y, z, w), returns .xy
inline void mat_xy(const float* m, float x, float y, float z, float w, float& ox, float& oy) {
  ox = ((m[0] * x + m[4] * y) + m[8] * z) + m[12] * w;
  oy = ((m[1] * x + m[5] * y) + m[9] * z) + m[13] * w;
}

float det4(const float* m) {  // determinant() of a column-major mat4 (cofactor expansion)
  auto a = [&](int r, int c) { return m[c * 4 + r]; };
  float s0 = a(0, 0) * a(1, 1) - a(1, 0) * a(0, 1);
  float s1 = a(0, 0) * a(1, 2) - a(1, 0) * a(0, 2);
  float s2 = a(0, 0) * a(1, 3) - a(1, 0) * a(0, 3);
  float s3 = a(0, 1) * a(1, 2) - a(1, 1) * a(0, 2);
  float s4 = a(0, 1) * a(1, 3) - a(1, 1) * a(0, 3);
  float s5 = a(0, 2) * a(1, 3) - a(1, 2) * a(0, 3);
  float c5 = a(2, 2) * a(3, 3) - a(3, 2) * a(2, 3);
  float c4 = a(2, 1) * a(3, 3) - a(3, 1) * a(2, 3);
  float c3 = a(2, 1) * a(3, 2) - a(3, 1) * a(2, 2);
  float c2 = a(2, 0) * a(3, 3) - a(3, 0) * a(2, 3);
  float c1 = a(2, 0) * a(3, 2) - a(3, 0) * a(2, 2);
  float c0 = a(2, 0) * a(3, 1) - a(3, 0) * a(2, 1);
  return s0 * c5 - s1 * c4 + s2 * c3 + s3 * c2 - s4 * c1 + s5 * c0;
}

// ------------------------------------------------------------------------
// Storage textures
// ------------------------------------------------------------------------
template <class T>
struct Tex {
  int w = 0, h = 0, c = 1;
  std::vector<T> d;
  void alloc(int W, int H, int C = 1) {
    w = W; h = H; c = C;
    d.assign((size_t)W * H * C, T(0));
  }
  bool in(long x, long y) const { return x >= 0 && y >= 0 && x < w && y < h; }
  T ld(long x, long y, int ch = 0) const { return in(x, y) ? d[((size_t)y * w + x) * c + ch] : T(0); }
  void st(long x, long y, T v, int ch = 0) {
    if (in(x, y)) d[((size_t)y * w + x) * c + ch] = v;
  }
  T& at(long x, long y, int ch = 0) { return d[((size_t)y * w + x) * c + ch]; }
  const T& at(long x, long y, int ch = 0) const { return d[((size_t)y * w + x) * c + ch]; }
};

struct Uniform {  // SimulationUniform, fluid_uniform.wgsl:3-10
  float dx, dt, rho;
  float g[2];
  float T[16];
  float size[2];
};

struct AF { float im, ip, jm, jp; };

// area_fraction, utils/area_fraction.wgsl:10-12
inline float area_fraction(float a, float b) { return wclamp(std::fmax(a, b) / std::fabs(a - b), 0.0f, 1.0f); }

// area_fractions, utils/area_fraction.wgsl:14-56
AF area_fractions(const Tex<float>& ls, long x, long y) {
  float c[9];
  int k = 0;
  for (int dy = -1; dy <= 1; ++dy)
    for (int dx = -1; dx <= 1; ++dx) c[k++] = ls.ld(x + dx, y + dy);
  float mm = 0.25f * (((c[0] + c[1]) + c[3]) + c[4]);
  float pm = 0.25f * (((c[1] + c[2]) + c[4]) + c[5]);
  float mp = 0.25f * (((c[3] + c[4]) + c[6]) + c[7]);
  float pp = 0.25f * (((c[4] + c[5]) + c[7]) + c[8]);
  return AF{area_fraction(mm, mp), area_fraction(pm, pp), area_fraction(mm, pm), area_fraction(mp, pp)};
}

// ------------------------------------------------------------------------
// Interpolation helpers, utils/coordinate.wgsl
// ------------------------------------------------------------------------
float interp2d_edge_x(const Tex<float>& u, float px, float py) {  // :51-65
  long i = f2i(wround(px)), j = f2i(std::floor(py));
  float fi = (px + 0.5f) - (float)i, fj = py - (float)j;
  return wmix(wmix(u.ld(i, j), u.ld(i + 1, j), fi), wmix(u.ld(i, j + 1), u.ld(i + 1, j + 1), fi), fj);
}
float interp2d_edge_y(const Tex<float>& v, float px, float py) {  // :67-81
  long i = f2i(std::floor(px)), j = f2i(wround(py));
  float fi = px - (float)i, fj = (py + 0.5f) - (float)j;
  return wmix(wmix(v.ld(i, j), v.ld(i + 1, j), fi), wmix(v.ld(i, j + 1), v.ld(i + 1, j + 1), fi), fj);
}
void runge_kutta(const Tex<float>& u, const Tex<float>& v, float px, float py, float dt, float& ox, float& oy) {  // :83-94
  float vx = interp2d_edge_x(u, px, py), vy = interp2d_edge_y(v, px, py);
  float h = 0.5f * dt;
  float mx = px - h * vx, my = py - h * vy;
  float wx = interp2d_edge_x(u, mx, my), wy = interp2d_edge_y(v, mx, my);
  ox = px - dt * wx;
  oy = py - dt * wy;
}

// ------------------------------------------------------------------------
// Obstacles, update_solid.wgsl:80-175 and solid_obstacle.wgsl:43-89
// ------------------------------------------------------------------------
const float LARGE_FLOAT_SOLID = 1.0e6f;

float dist_seg(float x, float y, float ax, float ay, float bx, float by) {  // :164-169
  float abx = bx - ax, aby = by - ay;
  float t = wclamp(((x - ax) * abx + (y - ay) * aby) / (abx * abx + aby * aby), 0.0f, 1.0f);
  float qx = ax + t * abx, qy = ay + t * aby;
  return wlen(x - qx, y - qy);
}
float sign_side(float x, float y, float ax, float ay, float bx, float by) {  // :171-175
  float x1 = x - ax, y1 = y - ay, abx = bx - ax, aby = by - ay;
  return wsign(y1 * abx - x1 * aby);
}
float level_obstacle(const efl_obstacle& o, float x, float y) {  // :121-162
  switch (o.shape) {
    case 0: {
      float dx = x - o.transform[12], dy = y - o.transform[13];
      return wlen(dx, dy) - o.values[0];
    }
    case 1: {  // level_rectangle :87-109
      float level = LARGE_FLOAT_SOLID;
      if (det4(o.transform) == 0.0f) return level;
      float x0, y0;
      mat_xy(o.inverse_transform, x, y, 0.0f, 1.0f, x0, y0);
      float hx = o.values[0], hy = o.values[1];
      bool ix = std::fabs(x0) < hx, iy = std::fabs(y0) < hy;
      if (ix) {
        if (iy) level = std::fmax(std::fabs(x0) - hx, std::fabs(y0) - hy);
        else level = std::fabs(y0) - hy;
      } else {
        if (iy) level = std::fabs(x0) - hx;
        else level = wlen(std::fabs(x0) - hx, std::fabs(y0) - hy);
      }
      return level;
    }
    case 2: {
      float ax, ay, bx, by;
      mat_xy(o.transform, o.values[0], o.values[1], 0.0f, 1.0f, ax, ay);
      mat_xy(o.transform, o.values[2], o.values[3], 0.0f, 1.0f, bx, by);
      return dist_seg(x, y, ax, ay, bx, by) - o.values[4];
    }
    case 4: {
      float p0x, p0y, p1x, p1y, p2x, p2y;
      mat_xy(o.transform, o.values[0], o.values[1], 0.0f, 1.0f, p0x, p0y);
      mat_xy(o.transform, o.values[2], o.values[3], 0.0f, 1.0f, p1x, p1y);
      mat_xy(o.transform, o.values[4], o.values[5], 0.0f, 1.0f, p2x, p2y);
      float d0 = dist_seg(x, y, p0x, p0y, p1x, p1y);
      float d1 = dist_seg(x, y, p1x, p1y, p2x, p2y);
      float d2 = dist_seg(x, y, p2x, p2y, p0x, p0y);
      float s0 = sign_side(x, y, p0x, p0y, p1x, p1y);
      float s1 = sign_side(x, y, p1x, p1y, p2x, p2y);
      float s2 = sign_side(x, y, p2x, p2y, p0x, p0y);
      float m = std::fmin(std::fmin(d0, d1), d2);
      if (s0 > 0.0f && s1 > 0.0f && s2 > 0.0f) return -m;
      return m;
    }
    default:
      return LARGE_FLOAT_SOLID;
  }
}
void velocity_at(const efl_obstacle& o, float x, float y, float& vx, float& vy) {  // :111-119
  float rx = x - o.transform[12], ry = y - o.transform[13];
  float w = o.angular_velocity;
  vx = o.linear_velocity[0] + (-w * ry);
  vy = o.linear_velocity[1] + w * rx;
}
void center_of_mass(const efl_obstacle& o, float& cx, float& cy) {  // solid_obstacle.wgsl:67-89
  switch (o.shape) {
    case 0: case 1: case 2:
      cx = o.transform[12]; cy = o.transform[13]; return;
    case 4: {
      float lx = ((o.values[0] + o.values[2]) + o.values[4]) / 3.0f;
      float ly = ((o.values[1] + o.values[3]) + o.values[5]) / 3.0f;
      mat_xy(o.transform, lx, ly, 0.0f, 1.0f, cx, cy);
      return;
    }
    default:
      cx = 0.0f; cy = 0.0f;
  }
}

// ------------------------------------------------------------------------
// Fluid sources, fluid_source/update_fluid_source.wgsl:72-107
// ------------------------------------------------------------------------
float level_aabb(float hx, float hy, float cx, float cy, float x, float y) {
  float level = LARGE_FLOAT_SOLID;
  float dx = std::fabs(cx - x) - hx, dy = std::fabs(cy - y) - hy;
  bool ix = dx < 0.0f, iy = dy < 0.0f;
  if (ix) level = iy ? std::fmax(dx, dy) : dy;
  else level = iy ? dx : wlen(dx, dy);
  return level;
}
float level_source(const efl_source& s, float x, float y) {
  switch (s.shape) {
    case 0: return wlen(s.center[0] - x, s.center[1] - y) - s.data[0];
    case 1: return level_aabb(s.data[0], s.data[1], s.center[0], s.center[1], x, y);
    default: return 0.0f;
  }
}

// ------------------------------------------------------------------------
// Instance
// ------------------------------------------------------------------------
const uint32_t LABEL_NONE = 0, LABEL_SOURCE = 1, LABEL_ACTIVE = 2;
const float FIM_LARGE = 1e30f;

}  // namespace

struct orc_inst {
  int W = 0, H = 0;
  Uniform uni{};
  efl_projection proj{};
  efl_reinit reinit{};
  std::vector<efl_source> src_init, src_step;
  std::vector<efl_force> forces;
  std::vector<efl_obstacle> obstacles;
  std::vector<efl_body> bodies;
  std::vector<std::array<float, 3>> body_state;  // previous accelerations
  std::vector<float> custom_ls;
  bool initialized = false;
  bool diag = false;
  efl_stats stats{};

  Tex<float> u0, u1, v0, v1, us, vs, div, p0, p1, ls0, ls1, grad, lss, afs, seeds[2], fts;
  Tex<int32_t> sid, uval[2], vval[2], bins;
  Tex<uint32_t> lab0, lab;
  // multigrid pyramid (multi_grid.rs:85-131); level 0 aliases p0/div/ls0/afs
  int nlev = 1;
  std::vector<Tex<float>> mx, mb, mls, mafs, mr;
  // PCG work vectors (level-0 sized)
  Tex<float> pcg_r, pcg_z, pcg_p, pcg_q;
};

namespace {

// dispatch coverage, pipeline.rs:175-187 and multi_grid.rs:405
struct Cov { long x0, x1, y0, y1; };
inline Cov cov_center(const orc_inst* o) { return {0, (long)(o->W / 8) * 8, 0, (long)(o->H / 8) * 8}; }
inline Cov cov_xmajor(const orc_inst* o) { return {0, (long)o->W + 1, 0, (long)(o->H / 64) * 64}; }
inline Cov cov_ymajor(const orc_inst* o) { return {0, (long)(o->W / 64) * 64, 0, (long)o->H + 1}; }

int nthreads() {
#ifdef _OPENMP
  return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
  return 1;
#endif
}

// ---- K1 initialize_grid_center, initialize_grid_center.wgsl:5-15 ----
void k_init_center(orc_inst* o) {
  Cov c = cov_center(o);
  for (long y = c.y0; y < c.y1; ++y)
    for (long x = c.x0; x < c.x1; ++x) {
      float level = (float)(uint32_t)(o->H - y);
      o->ls0.st(x, y, level);
      o->ls1.st(x, y, level);
      o->grad.st(x, y, 0.0f, 0);
      o->grad.st(x, y, -1.0f, 1);
    }
}

// ---- K2 initialize_grid_edge, initialize_grid_edge.wgsl:6-17, grid render_node.rs:228-248 ----
void k_init_edge(orc_inst* o) {
  Cov c = cov_xmajor(o);
  for (long y = c.y0; y < c.y1; ++y)
    for (long x = c.x0; x < c.x1; ++x) {
      o->u0.st(x, y, 0.0f);
      o->v0.st(y, x, 0.0f);
      o->u1.st(x, y, 0.0f);
      o->v1.st(y, x, 0.0f);
    }
}

// ---- K3 update_fluid_source, update_fluid_source.wgsl:30-70 ----
void k_fluid_source(orc_inst* o, const std::vector<efl_source>& srcs) {
  Cov c = cov_center(o);
  float dimx = (float)o->W, dimy = (float)o->H;
  uint32_t n = (uint32_t)std::min<size_t>(srcs.size(), EFL_MAX_FLUID_SOURCES);
  for (long y = c.y0; y < c.y1; ++y)  // raster order: a later invocation wins shared faces
    for (long x = c.x0; x < c.x1; ++x) {
      float uvx = (float)x / dimx, uvy = (float)y / dimy;
      float ndcx = uvx * 2.0f + (-1.0f), ndcy = uvy * (-2.0f) + 1.0f;  // bevy uv_to_ndc
      float px = (0.5f * dimx) * ndcx, py = (0.5f * dimy) * ndcy;
      float lv = o->ls1.ld(x, y);
      float vx = 0.0f, vy = 0.0f;
      bool has_source = false, vel_updated = false;
      for (uint32_t i = 0; i < n; ++i) {
        const efl_source& s = srcs[i];
        float sdf = level_source(s, px, py);
        if (s.mode == 0) {
          lv = std::fmin(lv, sdf);
          has_source = true;
          if (sdf < 0.0f) {
            vx = vx + s.velocity[0];
            vy = vy + s.velocity[1];
            vel_updated = true;
          }
        } else {
          if (!has_source) lv = std::fmax(lv, -sdf);
        }
      }
      o->ls1.st(x, y, lv);
      if (vel_updated) {
        o->u0.st(x, y, vx);
        o->u0.st(x + 1, y, vx);
        o->v0.st(x, y, vy);
        o->v0.st(x, y + 1, vy);
      }
    }
}

// ---- K4 update_solid, update_solid.wgsl:15-85 ----
void k_update_solid(orc_inst* o) {
  Cov c = cov_center(o);
  const Uniform& U = o->uni;
  auto to_world = [&](float fx, float fy, float& wx, float& wy) {
    float uvx = fx / (float)o->W, uvy = fy / (float)o->H;
    float xx = (uvx - 0.5f) * U.size[0], yy = (-uvy + 0.5f) * U.size[1];
    mat_xy(U.T, xx, yy, 0.0f, 1.0f, wx, wy);
  };
  long W = o->W, H = o->H;
#pragma omp parallel for schedule(static) num_threads(nthreads())
  for (long y = c.y0; y < c.y1; ++y)
    for (long x = c.x0; x < c.x1; ++x) {
      float cxw, cyw, exw, eyw, fxw, fyw;
      to_world((float)x, (float)y, cxw, cyw);
      to_world((float)x - 0.5f, (float)y, exw, eyw);
      to_world((float)x, (float)y - 0.5f, fxw, fyw);
      if (x == 0 || y == 0 || x == W - 1 || y == H - 1) {
        o->lss.st(x, y, 0.0f);
        o->us.st(x, y, 0.0f);
        o->vs.st(x, y, 0.0f);
        o->sid.st(x, y, -1);
        continue;
      }
      long tx = std::min(x, std::labs(W - 1 - x)), ty = std::min(y, std::labs(H - 1 - y));
      float level = (float)std::min(tx, ty);
      float uu = 0.0f, vv = 0.0f;
      int32_t sid = -1;
      for (const efl_obstacle& ob : o->obstacles) {
        float lo = level_obstacle(ob, cxw, cyw);
        if (lo < level) level = lo;
        if (lo < 0.5f) sid = (int32_t)ob.entity_id;
        float le = level_obstacle(ob, exw, eyw);
        if (le < 0.5f) {
          float a, b;
          velocity_at(ob, exw, eyw, a, b);
          uu = a;
        }
        float lf = level_obstacle(ob, fxw, fyw);
        if (lf < 0.5f) {
          float a, b;
          velocity_at(ob, fxw, fyw, a, b);
          vv = -b;
        }
      }
      o->us.st(x, y, uu);
      o->vs.st(x, y, vv);
      o->lss.st(x, y, level);
      o->sid.st(x, y, sid);
    }
}

// ---- K5 update_area_fraction, update_area_fraction.wgsl:6-14 ----
void k_area_fraction(orc_inst* o) {
  Cov c = cov_center(o);
#pragma omp parallel for schedule(static) num_threads(nthreads())
  for (long y = c.y0; y < c.y1; ++y)
    for (long x = c.x0; x < c.x1; ++x) {
      AF f = area_fractions(o->lss, x, y);
      o->afs.st(x, y, f.im, 0);
      o->afs.st(x, y, f.ip, 1);
      o->afs.st(x, y, f.jm, 2);
      o->afs.st(x, y, f.jp, 3);
    }
}

// ---- K6/K7 advect_u / advect_v, advect_velocity.wgsl:11-52 ----
void k_advect_velocity(orc_inst* o) {
  float dt = o->uni.dt;
  Cov cu = cov_xmajor(o);
  float dux = (float)o->u0.w, duy = (float)o->u0.h;
#pragma omp parallel for schedule(static) num_threads(nthreads())
  for (long y = cu.y0; y < cu.y1; ++y)
    for (long x = cu.x0; x < cu.x1; ++x) {
      float px = (float)x + (-0.5f), py = (float)y + 0.0f;
      float bx, by;
      runge_kutta(o->u0, o->v0, px, py, dt, bx, by);
      bool inside = bx >= 0.0f && by >= 0.0f && bx <= dux - 1.0f && by <= duy - 1.0f;
      o->u1.st(x, y, inside ? interp2d_edge_x(o->u0, bx, by) : o->u0.ld(x, y));
    }
  Cov cv = cov_ymajor(o);
  float dvx = (float)o->v0.w, dvy = (float)o->v0.h;
#pragma omp parallel for schedule(static) num_threads(nthreads())
  for (long y = cv.y0; y < cv.y1; ++y)
    for (long x = cv.x0; x < cv.x1; ++x) {
      float px = (float)x + 0.0f, py = (float)y + (-0.5f);
      float bx, by;
      runge_kutta(o->u0, o->v0, px, py, dt, bx, by);
      bool inside = bx >= 0.0f && by >= 0.0f && bx <= dvx - 1.0f && by <= dvy - 1.0f;
      o->v1.st(x, y, inside ? interp2d_edge_y(o->v0, bx, by) : o->v0.ld(x, y));
    }
}

// ---- K8/K9 apply_forces_u / apply_forces_v, apply_forces.wgsl:17-76 ----
inline float gaussian_2d(float x, float y, float x0, float y0) {
  const float sigma = 10.0f;
  float b = -1.0f / ((2.0f * sigma) * sigma);
  float dx = x - x0, dy = y - y0;
  return std::exp(b * (dx * dx + dy * dy));
}
void k_apply_forces(orc_inst* o) {
  const Uniform& U = o->uni;
  Cov cu = cov_xmajor(o);
  long nf = (long)o->forces.size();
#pragma omp parallel for schedule(static) num_threads(nthreads())
  for (long y = cu.y0; y < cu.y1; ++y)
    for (long x = cu.x0; x < cu.x1; ++x) {
      AF f = area_fractions(o->ls0, x, y);
      float fs = o->afs.ld(x, y, 0);
      if (f.im == 1.0f || fs == 0.0f) {
        o->u1.st(x, y, 0.0f);
        continue;
      }
      float net = U.g[0];
      for (long n = nf - 1; n >= 0; --n) {
        const efl_force& F = o->forces[n];
        net = net + F.force[0] * gaussian_2d((float)x, (float)y, F.position[0], F.position[1]);
      }
      float uv = o->u1.ld(x, y);
      o->u1.st(x, y, uv + (net * U.dt) / U.dx);
    }
  Cov cv = cov_ymajor(o);
#pragma omp parallel for schedule(static) num_threads(nthreads())
  for (long y = cv.y0; y < cv.y1; ++y)
    for (long x = cv.x0; x < cv.x1; ++x) {
      AF f = area_fractions(o->ls0, x, y);
      float fs = o->afs.ld(x, y, 2);
      if (f.jm == 1.0f || fs == 0.0f) {
        o->v1.st(x, y, 0.0f);
        continue;
      }
      float net = U.g[1];
      for (long n = nf - 1; n >= 0; --n) {
        const efl_force& F = o->forces[n];
        net = net + F.force[1] * gaussian_2d((float)x, (float)y, F.position[0], F.position[1]);
      }
      float vv = o->v1.ld(x, y);
      o->v1.st(x, y, vv + (net * U.dt) / U.dx);
    }
}

// ---- K10 divergence, divergence.wgsl:16-42 ----
void k_divergence(orc_inst* o) {
  Cov c = cov_center(o);
  float dx = o->uni.dx;
#pragma omp parallel for schedule(static) num_threads(nthreads())
  for (long y = c.y0; y < c.y1; ++y)
    for (long x = c.x0; x < c.x1; ++x) {
      AF f = area_fractions(o->lss, x, y);
      float u_ij = o->u1.ld(x, y), u_ip = o->u1.ld(x + 1, y);
      float v_ij = o->v1.ld(x, y), v_jp = o->v1.ld(x, y + 1);
      float du_f = f.ip * u_ip - f.im * u_ij;
      float dv_f = f.jp * v_jp - f.jm * v_ij;
      float us_m = o->us.ld(x, y), us_p = o->us.ld(x + 1, y);
      float vs_m = o->vs.ld(x, y), vs_p = o->vs.ld(x, y + 1);
      float du_s = (1.0f - f.ip) * us_p - (1.0f - f.im) * us_m;
      float dv_s = (1.0f - f.jp) * vs_p - (1.0f - f.jm) * vs_m;
      float rhs = -(((du_f + dv_f) + du_s) + dv_s) / dx;
      o->div.st(x, y, rhs);
    }
}

// ------------------------------------------------------------------------
// Pressure: the shared operator of gauss_seidel.wgsl:44-88 and
// residual.wgsl:12-57, written once per scalar type so that the FP64 solve
// uses the same discrete system.
// ------------------------------------------------------------------------
template <class T>
struct Lvl {
  int w, h;
  T* x;
  const T* b;
  const float* ls;
  const float* af;  // 4 per cell
  T* r;
  T scale;          // resolution_scale = 2^level
};

template <class T>
inline T gs_cell(const Lvl<T>& L, long x, long y, T weight, const Uniform& U) {  // update_pressure :44-88
  size_t id = (size_t)y * L.w + x;
  float lc = L.ls[id];
  if (lc >= 0.0f) return T(0);
  const float* f = L.af + id * 4;
  bool fully_solid = f[0] == 0.0f && f[1] == 0.0f && f[2] == 0.0f && f[3] == 0.0f;
  T p_old = L.x[id];
  if (fully_solid) return T(0);
  T denom = T(0);
  T dx = L.scale * (T)U.dx;
  T nume = dx * dx * (T)U.rho / (T)U.dt * L.b[id];
  static const int ox[4] = {-1, 1, 0, 0}, oy[4] = {0, 0, -1, 1};
  for (int i = 0; i < 4; ++i) {
    long jx = x + ox[i], jy = y + oy[i];
    if (jx >= 0 && jy >= 0 && jx < L.w && jy < L.h) {
      size_t jd = (size_t)jy * L.w + jx;
      float level = L.ls[jd];
      if (level < 0.0f) {
        denom += (T)f[i];
        nume += (T)f[i] * L.x[jd];
      } else {
        denom += (T)f[i] * (T(1) - (T)level / (T)lc);
      }
    }
  }
  if (std::fabs(denom) < T(1.0e-6)) return T(0);
  T p_new = nume / denom;
  return weight * p_new + (T(1) - weight) * p_old;
}

// one colour of gauss_seidel_red / gauss_seidel_black (:12-42) over a
// cover of (cw, ch) cells, guarded by the level dims
template <class T>
void gs_color(const Lvl<T>& L, int parity, T weight, const Uniform& U, long cw, long ch) {
#pragma omp parallel for schedule(static) num_threads(nthreads())
  for (long y = 0; y < std::min<long>(ch, L.h); ++y)
    for (long x = 0; x < std::min<long>(cw, L.w); ++x)
      if (((x + y) & 1) == parity) L.x[(size_t)y * L.w + x] = gs_cell(L, x, y, weight, U);
}

template <class T>
void residual_level(const Lvl<T>& L, const Uniform& U) {  // residual.wgsl:12-57
  T dx = (T)U.dx * L.scale;
  T factor = (T)U.dt / (T)U.rho / dx / dx;
#pragma omp parallel for schedule(static) num_threads(nthreads())
  for (long y = 0; y < L.h; ++y)
    for (long x = 0; x < L.w; ++x) {
      size_t id = (size_t)y * L.w + x;
      float phi = L.ls[id];
      if (phi >= 0.0f) { L.r[id] = T(0); continue; }
      const float* f = L.af + id * 4;
      if (f[0] == 0.0f && f[1] == 0.0f && f[2] == 0.0f && f[3] == 0.0f) { L.r[id] = T(0); continue; }
      T xc = L.x[id];
      T res = L.b[id];
      static const int ox[4] = {-1, 1, 0, 0}, oy[4] = {0, 0, -1, 1};
      for (int i = 0; i < 4; ++i) {
        long jx = x + ox[i], jy = y + oy[i];
        if (jx >= 0 && jy >= 0 && jx < L.w && jy < L.h) {
          size_t jd = (size_t)jy * L.w + jx;
          float pn = L.ls[jd];
          if (pn < 0.0f) res -= (T)f[i] * (xc - L.x[jd]) * factor;
          else res -= (T)f[i] * (T(1) - (T)pn / (T)phi) * xc * factor;
        }
      }
      L.r[id] = res;
    }
}

// restriction.wgsl:10-60 (fine r, ls, af -> coarse b, ls, af, x = 0).  The
// geometry is always float; b/x are T.
template <class T>
void restrict_level(const Lvl<T>& F, int cw, int chh, T* cb, float* cls, float* caf, T* cx) {
#pragma omp parallel for schedule(static) num_threads(nthreads())
  for (long y = 0; y < chh; ++y)
    for (long x = 0; x < cw; ++x) {
      auto ldr = [&](long a, long b) -> T { return (a < F.w && b < F.h) ? F.r[(size_t)b * F.w + a] : T(0); };
      auto ldl = [&](long a, long b) -> float { return (a < F.w && b < F.h) ? F.ls[(size_t)b * F.w + a] : 0.0f; };
      auto ldf = [&](long a, long b, int c) -> float { return (a < F.w && b < F.h) ? F.af[((size_t)b * F.w + a) * 4 + c] : 0.0f; };
      long fx = 2 * x, fy = 2 * y;
      T br = T(0.25) * (((ldr(fx, fy) + ldr(fx, fy + 1)) + ldr(fx + 1, fy)) + ldr(fx + 1, fy + 1));
      float phi0 = ldl(fx, fy), phi1 = ldl(fx, fy + 1), phi2 = ldl(fx + 1, fy), phi3 = ldl(fx + 1, fy + 1);
      float pr = phi0;
      if (std::fabs(phi1) < std::fabs(pr)) pr = phi1;
      if (std::fabs(phi2) < std::fabs(pr)) pr = phi2;
      if (std::fabs(phi3) < std::fabs(pr)) pr = phi3;
      size_t id = (size_t)y * cw + x;
      caf[id * 4 + 0] = 0.5f * (ldf(fx, fy, 0) + ldf(fx, fy + 1, 0));
      caf[id * 4 + 1] = 0.5f * (ldf(fx + 1, fy, 1) + ldf(fx + 1, fy + 1, 1));
      caf[id * 4 + 2] = 0.5f * (ldf(fx, fy, 2) + ldf(fx + 1, fy, 2));
      caf[id * 4 + 3] = 0.5f * (ldf(fx, fy + 1, 3) + ldf(fx + 1, fy + 1, 3));
      cb[id] = br;
      cls[id] = pr;
      cx[id] = T(0);
    }
}

// prolongation.wgsl:5-34
template <class T>
void prolong_level(const Lvl<T>& F, int cw, int chh, const T* cx) {
#pragma omp parallel for schedule(static) num_threads(nthreads())
  for (long y = 0; y < chh; ++y)
    for (long x = 0; x < cw; ++x) {
      T corr = cx[(size_t)y * cw + x];
      static const int ox[4] = {0, 1, 0, 1}, oy[4] = {0, 0, 1, 1};
      for (int i = 0; i < 4; ++i) {
        long fx = 2 * x + ox[i], fy = 2 * y + oy[i];
        if (fx < F.w && fy < F.h) {
          size_t fd = (size_t)fy * F.w + fx;
          if (F.ls[fd] < 0.0f) F.x[fd] = corr + F.x[fd];
        }
      }
    }
}

// A multigrid hierarchy over which a V-cycle runs, scalar type T.
template <class T>
struct Hier {
  int nlev;
  std::vector<int> w, h;
  std::vector<T*> x, r;
  std::vector<const T*> b;
  std::vector<float*> ls, af;
  std::vector<T*> bmut;  // writable coarse b (level >= 1)
};

// v_cycle, multi_grid.rs:390-475.  symmetric = post-smooth sweeps run
// black-then-red and the coarsest level runs n (red, black) then n (black,
// red), so that the cycle is a symmetric preconditioner; the reference order
// is symmetric = false.
template <class T>
void v_cycle(Hier<T>& Hh, int i, const efl_projection& cfg, const Uniform& U, bool symmetric) {
  Lvl<T> L{Hh.w[i], Hh.h[i], Hh.x[i], Hh.b[i], Hh.ls[i], Hh.af[i], Hh.r[i], (T)(float)(1 << i)};
  long cw = ((Hh.w[i] + 7) / 8) * 8, chh = ((Hh.h[i] + 7) / 8) * 8;
  T one = T(1);
  if (i == Hh.nlev - 1) {
    for (uint32_t k = 0; k < cfg.mg_coarsest_iterations; ++k) {
      gs_color(L, 1, one, U, cw, chh);
      gs_color(L, 0, one, U, cw, chh);
    }
    if (symmetric)
      for (uint32_t k = 0; k < cfg.mg_coarsest_iterations; ++k) {
        gs_color(L, 0, one, U, cw, chh);
        gs_color(L, 1, one, U, cw, chh);
      }
    return;
  }
  for (uint32_t k = 0; k < cfg.mg_pre_iterations; ++k) {
    gs_color(L, 1, one, U, cw, chh);
    gs_color(L, 0, one, U, cw, chh);
  }
  residual_level(L, U);
  restrict_level(L, Hh.w[i + 1], Hh.h[i + 1], Hh.bmut[i + 1], Hh.ls[i + 1], Hh.af[i + 1], Hh.x[i + 1]);
  v_cycle(Hh, i + 1, cfg, U, symmetric);
  prolong_level(L, Hh.w[i + 1], Hh.h[i + 1], Hh.x[i + 1]);
  for (uint32_t k = 0; k < cfg.mg_post_iterations; ++k) {
    if (symmetric) {
      gs_color(L, 0, one, U, cw, chh);
      gs_color(L, 1, one, U, cw, chh);
    } else {
      gs_color(L, 1, one, U, cw, chh);
      gs_color(L, 0, one, U, cw, chh);
    }
  }
}

Hier<float> float_hier(orc_inst* o, float* x0, const float* b0, float* r0) {
  Hier<float> Hh;
  Hh.nlev = o->nlev;
  for (int i = 0; i < o->nlev; ++i) {
    Hh.w.push_back(i == 0 ? o->W : o->mx[i].w);
    Hh.h.push_back(i == 0 ? o->H : o->mx[i].h);
    Hh.x.push_back(i == 0 ? x0 : o->mx[i].d.data());
    Hh.b.push_back(i == 0 ? b0 : o->mb[i].d.data());
    Hh.bmut.push_back(i == 0 ? nullptr : o->mb[i].d.data());
    Hh.ls.push_back(i == 0 ? o->ls0.d.data() : o->mls[i].d.data());
    Hh.af.push_back(i == 0 ? o->afs.d.data() : o->mafs[i].d.data());
    Hh.r.push_back(i == 0 ? r0 : o->mr[i].d.data());
  }
  return Hh;
}

// Active cells of the pressure system: phi < 0 and not fully solid.
inline bool active_cell(const orc_inst* o, size_t id) {
  if (o->ls0.d[id] >= 0.0f) return false;
  const float* f = &o->afs.d[id * 4];
  return !(f[0] == 0.0f && f[1] == 0.0f && f[2] == 0.0f && f[3] == 0.0f);
}

double masked_norm2(const orc_inst* o, const float* v) {
  double s = 0.0;
  size_t n = (size_t)o->W * o->H;
  for (size_t i = 0; i < n; ++i)
    if (active_cell(o, i)) s += (double)v[i] * (double)v[i];
  return std::sqrt(s);
}

double rel_residual_now(orc_inst* o) {
  Lvl<float> L{o->W, o->H, o->p0.d.data(), o->div.d.data(), o->ls0.d.data(), o->afs.d.data(), o->mr[0].d.data(), 1.0f};
  residual_level(L, o->uni);
  double bn = masked_norm2(o, o->div.d.data());
  double rn = masked_norm2(o, o->mr[0].d.data());
  return bn > 0.0 ? rn / bn : rn;
}

// A x for the FP32 system (residual.wgsl's operator with b = 0, negated).
void apply_A(orc_inst* o, const float* x, float* out) {
  std::vector<float> zero((size_t)o->W * o->H, 0.0f);
  Lvl<float> L{o->W, o->H, const_cast<float*>(x), zero.data(), o->ls0.d.data(), o->afs.d.data(), out, 1.0f};
  residual_level(L, o->uni);
  size_t n = (size_t)o->W * o->H;
  for (size_t i = 0; i < n; ++i) out[i] = -out[i];
}

double dotm(const orc_inst* o, const float* a, const float* b) {
  double s = 0.0;
  size_t n = (size_t)o->W * o->H;
  for (size_t i = 0; i < n; ++i)
    if (active_cell(o, i)) s += (double)a[i] * (double)b[i];
  return s;
}

// M z = r with one symmetric V-cycle from a zero guess.
void precond(orc_inst* o, const float* r, float* z) {
  size_t n = (size_t)o->W * o->H;
  std::fill(z, z + n, 0.0f);
  Hier<float> Hh = float_hier(o, z, r, o->mr[0].d.data());
  v_cycle(Hh, 0, o->proj, o->uni, true);
  for (size_t i = 0; i < n; ++i)
    if (!active_cell(o, i)) z[i] = 0.0f;
}

// Fluid regions that the pressure system leaves without a Dirichlet term:
// components of the active cells (linked across open faces, f > 0) none of
// whose open faces leads to air, a fully solid fluid cell or a cell outside
// the grid's stencil.  On such a component the operator is singular (constants
// are its null space) and the system has a solution only if the divergence
// sums to zero there, which fluid trapped between bodies that move toward
// each other breaks.  This subtracts the component's mean divergence (double
// sum over the cells in raster order, then one float subtraction per cell),
// the compatibility condition of the Neumann problem.  Returns the number of
// components changed.  (No reference counterpart: the reference's fixed
// sweep counts never need the system to be solvable.)
int repair_sealed(orc_inst* o) {
  const int W = o->W, H = o->H;
  const size_t n = (size_t)W * H;
  std::vector<int32_t> comp(n, -1);
  std::vector<size_t> stack, cells;
  int repaired = 0, next = 0;
  static const int ox[4] = {-1, 1, 0, 0}, oy[4] = {0, 0, -1, 1};
  for (size_t s0 = 0; s0 < n; ++s0) {
    if (comp[s0] >= 0 || !active_cell(o, s0)) continue;
    const int id = next++;
    bool sealed = true;
    cells.clear();
    stack.assign(1, s0);
    comp[s0] = id;
    while (!stack.empty()) {
      const size_t c = stack.back();
      stack.pop_back();
      cells.push_back(c);
      const long cx = (long)(c % W), cy = (long)(c / W);
      for (int k = 0; k < 4; ++k) {
        if (!(o->afs.d[c * 4 + k] > 0.0f)) continue;  // closed face: no coupling
        const long jx = cx + ox[k], jy = cy + oy[k];
        if (jx < 0 || jy < 0 || jx >= W || jy >= H) continue;  // outside: no term
        const size_t j = (size_t)jy * W + jx;
        if (!active_cell(o, j)) { sealed = false; continue; }  // air or fully solid fluid: Dirichlet
        if (comp[j] < 0) {
          comp[j] = id;
          stack.push_back(j);
        }
      }
    }
    if (!sealed) continue;
    std::sort(cells.begin(), cells.end());
    double sum = 0.0;
    for (size_t c : cells) sum += (double)o->div.d[c];
    const float mean = (float)(sum / (double)cells.size());
    if (mean == 0.0f) continue;
    for (size_t c : cells) o->div.d[c] = o->div.d[c] - mean;
    ++repaired;
  }
  return repaired;
}

// MG-preconditioned CG from the current p0 (see k_projection); returns the
// iteration count and sets *rel_out, *conv.
uint32_t mgpcg(orc_inst* o, double* rel_out) {
  const efl_projection& P = o->proj;
  size_t n = (size_t)o->W * o->H;
  float* x = o->p0.d.data();
  float* r = o->pcg_r.d.data();
  float* z = o->pcg_z.d.data();
  float* p = o->pcg_p.d.data();
  float* q = o->pcg_q.d.data();
  for (size_t i = 0; i < n; ++i)
    if (!active_cell(o, i)) x[i] = 0.0f;
  apply_A(o, x, q);
  for (size_t i = 0; i < n; ++i) r[i] = active_cell(o, i) ? o->div.d[i] - q[i] : 0.0f;
  double bn = masked_norm2(o, o->div.d.data());
  double rn = masked_norm2(o, r);
  double rel = bn > 0 ? rn / bn : rn;
  uint32_t maxc = P.max_cycles ? P.max_cycles : 1;
  uint32_t k = 0;
  if (!(P.rel_tol > 0.0f && rel <= P.rel_tol)) {
    precond(o, r, z);
    std::copy(z, z + n, p);
    double rz = dotm(o, r, z);
    for (; k < maxc;) {
      apply_A(o, p, q);
      double pq = dotm(o, p, q);
      if (!(pq > 0.0)) break;
      float alpha = (float)(rz / pq);
      for (size_t i = 0; i < n; ++i) {
        x[i] = x[i] + alpha * p[i];
        r[i] = r[i] - alpha * q[i];
      }
      ++k;
      rn = masked_norm2(o, r);
      rel = bn > 0 ? rn / bn : rn;
      if (P.rel_tol > 0.0f && rel <= P.rel_tol) break;
      precond(o, r, z);
      double rz2 = dotm(o, r, z);
      float beta = (float)(rz2 / rz);
      rz = rz2;
      for (size_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    }
  }
  *rel_out = rel;
  return k;
}

// ProjectionMethod dispatch, projection.rs:89-164, plus the residual-target
// loop (MULTIGRID with max_cycles > 1 or rel_tol > 0) and MGPCG.
void k_projection(orc_inst* o) {
  const efl_projection& P = o->proj;
  const Uniform& U = o->uni;
  o->stats.cycles = 0;
  o->stats.converged = 1;
  o->stats.rel_residual = -1.0f;
  if (P.kind == EFL_PROJECTION_GAUSS_SEIDEL) {  // gauss_seidel.rs:120-148, weight 1.9 (:72)
    Lvl<float> L{o->W, o->H, o->p0.d.data(), o->div.d.data(), o->ls0.d.data(), o->afs.d.data(), o->mr[0].d.data(), 1.0f};
    Cov c = cov_center(o);
    for (uint32_t k = 0; k < P.gs_iterations; ++k) {
      gs_color(L, 1, 1.9f, U, c.x1, c.y1);
      gs_color(L, 0, 1.9f, U, c.x1, c.y1);
    }
    o->stats.cycles = P.gs_iterations;
    return;
  }
  if (P.kind == EFL_PROJECTION_JACOBI) {  // projection.rs:118-145, jacobi_iteration.wgsl:18-91
    Cov c = cov_center(o);
    float factor = U.dx * U.rho / U.dt;
    for (int it = 0; it < 100; ++it) {
      Tex<float>& src = (it & 1) ? o->p1 : o->p0;
      Tex<float>& dst = (it & 1) ? o->p0 : o->p1;
#pragma omp parallel for schedule(static) num_threads(nthreads())
      for (long y = c.y0; y < c.y1; ++y)
        for (long x = c.x0; x < c.x1; ++x) {
          float lc = o->ls0.ld(x, y);
          float out;
          if (lc >= 0.0f) {
            out = 0.0f;
          } else {
            AF f = area_fractions(o->lss, x, y);
            bool fs = f.im == 0.0f && f.ip == 0.0f && f.jm == 0.0f && f.jp == 0.0f;
            if (fs) {
              out = src.ld(x, y);
            } else {
              float pm = src.ld(x - 1, y), pp = src.ld(x + 1, y), qm = src.ld(x, y - 1), qp = src.ld(x, y + 1);
              float ffm = std::fmin(0.0f, o->ls0.ld(x - 1, y) / lc);
              float ffp = std::fmin(0.0f, o->ls0.ld(x + 1, y) / lc);
              float fgm = std::fmin(0.0f, o->ls0.ld(x, y - 1) / lc);
              float fgp = std::fmin(0.0f, o->ls0.ld(x, y + 1) / lc);
              float coef = ((f.im * (1.0f - ffm) + f.ip * (1.0f - ffp)) + f.jm * (1.0f - fgm)) + f.jp * (1.0f - fgp);
              if (std::fabs(coef) < 1.0e-6f) {
                out = 0.0f;
              } else {
                float d = o->div.ld(x, y);
                auto stp = [](float v) { return v >= 0.0f ? 1.0f : 0.0f; };
                float a = f.im * stp(ffm) * pm, b = f.ip * stp(ffp) * pp;
                float cc = f.jm * stp(fgm) * qm, dd = f.jp * stp(fgp) * qp;
                out = ((((a + b) + cc) + dd) - factor * d) / coef;
              }
            }
          }
          dst.st(x, y, out);
        }
    }
    o->stats.cycles = 100;
    return;
  }
  if (P.kind == EFL_PROJECTION_MULTIGRID) {  // multi_grid.rs:354-475
    uint32_t maxc = P.max_cycles ? P.max_cycles : 1;
    uint32_t every = P.check_every ? P.check_every : 1;
    Hier<float> Hh = float_hier(o, o->p0.d.data(), o->div.d.data(), o->mr[0].d.data());
    uint32_t k = 0;
    for (; k < maxc;) {
      v_cycle(Hh, 0, P, U, false);
      ++k;
      if (P.rel_tol > 0.0f && (k % every == 0 || k == maxc)) {
        double rr = rel_residual_now(o);
        o->stats.rel_residual = (float)rr;
        if (rr <= P.rel_tol) break;
      }
    }
    o->stats.cycles = k;
    if (P.rel_tol > 0.0f) o->stats.converged = o->stats.rel_residual <= P.rel_tol;
    return;
  }
  if (P.kind == EFL_PROJECTION_MGPCG) {
    double rel = 0.0;
    uint32_t k = mgpcg(o, &rel);
    bool ok = P.rel_tol > 0.0f ? rel <= P.rel_tol : true;
    // a solve that failed may have met a sealed, incompatible region: make
    // the divergence compatible there and solve again from zero
    if (!ok && repair_sealed(o) > 0) {
      std::fill(o->p0.d.begin(), o->p0.d.end(), 0.0f);
      k += mgpcg(o, &rel);
      ok = P.rel_tol > 0.0f ? rel <= P.rel_tol : true;
    }
    o->stats.cycles = k;
    o->stats.rel_residual = (float)rel;
    o->stats.converged = ok;
    return;
  }
}

// ---- K18/K19 solve_velocity_u / _v, solve_velocity_u.wgsl:15-47, solve_velocity_v.wgsl:15-45 ----
void k_solve_velocity(orc_inst* o) {
  const Uniform& U = o->uni;
  float factor = U.dt / (U.dx * U.rho);
  Cov cu = cov_xmajor(o);
#pragma omp parallel for schedule(static) num_threads(nthreads())
  for (long y = cu.y0; y < cu.y1; ++y)
    for (long x = cu.x0; x < cu.x1; ++x) {
      float fr = o->afs.ld(x, y, 0);
      if (fr == 0.0f) { o->u0.st(x, y, o->us.ld(x, y)); continue; }
      float pij = o->p0.ld(x, y), pim = o->p0.ld(x - 1, y);
      float lp = o->ls0.ld(x, y), lm = o->ls0.ld(x - 1, y);
      if (lm >= 0.0f && lp < 0.0f) pim = lm / lp * pij;
      else if (lm < 0.0f && lp >= 0.0f) pij = lp / lm * pim;
      else if (lm >= 0.0f && lp >= 0.0f) { o->u0.st(x, y, 0.0f); continue; }
      float u = o->u1.ld(x, y);
      o->u0.st(x, y, u - factor * (pij - pim));
    }
  Cov cv = cov_ymajor(o);
#pragma omp parallel for schedule(static) num_threads(nthreads())
  for (long y = cv.y0; y < cv.y1; ++y)
    for (long x = cv.x0; x < cv.x1; ++x) {
      float fr = o->afs.ld(x, y, 2);
      if (fr == 0.0f) { o->v0.st(x, y, o->vs.ld(x, y)); continue; }
      float pij = o->p0.ld(x, y), pjm = o->p0.ld(x, y - 1);
      float lp = o->ls0.ld(x, y), lm = o->ls0.ld(x, y - 1);
      if (lm >= 0.0f && lp < 0.0f) pjm = lm / lp * pij;
      else if (lm < 0.0f && lp >= 0.0f) pij = lp / lm * pjm;
      float v = o->v1.ld(x, y);
      o->v0.st(x, y, v - factor * (pij - pjm));
    }
}

// ---- K20-K23 + loop, render_node.rs:540-608, extrapolate/*.wgsl ----
void k_extrapolate(orc_inst* o) {
  Cov cu = cov_xmajor(o), cv = cov_ymajor(o);
  // initialize_u_valid.wgsl:6-30
#pragma omp parallel for schedule(static) num_threads(nthreads())
  for (long y = cu.y0; y < cu.y1; ++y)
    for (long x = cu.x0; x < cu.x1; ++x) {
      const Tex<float>& L = o->ls0;
      float c0 = L.ld(x - 1, y - 1), c1 = L.ld(x, y - 1), c2 = L.ld(x - 1, y), c3 = L.ld(x, y);
      float c4 = L.ld(x - 1, y + 1), c5 = L.ld(x, y + 1);
      float vm = 0.25f * (((c0 + c1) + c2) + c3), vp = 0.25f * (((c2 + c3) + c4) + c5);
      o->uval[0].st(x, y, area_fraction(vm, vp) > 0.0f ? 0 : 1);
    }
  // initialize_v_valid.wgsl:6-30
#pragma omp parallel for schedule(static) num_threads(nthreads())
  for (long y = cv.y0; y < cv.y1; ++y)
    for (long x = cv.x0; x < cv.x1; ++x) {
      const Tex<float>& L = o->ls0;
      float c0 = L.ld(x - 1, y - 1), c1 = L.ld(x - 1, y), c2 = L.ld(x, y - 1), c3 = L.ld(x, y);
      float c4 = L.ld(x + 1, y - 1), c5 = L.ld(x + 1, y);
      float vm = 0.25f * (((c0 + c1) + c2) + c3), vp = 0.25f * (((c2 + c3) + c4) + c5);
      o->vval[0].st(x, y, area_fraction(vm, vp) > 0.0f ? 0 : 1);
    }
  // extrapolate_u.wgsl:5-63 / extrapolate_v.wgsl:5-61 (identical bodies)
  auto pass = [&](Tex<float>& f, const Tex<int32_t>& vin, Tex<int32_t>& vout, Cov c) {
    long dw = vin.w, dh = vin.h;
#pragma omp parallel for schedule(static) num_threads(nthreads())
    for (long y = c.y0; y < c.y1; ++y)
      for (long x = c.x0; x < c.x1; ++x) {
        if (vin.ld(x, y) == 1) { vout.st(x, y, 1); continue; }
        int count = 0;
        float nu = 0.0f;
        if (x + 1 < dw && y < dh && vin.ld(x + 1, y) == 1) { nu += f.ld(x + 1, y); count += 1; }
        if (x - 1 > 0 && y > 0 && vin.ld(x - 1, y) == 1) { nu += f.ld(x - 1, y); count += 1; }
        if (x < dw && y + 1 < dh && vin.ld(x, y + 1) == 1) { nu += f.ld(x, y + 1); count += 1; }
        if (x > 0 && y - 1 > 0 && vin.ld(x, y - 1) == 1) { nu += f.ld(x, y - 1); count += 1; }
        if (count > 0) {
          nu /= (float)count;
          f.st(x, y, nu);
          vout.st(x, y, 1);
        } else {
          vout.st(x, y, 0);
        }
      }
  };
  for (int k = 0; k < 10 / 2; ++k) {
    pass(o->u0, o->uval[0], o->uval[1], cu);
    pass(o->u0, o->uval[1], o->uval[0], cu);
    pass(o->v0, o->vval[0], o->vval[1], cv);
    pass(o->v0, o->vval[1], o->vval[0], cv);
  }
}

// ---- K24 advect_levelset with CUBIC, advect_levelset.wgsl:11-74, advect_levelset.rs:62 ----
inline float cubic1d(float y0, float y1, float y2, float y3, float t) {
  float d1 = 0.5f * (y2 - y0), d2 = 0.5f * (y3 - y1);
  float a0 = y1, a1 = d1;
  float a2 = ((-2.0f * d1) - d2) + 3.0f * (y2 - y1);
  float a3 = (d1 + d2) - 2.0f * (y2 - y1);
  return (((a3 * t * t * t) + (a2 * t * t)) + (a1 * t)) + a0;
}
void k_advect_levelset(orc_inst* o) {
  Cov c = cov_center(o);
  float dt = o->uni.dt;
  float sx = (float)o->W - 1.0f, sy = (float)o->H - 1.0f;
#pragma omp parallel for schedule(static) num_threads(nthreads())
  for (long y = c.y0; y < c.y1; ++y)
    for (long x = c.x0; x < c.x1; ++x) {
      float nx, ny;
      runge_kutta(o->u0, o->v0, (float)x, (float)y, dt, nx, ny);
      nx = wclamp(nx, 0.0f, sx);
      ny = wclamp(ny, 0.0f, sy);
      long bx = f2i(nx), by = f2i(ny);
      float tx = nx - (float)bx, ty = ny - (float)by;
      float r[4];
      for (int k = 0; k < 4; ++k) {
        long yy = by - 1 + k;
        r[k] = cubic1d(o->ls0.ld(bx - 1, yy), o->ls0.ld(bx, yy), o->ls0.ld(bx + 1, yy), o->ls0.ld(bx + 2, yy), tx);
      }
      float nl = cubic1d(r[0], r[1], r[2], r[3], ty);
      if (std::fabs(nl) < 1000.0f) o->ls1.st(x, y, nl);
    }
}

// ---- K25-K27 jump flooding, reinitialize_levelset.rs:171-210, jump_flooding.rs:200-209 ----
inline bool valid_seed(float sx, float sy) { return sx != -1.0f && sy != -1.0f; }
void k_jump_flooding(orc_inst* o) {
  Cov c = cov_center(o);
  long W = o->W, H = o->H;
  // initialize.wgsl:8-46
#pragma omp parallel for schedule(static) num_threads(nthreads())
  for (long y = c.y0; y < c.y1; ++y)
    for (long x = c.x0; x < c.x1; ++x) {
      float md = 10.0f, sx = -1.0f, sy = -1.0f;
      float level = o->ls1.ld(x, y);
      for (int i = -1; i <= 1; ++i)
        for (int j = -1; j <= 1; ++j) {
          if (i == 0 && j == 0) continue;
          long nx = x + i, ny = y + j;
          if (nx < 0 || ny < 0 || nx >= W || ny >= H) continue;
          float nl = o->ls1.ld(nx, ny);
          if (nl >= 0.0f && !(level >= 0.0f)) {
            float d = level / (level - nl);
            if (std::fabs(d) < md) {
              md = std::fabs(d);
              sx = (float)x + d * (float)i;
              sy = (float)y + d * (float)j;
            }
          }
        }
      o->seeds[0].st(x, y, sx, 0);
      o->seeds[0].st(x, y, sy, 1);
    }
  int max_power = (int)std::floor(std::log2((float)std::max(W, H)) - 1.0f);
  uint32_t step = 1u << (max_power + 1);
  int src = 0, dst = 1;
  for (int k = 0; k < max_power + 1; ++k) {
    step /= 2;
    long st = (long)step;
    Tex<float>& si = o->seeds[src];
    Tex<float>& so = o->seeds[dst];
    // iterate.wgsl:18-56
#pragma omp parallel for schedule(static) num_threads(nthreads())
    for (long y = c.y0; y < c.y1; ++y)
      for (long x = c.x0; x < c.x1; ++x) {
        float cx = si.ld(x, y, 0), cy = si.ld(x, y, 1);
        float bx = cx, by = cy, bd = 1e10f;
        if (valid_seed(cx, cy)) bd = wlen(cx - (float)x, cy - (float)y);
        for (int i = -1; i <= 1; ++i)
          for (int j = -1; j <= 1; ++j) {
            if (i == 0 && j == 0) continue;
            long nx = x + i * st, ny = y + j * st;
            if (nx < 0 || ny < 0 || nx >= si.w || ny >= si.h) continue;
            float qx = si.ld(nx, ny, 0), qy = si.ld(nx, ny, 1);
            if (!valid_seed(qx, qy)) continue;
            float d = wlen(qx - (float)x, qy - (float)y);
            if (d < bd) { bx = qx; by = qy; bd = d; }
          }
        so.st(x, y, bx, 0);
        so.st(x, y, by, 1);
      }
    std::swap(src, dst);
  }
  // calculate_sdf.wgsl:9-23
  const Tex<float>& fin = o->seeds[src];
#pragma omp parallel for schedule(static) num_threads(nthreads())
  for (long y = c.y0; y < c.y1; ++y)
    for (long x = c.x0; x < c.x1; ++x) {
      float sdf = wlen(fin.ld(x, y, 0) - (float)x, fin.ld(x, y, 1) - (float)y);
      float level = o->ls1.ld(x, y);
      float sg = level < 0.0f ? -1.0f : 1.0f;
      o->ls0.st(x, y, sdf * sg);
    }
}

// ---- K28-K30 fast iterative method, reinitialize_levelset.rs:211-241 ----
inline float abs_get_phi(const Tex<float>& phi, long x, long y) {  // update.wgsl:73-83
  if (x < 0 || y < 0 || x >= phi.w || y >= phi.h) return FIM_LARGE;
  return std::fabs(phi.at(x, y));
}
inline float solve_quadratic_2d(const Tex<float>& phi, long x, long y) {  // update.wgsl:56-71
  const float SQRT2 = std::sqrt(2.0f);
  float xm = std::fmin(abs_get_phi(phi, x - 1, y), abs_get_phi(phi, x + 1, y));
  float ym = std::fmin(abs_get_phi(phi, x, y - 1), abs_get_phi(phi, x, y + 1));
  float d = xm - ym;
  if (d > SQRT2) return ym + 1.0f;
  if (d < -SQRT2) return xm + 1.0f;
  return 0.5f * ((xm + ym) + std::sqrt(2.0f - d * d));
}

// The reference update kernel (update.wgsl:13-54) reads and writes
// neighbours' phi and labels inside one dispatch.  The oracle runs it as two
// phases per iteration, every read in a phase seeing the state at the start
// of that phase:
//   A: each ACTIVE cell c: q = solve(phi) * sign(phi_c); phi_c <- q;
//      c converged iff |phi_c - q| <= 1e-6 (then its label becomes NONE).
//   B: each cell n that was NONE at the start of the iteration and has a
//      4-neighbour (order -x, +x, -y, +y) that converged in A: q_n =
//      solve(phi after A) at n; if |phi_n| > q_n then phi_n <- q_n *
//      sign(phi of the first such neighbour before A) and n becomes ACTIVE.
void fim_update(orc_inst* o, Tex<float>& phiA, Tex<uint32_t>& flags) {
  Cov c = cov_center(o);
  Tex<float>& phi = o->ls0;
  Tex<uint32_t>& lab = o->lab;
  phiA.d = phi.d;
  flags.d.assign(flags.d.size(), 0u);
  // phase A
#pragma omp parallel for schedule(static) num_threads(nthreads())
  for (long y = c.y0; y < c.y1; ++y)
    for (long x = c.x0; x < c.x1; ++x) {
      if (lab.at(x, y) != LABEL_ACTIVE) continue;
      float p = phi.at(x, y);
      float q = solve_quadratic_2d(phi, x, y) * wsign(p);
      phiA.at(x, y) = q;
      if (!(std::fabs(p - q) > 1e-6f)) flags.at(x, y) = 1u;  // converged
    }
  // phase B
  Tex<float> phiB = phiA;
  std::vector<uint32_t> newlab = lab.d;
#pragma omp parallel for schedule(static) num_threads(nthreads())
  for (long y = c.y0; y < c.y1; ++y)
    for (long x = c.x0; x < c.x1; ++x) {
      uint32_t l = lab.at(x, y);
      size_t id = (size_t)y * phi.w + x;
      if (l == LABEL_ACTIVE) {
        if (flags.at(x, y)) newlab[id] = LABEL_NONE;
        continue;
      }
      if (l == LABEL_SOURCE) continue;
      static const int ox[4] = {-1, 1, 0, 0}, oy[4] = {0, 0, -1, 1};
      long trig = -1;
      for (int i = 0; i < 4; ++i) {
        long nx = x + ox[i], ny = y + oy[i];
        if (nx < 0 || ny < 0 || nx >= phi.w || ny >= phi.h) continue;
        // only invocations inside the dispatch cover can converge
        if (lab.at(nx, ny) == LABEL_ACTIVE && flags.at(nx, ny)) { trig = i; break; }
      }
      if (trig < 0) continue;
      float pn = std::fabs(phiA.at(x, y));
      float qn = solve_quadratic_2d(phiA, x, y);
      if (pn > qn) {
        float sg = wsign(phi.at(x + ox[trig], y + oy[trig]));
        phiB.at(x, y) = qn * sg;
        newlab[id] = LABEL_ACTIVE;
      }
    }
  phi.d.swap(phiB.d);
  lab.d.swap(newlab);
}

void k_fast_iterative(orc_inst* o) {
  Cov c = cov_center(o);
  long W = o->W, H = o->H;
  static const int ox[4] = {-1, 1, 0, 0}, oy[4] = {0, 0, -1, 1};
  // initialize.wgsl:10-40
#pragma omp parallel for schedule(static) num_threads(nthreads())
  for (long y = c.y0; y < c.y1; ++y)
    for (long x = c.x0; x < c.x1; ++x) {
      float level = o->ls1.ld(x, y);
      bool src = false;
      for (int i = 0; i < 4; ++i) {
        long nx = x + ox[i], ny = y + oy[i];
        if (nx >= 0 && ny >= 0 && nx < W && ny < H) {
          if (level * o->ls1.ld(nx, ny) <= 0.0f) { src = true; break; }
        }
      }
      if (src) {
        o->ls0.st(x, y, level);
        o->lab0.st(x, y, LABEL_SOURCE);
      } else {
        o->ls0.st(x, y, wsign(level) * FIM_LARGE);
        o->lab0.st(x, y, LABEL_NONE);
      }
    }
  // initialize_active_label.wgsl:8-40
#pragma omp parallel for schedule(static) num_threads(nthreads())
  for (long y = c.y0; y < c.y1; ++y)
    for (long x = c.x0; x < c.x1; ++x) {
      uint32_t l = o->lab0.ld(x, y);
      if (l == LABEL_SOURCE) { o->lab.st(x, y, LABEL_SOURCE); continue; }
      uint32_t out = LABEL_NONE;
      for (int i = 0; i < 4; ++i) {
        long nx = x + ox[i], ny = y + oy[i];
        if (nx >= 0 && ny >= 0 && nx < W && ny < H && o->lab0.ld(nx, ny) == LABEL_SOURCE) { out = LABEL_ACTIVE; break; }
      }
      o->lab.st(x, y, out);
    }
  Tex<float> phiA = o->ls0;
  Tex<uint32_t> flags = o->lab;
  for (uint32_t k = 0; k < o->reinit.fim_iterations; ++k) fim_update(o, phiA, flags);
}

void k_reinit(orc_inst* o) {
  if (o->reinit.kind == EFL_REINIT_JUMP_FLOODING) k_jump_flooding(o);
  else k_fast_iterative(o);
}

// ---- K31 levelset_gradient, levelset_gradient.wgsl:4-21 ----
void k_levelset_gradient(orc_inst* o) {
  Cov c = cov_center(o);
#pragma omp parallel for schedule(static) num_threads(nthreads())
  for (long y = c.y0; y < c.y1; ++y)
    for (long x = c.x0; x < c.x1; ++x) {
      if (x == 0 || y == 0) continue;
      float gx = 0.5f * (o->ls0.ld(x + 1, y) - o->ls0.ld(x - 1, y));
      float gy = 0.5f * (o->ls0.ld(x, y + 1) - o->ls0.ld(x, y - 1));
      o->grad.st(x, y, gx, 0);
      o->grad.st(x, y, gy, 1);
    }
}

// ---- K32/K33 sample_forces_to_solid + accumulate_forces, fluid_to_solid/*.wgsl;
//      bins cleared first as initialize_buffer does each frame (fluid_to_solid.rs:211-229) ----
inline int32_t f32_to_i32(float v) { return (int32_t)f2i(v * (1.0f / 0.1f)); }
void k_fluid_to_solid(orc_inst* o) {
  std::fill(o->bins.d.begin(), o->bins.d.end(), 0);
  Cov c = cov_center(o);
  const Uniform& U = o->uni;
  uint32_t nob = (uint32_t)o->obstacles.size();
  for (long y = c.y0; y < c.y1; ++y)
    for (long x = c.x0; x < c.x1; ++x) {
      AF f = area_fractions(o->lss, x, y);
      if (f.im == 0.0f && f.ip == 0.0f && f.jm == 0.0f && f.jp == 0.0f) continue;
      int32_t id = o->sid.ld(x, y);
      if (id == -1) continue;
      float p = o->p0.ld(x, y);
      float fx = ((-(f.ip - f.im)) * p) * U.dx;
      float fy = ((-(f.jp - f.jm)) * p) * U.dx;
      efl_obstacle ob;
      std::memset(&ob, 0, sizeof(ob));
      if (id >= 0 && (uint32_t)id < nob) ob = o->obstacles[id];
      float cx, cy;
      center_of_mass(ob, cx, cy);
      float uvx = (float)x / U.size[0], uvy = (float)y / U.size[1];
      float xx = (uvx - 0.5f) * U.size[0], yy = (-uvy + 0.5f) * U.size[1];
      float wx, wy;
      mat_xy(U.T, xx, yy, 0.0f, 1.0f, wx, wy);
      float rx = wx - cx, ry = wy - cy;
      float tq = rx * fy - ry * fx;
      // no counterpart: torque about the centre in the world frame (grid y
      // points down, so the world force is (fx, -fy))
      float tw = -rx * fy - ry * fx;
      if (id >= 0 && id < EFL_MAX_SOLIDS) {  // atomicAdd into bins (wrapping i32 add)
        o->bins.at(id, 0) = (int32_t)((uint32_t)o->bins.at(id, 0) + (uint32_t)f32_to_i32(fx));
        o->bins.at(id, 1) = (int32_t)((uint32_t)o->bins.at(id, 1) + (uint32_t)f32_to_i32(fy));
        o->bins.at(id, 2) = (int32_t)((uint32_t)o->bins.at(id, 2) + (uint32_t)f32_to_i32(tq));
        o->bins.at(id, 3) = (int32_t)((uint32_t)o->bins.at(id, 3) + (uint32_t)f32_to_i32(tw));
      }
    }
  for (int i = 0; i < EFL_MAX_SOLIDS; ++i) {  // accumulate_forces.wgsl:16-32
    o->fts.at(i, 0, 0) = (float)o->bins.at(i, 0) * 0.1f;
    o->fts.at(i, 0, 1) = (float)o->bins.at(i, 1) * 0.1f;
    o->fts.at(i, 0, 2) = (float)o->bins.at(i, 2) * 0.1f;
    o->fts.at(i, 0, 3) = (float)o->bins.at(i, 3) * 0.1f;
  }
}

void measure_state(orc_inst* o, float dt, float* rel, float* mdiv);

// Rigid bodies after the step (the role avian2d plays for the reference):
// a = (F / dx^2 + m_a a_prev + m g / dx) / (m + m_a) in cells/s^2 with
// m = density * area, m_a = 4 rho area (an added-mass bound, with the previous
// acceleration, which keeps explicit coupling of light bodies stable) and the
// same form for the world-frame torque with J = density * inertia,
// J_a = 2 rho inertia; bodies held inside the wall ring.
void integrate_bodies(orc_inst* o, float dt) {
  size_t n = o->obstacles.size();
  if (o->bodies.size() != n || n == 0) return;
  if (o->body_state.size() != n) o->body_state.assign(n, {0.0f, 0.0f, 0.0f});
  float dx = o->uni.dx;
  for (size_t k = 0; k < n; ++k) {
    efl_obstacle& ob = o->obstacles[k];
    const efl_body& bd = o->bodies[k];
    if (!bd.dynamic) continue;
    float m = bd.density * bd.area, J = bd.density * bd.inertia;
    float ma = 4.0f * o->uni.rho * bd.area, Ja = 2.0f * o->uni.rho * bd.inertia;
    float Fx = 0.0f, Fy = 0.0f, Tq = 0.0f;
    uint32_t id = ob.entity_id;
    if (id < (uint32_t)EFL_MAX_SOLIDS) {
      float s = 1.0f / (dx * dx);
      Fx = o->fts.at(id, 0, 0) * s;
      Fy = -o->fts.at(id, 0, 1) * s;
      Tq = o->fts.at(id, 0, 3) * s;
    }
    std::array<float, 3>& prev = o->body_state[k];
    float ax = (Fx + ma * prev[0] + m * o->uni.g[0] / dx) / (m + ma);
    float ay = (Fy + ma * prev[1] + m * (-o->uni.g[1]) / dx) / (m + ma);
    float al = (J + Ja) > 0.0f ? (Tq + Ja * prev[2]) / (J + Ja) : 0.0f;
    prev = {ax, ay, al};
    float vx = ob.linear_velocity[0] + ax * dt;
    float vy = ob.linear_velocity[1] + ay * dt;
    float w = ob.angular_velocity + al * dt;
    const float* v = ob.values;
    float R = 0.0f;
    switch (ob.shape) {
      case 0: R = v[0]; break;
      case 1: R = wlen(v[0], v[1]); break;
      case 2: R = std::fmax(wlen(v[0], v[1]), wlen(v[2], v[3])) + v[4]; break;
      case 4: R = std::fmax(std::fmax(wlen(v[0], v[1]), wlen(v[2], v[3])), wlen(v[4], v[5])); break;
      default: break;
    }
    // at most one cell per step, translating and rotating at the bounding radius
    const float vmax = 1.0f / dt;
    const float sp = wlen(vx, vy);
    if (sp > vmax) {
      const float sc = vmax / sp;
      vx = vx * sc;
      vy = vy * sc;
    }
    if (R > 0.0f && std::fabs(w) * R > vmax) w = std::copysign(vmax / R, w);
    float tx = ob.transform[12] + vx * dt, ty = ob.transform[13] + vy * dt;
    float ang = std::atan2(ob.transform[1], ob.transform[0]) + w * dt;
    float lx = std::fmax(0.5f * (float)o->W - 1.0f - R, 0.0f), ly = std::fmax(0.5f * (float)o->H - 1.0f - R, 0.0f);
    if (tx > lx) { tx = lx; vx = std::fmin(vx, 0.0f); }
    if (tx < -lx) { tx = -lx; vx = std::fmax(vx, 0.0f); }
    if (ty > ly) { ty = ly; vy = std::fmin(vy, 0.0f); }
    if (ty < -ly) { ty = -ly; vy = std::fmax(vy, 0.0f); }
    float c = std::cos(ang), sn = std::sin(ang);
    float T[16] = {c, sn, 0.f, 0.f, -sn, c, 0.f, 0.f, 0.f, 0.f, 1.f, 0.f, tx, ty, 0.f, 1.f};
    float Ti[16] = {c, -sn, 0.f, 0.f, sn, c, 0.f, 0.f, 0.f, 0.f, 1.f, 0.f, -(c * tx + sn * ty), -(-sn * tx + c * ty), 0.f, 1.f};
    std::memcpy(ob.transform, T, sizeof(T));
    std::memcpy(ob.inverse_transform, Ti, sizeof(Ti));
    ob.linear_velocity[0] = vx;
    ob.linear_velocity[1] = vy;
    ob.angular_velocity = w;
  }
}

// EulerFluidNode::run Uninitialized branch, render_node.rs:219-277
void initialize_all(orc_inst* o) {
  k_init_center(o);
  k_init_edge(o);
  k_fluid_source(o, o->src_init);
  if (!o->custom_ls.empty()) {
    for (size_t i = 0; i < o->ls1.d.size(); ++i) o->ls1.d[i] = std::fmin(o->ls1.d[i], o->custom_ls[i]);
  }
  k_reinit(o);
  o->initialized = true;
  o->stats.steps = 0;
}

// EulerFluidNode::run Initialized branch, render_node.rs:278-466
void step_all(orc_inst* o, float dt) {
  o->uni.dt = dt;
  k_update_solid(o);
  k_area_fraction(o);
  k_advect_velocity(o);
  k_apply_forces(o);
  k_divergence(o);
  k_projection(o);
  if (o->diag && o->stats.rel_residual < 0.0f) o->stats.rel_residual = (float)rel_residual_now(o);
  k_solve_velocity(o);
  if (o->diag) measure_state(o, dt, nullptr, &o->stats.max_norm_divergence);
  k_extrapolate(o);
  k_advect_levelset(o);
  k_fluid_source(o, o->src_step);
  k_reinit(o);
  k_levelset_gradient(o);
  k_fluid_to_solid(o);
  o->forces.clear();  // apply_forces_and_clear, apply_forces.rs:157-167
  o->stats.steps += 1;
}

template <class T>
struct FieldRef { Tex<T>* t; };

void* field_ptr(orc_inst* o, int f, size_t* count, uint32_t* w, uint32_t* h, uint32_t* c) {
  auto set = [&](auto& t) -> void* {
    if (count) *count = t.d.size();
    if (w) *w = t.w;
    if (h) *h = t.h;
    if (c) *c = t.c;
    return (void*)t.d.data();
  };
  switch (f) {
    case EFL_FIELD_U0: return set(o->u0);
    case EFL_FIELD_U1: return set(o->u1);
    case EFL_FIELD_V0: return set(o->v0);
    case EFL_FIELD_V1: return set(o->v1);
    case EFL_FIELD_U_SOLID: return set(o->us);
    case EFL_FIELD_V_SOLID: return set(o->vs);
    case EFL_FIELD_SOLID_ID: return set(o->sid);
    case EFL_FIELD_U_VALID0: return set(o->uval[0]);
    case EFL_FIELD_U_VALID1: return set(o->uval[1]);
    case EFL_FIELD_V_VALID0: return set(o->vval[0]);
    case EFL_FIELD_V_VALID1: return set(o->vval[1]);
    case EFL_FIELD_DIV: return set(o->div);
    case EFL_FIELD_P0: return set(o->p0);
    case EFL_FIELD_P1: return set(o->p1);
    case EFL_FIELD_LS_AIR0: return set(o->ls0);
    case EFL_FIELD_LS_AIR1: return set(o->ls1);
    case EFL_FIELD_GRAD_LS: return set(o->grad);
    case EFL_FIELD_LS_SOLID: return set(o->lss);
    case EFL_FIELD_AREA_FRACTION: return set(o->afs);
    case EFL_FIELD_FIM_LABELS0: return set(o->lab0);
    case EFL_FIELD_FIM_LABELS: return set(o->lab);
    case EFL_FIELD_JFA_SEEDS0: return set(o->seeds[0]);
    case EFL_FIELD_JFA_SEEDS1: return set(o->seeds[1]);
    case EFL_FIELD_SOLID_BINS: return set(o->bins);
    case EFL_FIELD_SOLID_FORCES: return set(o->fts);
    case EFL_FIELD_MG_R0: return set(o->mr[0]);
    default: return nullptr;
  }
}

void measure_state(orc_inst* o, float dt, float* rel, float* mdiv) {
  if (rel) *rel = (float)rel_residual_now(o);
  if (mdiv) {
    // K10 numerator over fluid, non-solid cells, on the post-projection u0/v0
    float m = 0.0f;
    for (long y = 0; y < o->H; ++y)
      for (long x = 0; x < o->W; ++x) {
        if (!(o->ls0.at(x, y) < 0.0f)) continue;
        AF f = area_fractions(o->lss, x, y);
        if (f.im == 0.0f && f.ip == 0.0f && f.jm == 0.0f && f.jp == 0.0f) continue;
        float du_f = f.ip * o->u0.ld(x + 1, y) - f.im * o->u0.ld(x, y);
        float dv_f = f.jp * o->v0.ld(x, y + 1) - f.jm * o->v0.ld(x, y);
        float du_s = (1.0f - f.ip) * o->us.ld(x + 1, y) - (1.0f - f.im) * o->us.ld(x, y);
        float dv_s = (1.0f - f.jp) * o->vs.ld(x, y + 1) - (1.0f - f.jm) * o->vs.ld(x, y);
        float d = std::fabs(((du_f + dv_f) + du_s) + dv_s);
        if (d > m) m = d;
      }
    *mdiv = dt * m;
  }
}

}  // namespace

// ------------------------------------------------------------------------
// C API
// ------------------------------------------------------------------------
extern "C" {

int orc_set_threads(int n) {
  g_threads = n;
  return 0;
}

orc_inst* orc_create(const efl_settings* s) {
  if (!s || s->width < 8 || s->height < 8 || !(s->length_unit > 0.0f)) return nullptr;
  orc_inst* o = new orc_inst();
  int W = (int)s->width, H = (int)s->height;
  o->W = W;
  o->H = H;
  o->uni.dx = 1.0f / s->length_unit;
  o->uni.dt = 0.0f;
  o->uni.rho = s->rho;
  o->uni.g[0] = s->gravity[0];
  o->uni.g[1] = s->gravity[1];
  bool zero = true;
  for (int i = 0; i < 16; ++i) zero = zero && s->transform[i] == 0.0f;
  for (int i = 0; i < 16; ++i) o->uni.T[i] = zero ? ((i % 5) == 0 ? 1.0f : 0.0f) : s->transform[i];
  o->uni.size[0] = (float)W;
  o->uni.size[1] = (float)H;
  o->proj = s->projection;
  o->reinit = s->reinit;
  o->u0.alloc(W + 1, H); o->u1.alloc(W + 1, H); o->us.alloc(W + 1, H);
  o->v0.alloc(W, H + 1); o->v1.alloc(W, H + 1); o->vs.alloc(W, H + 1);
  o->uval[0].alloc(W + 1, H); o->uval[1].alloc(W + 1, H);
  o->vval[0].alloc(W, H + 1); o->vval[1].alloc(W, H + 1);
  o->sid.alloc(W, H);
  o->div.alloc(W, H); o->p0.alloc(W, H); o->p1.alloc(W, H);
  o->ls0.alloc(W, H); o->ls1.alloc(W, H); o->lss.alloc(W, H);
  o->grad.alloc(W, H, 2); o->afs.alloc(W, H, 4);
  o->seeds[0].alloc(W, H, 2); o->seeds[1].alloc(W, H, 2);
  o->lab0.alloc(W, H); o->lab.alloc(W, H);
  o->bins.alloc(EFL_MAX_SOLIDS, 4);  // planar: force x, force y, torque, world torque
  o->fts.alloc(EFL_MAX_SOLIDS, 1, 4);
  o->pcg_r.alloc(W, H); o->pcg_z.alloc(W, H); o->pcg_p.alloc(W, H); o->pcg_q.alloc(W, H);
  // multi_grid.rs:95-118
  int mn = std::min(W, H);
  int lg = (int)std::log2((float)mn);
  int nl = std::max(1, lg - 2);
  o->nlev = nl;
  o->mx.resize(nl); o->mb.resize(nl); o->mls.resize(nl); o->mafs.resize(nl); o->mr.resize(nl);
  o->mr[0].alloc(W, H);
  int w = W, h = H;
  for (int i = 1; i < nl; ++i) {
    w /= 2; h /= 2;
    o->mx[i].alloc(w, h); o->mb[i].alloc(w, h); o->mls[i].alloc(w, h);
    o->mafs[i].alloc(w, h, 4); o->mr[i].alloc(w, h);
  }
  return o;
}

void orc_destroy(orc_inst* o) { delete o; }

int orc_set_projection(orc_inst* o, const efl_projection* p) { o->proj = *p; return 0; }
int orc_set_reinit(orc_inst* o, const efl_reinit* r) { o->reinit = *r; return 0; }
int orc_set_uniforms(orc_inst* o, float rho, const float g[2], const float T[16]) {
  o->uni.rho = rho;
  if (g) { o->uni.g[0] = g[0]; o->uni.g[1] = g[1]; }
  if (T) for (int i = 0; i < 16; ++i) o->uni.T[i] = T[i];
  return 0;
}
int orc_set_sources(orc_inst* o, const efl_source* s, uint32_t n, int oneshot) {
  if (n > EFL_MAX_FLUID_SOURCES) return EFL_E_RANGE;
  std::vector<efl_source>& dst = oneshot ? o->src_init : o->src_step;
  dst.assign(s, s + n);
  return 0;
}
int orc_set_obstacles(orc_inst* o, const efl_obstacle* ob, uint32_t n) {
  if (n > EFL_MAX_SOLIDS) return EFL_E_RANGE;
  o->obstacles.assign(ob, ob + n);
  return 0;
}
int orc_set_forces(orc_inst* o, const efl_force* f, uint32_t n) {
  o->forces.assign(f, f + n);
  return 0;
}
int orc_init_levelset(orc_inst* o, const float* ls) {
  if (!ls) o->custom_ls.clear();
  else o->custom_ls.assign(ls, ls + (size_t)o->W * o->H);
  return 0;
}
int orc_reset(orc_inst* o) { o->initialized = false; return 0; }
int orc_initialize(orc_inst* o) { initialize_all(o); return 0; }
int orc_step(orc_inst* o, float dt) {
  if (!o->initialized) {
    initialize_all(o);
  } else {
    step_all(o, dt);
    integrate_bodies(o, dt);
  }
  return 0;
}

int orc_set_bodies(orc_inst* o, const efl_body* b, uint32_t n) {
  if (n && n != o->obstacles.size()) return EFL_E_INVALID;
  o->bodies.assign(b, b + n);
  return 0;
}

int orc_get_obstacles(orc_inst* o, efl_obstacle* out, uint32_t n) {
  if (n > o->obstacles.size()) return EFL_E_RANGE;
  std::memcpy(out, o->obstacles.data(), sizeof(efl_obstacle) * n);
  return 0;
}

int orc_run_stage(orc_inst* o, int stage, float dt) {
  o->uni.dt = dt;
  switch (stage) {
    case EFL_STAGE_INIT_GRID_CENTER: k_init_center(o); break;
    case EFL_STAGE_INIT_GRID_EDGE: k_init_edge(o); break;
    case EFL_STAGE_SOURCE_ONESHOT: k_fluid_source(o, o->src_init); break;
    case EFL_STAGE_UPDATE_SOLID: k_update_solid(o); break;
    case EFL_STAGE_AREA_FRACTION: k_area_fraction(o); break;
    case EFL_STAGE_ADVECT_VELOCITY: k_advect_velocity(o); break;
    case EFL_STAGE_APPLY_FORCES: k_apply_forces(o); break;
    case EFL_STAGE_DIVERGENCE: k_divergence(o); break;
    case EFL_STAGE_PROJECTION: k_projection(o); break;
    case EFL_STAGE_SOLVE_VELOCITY: k_solve_velocity(o); break;
    case EFL_STAGE_EXTRAPOLATE_VELOCITY: k_extrapolate(o); break;
    case EFL_STAGE_ADVECT_LEVELSET: k_advect_levelset(o); break;
    case EFL_STAGE_SOURCE: k_fluid_source(o, o->src_step); break;
    case EFL_STAGE_REINITIALIZE: k_reinit(o); break;
    case EFL_STAGE_LEVELSET_GRADIENT: k_levelset_gradient(o); break;
    case EFL_STAGE_FLUID_TO_SOLID: k_fluid_to_solid(o); break;
    case EFL_STAGE_GS_SWEEP:
    case EFL_STAGE_GS_SWEEP_SPLIT: {
      Lvl<float> L{o->W, o->H, o->p0.d.data(), o->div.d.data(), o->ls0.d.data(), o->afs.d.data(), o->mr[0].d.data(), 1.0f};
      Cov c = cov_center(o);
      gs_color(L, 1, 1.9f, o->uni, c.x1, c.y1);
      gs_color(L, 0, 1.9f, o->uni, c.x1, c.y1);
      break;
    }
    case EFL_STAGE_MG_SMOOTH: {  // two red-black sweeps at weight 1 on the whole level 0
      Lvl<float> L{o->W, o->H, o->p0.d.data(), o->div.d.data(), o->ls0.d.data(), o->afs.d.data(), o->mr[0].d.data(), 1.0f};
      for (int k = 0; k < 2; ++k) {
        gs_color(L, 1, 1.0f, o->uni, o->W, o->H);
        gs_color(L, 0, 1.0f, o->uni, o->W, o->H);
      }
      break;
    }
    case EFL_STAGE_MG_VCYCLE: {
      Hier<float> Hh = float_hier(o, o->p0.d.data(), o->div.d.data(), o->mr[0].d.data());
      v_cycle(Hh, 0, o->proj, o->uni, false);
      break;
    }
    case EFL_STAGE_MG_RESIDUAL: {
      Lvl<float> L{o->W, o->H, o->p0.d.data(), o->div.d.data(), o->ls0.d.data(), o->afs.d.data(), o->mr[0].d.data(), 1.0f};
      residual_level(L, o->uni);
      break;
    }
    default:
      return EFL_E_INVALID;
  }
  return 0;
}

int orc_field_info(const orc_inst* o, int f, size_t* count, uint32_t* w, uint32_t* h, uint32_t* c) {
  return field_ptr(const_cast<orc_inst*>(o), f, count, w, h, c) ? 0 : EFL_E_INVALID;
}
int orc_get_field(const orc_inst* o, int f, void* out, size_t count) {
  size_t n = 0;
  void* p = field_ptr(const_cast<orc_inst*>(o), f, &n, nullptr, nullptr, nullptr);
  if (!p || count != n) return EFL_E_INVALID;
  std::memcpy(out, p, n * 4);
  return 0;
}
int orc_set_field(orc_inst* o, int f, const void* in, size_t count) {
  size_t n = 0;
  void* p = field_ptr(o, f, &n, nullptr, nullptr, nullptr);
  if (!p || count != n) return EFL_E_INVALID;
  std::memcpy(p, in, n * 4);
  return 0;
}
int orc_get_stats(const orc_inst* o, efl_stats* out) { *out = o->stats; return 0; }
int orc_set_diagnostics(orc_inst* o, int enabled) { o->diag = enabled != 0; return 0; }
int orc_measure(orc_inst* o, float dt, float* rel, float* mdiv) {
  measure_state(o, dt, rel, mdiv);
  return 0;
}

// Fluid diagnostics (diagnostics/plugin.rs:20-104): the volume
// (calculate_volume.wgsl:12-75: cells with levelset_air < 0), and the largest
// and smallest cell-centred speed (max_velocity.wgsl:13-81, min_velocity.wgsl:
// 13-87: |(u(x)+u(x+1), v(y)+v(y+1)) / 2|, the minimum over cells with
// levelset_air < 0, others 1e30), each in two passes.  Pass 1 runs one 16 x 16
// workgroup per block of the floor(W/16) x floor(H/16) cover
// (calculate_volume.rs:51,133; max_velocity.rs:53,135; min_velocity.rs:55,
// 138) and folds its 256 values by halving strides 128..1.  Pass 2 is one
// workgroup of 512: lane l adds the partials l, l + 512, ... in order (the
// velocity passes add their partial maxima and minima as well) and the 512
// lane values fold by halving strides 256..1 (sum for the volume, max / min
// for the speeds).  FluidTextures.u / v / levelset_air are u0 / v0 / ls0
// (setup_components.rs:115-122).
int orc_fluid_diagnostics(const orc_inst* o, float* out) {
  const int W = o->W, H = o->H, gx = W / 16, gy = H / 16, np = gx * gy;
  std::vector<float> pv(np), px(np), pn(np);
  for (int by = 0; by < gy; ++by)
    for (int bx = 0; bx < gx; ++bx) {
      float wv[256], wx[256], wn[256];
      for (int lid = 0; lid < 256; ++lid) {
        const int x = bx * 16 + lid % 16, y = by * 16 + lid / 16;
        const float level = o->ls0.at(x, y);
        const float uc = 0.5f * (o->u0.at(x, y) + o->u0.at(x + 1, y));
        const float vc = 0.5f * (o->v0.at(x, y) + o->v0.at(x, y + 1));
        const float mag = std::sqrt(uc * uc + vc * vc);
        wv[lid] = level < 0.0f ? 1.0f : 0.0f;
        wx[lid] = mag;
        wn[lid] = level < 0.0f ? mag : 1e30f;
      }
      for (int stride = 128;; stride /= 2) {
        for (int lid = 0; lid < stride; ++lid) {
          wv[lid] += wv[lid + stride];
          wx[lid] = std::fmax(wx[lid], wx[lid + stride]);
          wn[lid] = std::fmin(wn[lid], wn[lid + stride]);
        }
        if (stride == 1) break;
      }
      const int wid = bx + by * gx;
      pv[wid] = wv[0];
      px[wid] = wx[0];
      pn[wid] = wn[0];
    }
  float tv[512], tx[512], tn[512];
  for (int l = 0; l < 512; ++l) {
    float sv = 0.0f, sx = 0.0f, sn = 0.0f;
    for (int i = l; i < np; i += 512) {
      sv += pv[i];
      sx += px[i];
      sn += pn[i];
    }
    tv[l] = sv;
    tx[l] = sx;
    tn[l] = sn;
  }
  for (int stride = 256; stride > 0; stride /= 2)
    for (int l = 0; l < stride; ++l) {
      tv[l] += tv[l + stride];
      tx[l] = std::fmax(tx[l], tx[l + stride]);
      tn[l] = std::fmin(tn[l], tn[l + stride]);
    }
  out[0] = tv[0];
  out[1] = tn[0];
  out[2] = tx[0];
  return 0;
}

namespace {
Tex<float>* mg_tex(orc_inst* o, int level, int which) {
  if (level < 0 || level >= o->nlev) return nullptr;
  if (level == 0) {
    switch (which) {
      case 0: return &o->p0;
      case 1: return &o->div;
      case 2: return &o->ls0;
      case 3: return &o->afs;
      case 4: return &o->mr[0];
      default: return nullptr;
    }
  }
  switch (which) {
    case 0: return &o->mx[level];
    case 1: return &o->mb[level];
    case 2: return &o->mls[level];
    case 3: return &o->mafs[level];
    case 4: return &o->mr[level];
    default: return nullptr;
  }
}
}  // namespace

int orc_mg_level_size(const orc_inst* o, int level, uint32_t* w, uint32_t* h) {
  if (level < 0 || level >= o->nlev) return -1;
  *w = (uint32_t)(level == 0 ? o->W : o->mx[level].w);
  *h = (uint32_t)(level == 0 ? o->H : o->mx[level].h);
  return 0;
}

int orc_mg_level(orc_inst* o, int level, int which, float* data, size_t count, int set) {
  Tex<float>* t = mg_tex(o, level, which);
  if (!t || count != t->d.size()) return -1;
  if (set) std::copy(data, data + count, t->d.begin());
  else std::copy(t->d.begin(), t->d.end(), data);
  return 0;
}

int orc_mg_restrict(orc_inst* o, int level) {
  if (level < 0 || level + 1 >= o->nlev) return -1;
  Hier<float> Hh = float_hier(o, o->p0.d.data(), o->div.d.data(), o->mr[0].d.data());
  Lvl<float> L{Hh.w[level], Hh.h[level], Hh.x[level], Hh.b[level], Hh.ls[level], Hh.af[level], Hh.r[level],
               (float)(1 << level)};
  restrict_level(L, Hh.w[level + 1], Hh.h[level + 1], Hh.bmut[level + 1], Hh.ls[level + 1], Hh.af[level + 1],
                 Hh.x[level + 1]);
  return 0;
}

int orc_mg_prolong(orc_inst* o, int level) {
  if (level < 0 || level + 1 >= o->nlev) return -1;
  Hier<float> Hh = float_hier(o, o->p0.d.data(), o->div.d.data(), o->mr[0].d.data());
  Lvl<float> L{Hh.w[level], Hh.h[level], Hh.x[level], Hh.b[level], Hh.ls[level], Hh.af[level], Hh.r[level],
               (float)(1 << level)};
  prolong_level(L, Hh.w[level + 1], Hh.h[level + 1], Hh.x[level + 1]);
  return 0;
}

int orc_residual_of(orc_inst* o, float dt, const float* x, float* r_out) {
  o->uni.dt = dt;
  Lvl<float> L{o->W, o->H, const_cast<float*>(x), o->div.d.data(), o->ls0.d.data(), o->afs.d.data(), r_out, 1.0f};
  residual_level(L, o->uni);
  return 0;
}

// FP64 solve of the same discrete system: CG preconditioned by a double
// precision symmetric V-cycle over the same (float) geometry pyramid.
int orc_solve_fp64(orc_inst* o, float dt, double rel_tol, double* p_out, double* rel_out, int* it_out) {
  o->uni.dt = dt;
  int W = o->W, H = o->H;
  size_t n = (size_t)W * H;
  int nl = o->nlev;
  // geometry pyramid (float, restriction.wgsl), built by a float restriction pass
  std::vector<std::vector<double>> X(nl), B(nl), R(nl);
  std::vector<std::vector<float>> LS(nl), AFv(nl);
  std::vector<int> w(nl), h(nl);
  w[0] = W; h[0] = H;
  LS[0] = o->ls0.d; AFv[0] = o->afs.d;
  for (int i = 1; i < nl; ++i) { w[i] = w[i - 1] / 2; h[i] = h[i - 1] / 2; }
  for (int i = 0; i < nl; ++i) {
    X[i].assign((size_t)w[i] * h[i], 0.0);
    B[i].assign((size_t)w[i] * h[i], 0.0);
    R[i].assign((size_t)w[i] * h[i], 0.0);
    if (i > 0) { LS[i].assign((size_t)w[i] * h[i], 0.0f); AFv[i].assign((size_t)w[i] * h[i] * 4, 0.0f); }
  }
  Hier<double> Hh;
  Hh.nlev = nl;
  for (int i = 0; i < nl; ++i) {
    Hh.w.push_back(w[i]); Hh.h.push_back(h[i]);
    Hh.x.push_back(X[i].data()); Hh.b.push_back(B[i].data()); Hh.bmut.push_back(B[i].data());
    Hh.ls.push_back(LS[i].data()); Hh.af.push_back(AFv[i].data()); Hh.r.push_back(R[i].data());
  }
  efl_projection cfg = o->proj;
  if (cfg.mg_pre_iterations == 0) cfg.mg_pre_iterations = 2;
  if (cfg.mg_post_iterations == 0) cfg.mg_post_iterations = 2;
  if (cfg.mg_coarsest_iterations == 0) cfg.mg_coarsest_iterations = 3;
  auto act = [&](size_t i) { return active_cell(o, i); };
  std::vector<double> bb(n), x(n, 0.0), r(n), z(n), p(n), q(n), zero(n, 0.0);
  for (size_t i = 0; i < n; ++i) bb[i] = act(i) ? (double)o->div.d[i] : 0.0;
  auto Aop = [&](const std::vector<double>& in, std::vector<double>& out) {
    Lvl<double> L{W, H, const_cast<double*>(in.data()), zero.data(), LS[0].data(), AFv[0].data(), out.data(), 1.0};
    residual_level(L, o->uni);
    for (size_t i = 0; i < n; ++i) out[i] = act(i) ? -out[i] : 0.0;
  };
  auto M = [&](const std::vector<double>& in, std::vector<double>& out) {
    std::fill(X[0].begin(), X[0].end(), 0.0);
    Hh.b[0] = in.data();
    v_cycle(Hh, 0, cfg, o->uni, true);
    for (size_t i = 0; i < n; ++i) out[i] = act(i) ? X[0][i] : 0.0;
  };
  auto dot = [&](const std::vector<double>& a, const std::vector<double>& b) {
    double s = 0.0;
    for (size_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
  };
  double bn = std::sqrt(dot(bb, bb));
  r = bb;
  double rel = 1.0;
  int it = 0;
  if (bn > 0.0) {
    M(r, z);
    p = z;
    double rz = dot(r, z);
    for (it = 0; it < 2000; ++it) {
      Aop(p, q);
      double pq = dot(p, q);
      if (!(pq > 0.0)) break;
      double alpha = rz / pq;
      for (size_t i = 0; i < n; ++i) { x[i] += alpha * p[i]; r[i] -= alpha * q[i]; }
      rel = std::sqrt(dot(r, r)) / bn;
      if (rel <= rel_tol) { ++it; break; }
      M(r, z);
      double rz2 = dot(r, z);
      double beta = rz2 / rz;
      rz = rz2;
      for (size_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    }
  } else {
    rel = 0.0;
  }
  std::copy(x.begin(), x.end(), p_out);
  if (rel_out) *rel_out = rel;
  if (it_out) *it_out = it;
  return 0;
}

}  // extern "C"
