// euler_fluid.cpp - FluidWorld and the conversions of the C++ host mirror.
// Every call into the device goes through the C ABI (include/efl.h); nothing
// here computes fluid state.
#include "euler_fluid.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

namespace bevy_eulerian_fluid {

namespace {

void warn(const char* msg) { std::fprintf(stderr, "WARN bevy_eulerian_fluid: %s\n", msg); }

}  // namespace

// ---------------------------------------------------------------- Transform

Mat4 Transform::to_matrix() const {
  const float c = std::cos(rotation), s = std::sin(rotation);
  Mat4 m{};
  m[0] = c * scale.x;
  m[1] = s * scale.x;
  m[4] = -s * scale.y;
  m[5] = c * scale.y;
  m[10] = 1.0f;
  m[12] = translation.x;
  m[13] = translation.y;
  m[15] = 1.0f;
  return m;
}

Mat4 Transform::inverse_matrix() const {
  // (T R S)^-1 = S^-1 R^T T^-1
  const float c = std::cos(rotation), s = std::sin(rotation);
  const float ix = 1.0f / scale.x, iy = 1.0f / scale.y;
  Mat4 m{};
  m[0] = c * ix;
  m[1] = -s * iy;
  m[4] = s * ix;
  m[5] = c * iy;
  m[10] = 1.0f;
  m[12] = -(c * translation.x + s * translation.y) * ix;
  m[13] = -(-s * translation.x + c * translation.y) * iy;
  m[15] = 1.0f;
  return m;
}

// -------------------------------------------------------------- FluidPlugin

FluidPlugin::FluidPlugin(float length_unit) : length_unit_(length_unit) {
  if (!(length_unit > 0.0f)) throw FluidError("length_unit must be positive value.");
}

// ------------------------------------------------------- ProjectionMethod

std::string ProjectionMethod::to_string() const {
  switch (kind_) {
    case Kind::Jacobi: return "Jacobi";
    case Kind::GaussSeidel: return "Gauss-Seidel(" + std::to_string(gauss_seidel_.num_iterations) + ")";
    case Kind::MultiGrid:
    case Kind::MgPcg: {
      std::string s = kind_ == Kind::MultiGrid ? "Multigrid(" : "MG-PCG(";
      return s + std::to_string(multi_grid_.pre_smooth_config.num_iterations) + "-" +
             std::to_string(multi_grid_.coarsest_config.num_iterations) + "-" +
             std::to_string(multi_grid_.post_smooth_config.num_iterations) + ")";
    }
  }
  return "?";
}

efl_projection ProjectionMethod::to_c() const {
  efl_projection p{};
  p.gs_iterations = gauss_seidel_.num_iterations;
  p.mg_pre_iterations = multi_grid_.pre_smooth_config.num_iterations;
  p.mg_post_iterations = multi_grid_.post_smooth_config.num_iterations;
  p.mg_coarsest_iterations = multi_grid_.coarsest_config.num_iterations;
  p.max_cycles = target_.max_cycles;
  p.rel_tol = target_.rel_tol;
  p.check_every = target_.check_every;
  switch (kind_) {
    case Kind::Jacobi: p.kind = EFL_PROJECTION_JACOBI; break;
    case Kind::GaussSeidel: p.kind = EFL_PROJECTION_GAUSS_SEIDEL; break;
    case Kind::MultiGrid: p.kind = EFL_PROJECTION_MULTIGRID; break;
    case Kind::MgPcg: p.kind = EFL_PROJECTION_MGPCG; break;
  }
  return p;
}

std::string ReinitializeMethod::to_string() const {
  return jfa_ ? std::string("Jump Flooding") : "Fast Iterative(" + std::to_string(fim_.num_iterations) + ")";
}

efl_reinit ReinitializeMethod::to_c() const {
  efl_reinit r{};
  r.kind = jfa_ ? EFL_REINIT_JUMP_FLOODING : EFL_REINIT_FAST_ITERATIVE;
  r.fim_iterations = fim_.num_iterations;
  return r;
}

// ------------------------------------------------------------------ solids

ShapeVariant ShapeVariant::from_ball(float radius) {
  ShapeVariant v;
  v.shape = (uint32_t)ShapeType::Ball;
  v.values[0] = radius;
  return v;
}
ShapeVariant ShapeVariant::from_rectangle(Vec2 h) {
  ShapeVariant v;
  v.shape = (uint32_t)ShapeType::Cuboid;
  v.values[0] = h.x;
  v.values[1] = h.y;
  return v;
}
ShapeVariant ShapeVariant::from_triangle(Vec2 a, Vec2 b, Vec2 c) {
  ShapeVariant v;
  v.shape = (uint32_t)ShapeType::Triangle;
  const float vals[6] = {a.x, a.y, b.x, b.y, c.x, c.y};
  std::memcpy(v.values, vals, sizeof(vals));
  return v;
}
ShapeVariant ShapeVariant::from_capsule(Vec2 a, Vec2 b, float radius) {
  ShapeVariant v;
  v.shape = (uint32_t)ShapeType::Capsule;
  const float vals[6] = {a.x, a.y, b.x, b.y, radius, 0.0f};
  std::memcpy(v.values, vals, sizeof(vals));
  return v;
}

efl_obstacle SolidObstacle::to_c() const {
  efl_obstacle o{};
  o.entity_id = entity_id;
  o.shape = shape.shape;
  std::memcpy(o.values, shape.values, sizeof(o.values));
  std::memcpy(o.transform, transform.data(), sizeof(o.transform));
  std::memcpy(o.inverse_transform, inverse_transform.data(), sizeof(o.inverse_transform));
  o.linear_velocity[0] = linear_velocity.x;
  o.linear_velocity[1] = linear_velocity.y;
  o.angular_velocity = angular_velocity;
  return o;
}

Collider Collider::circle(float r) {
  Collider c;
  c.type = ShapeType::Ball;
  c.values[0] = r;
  return c;
}
Collider Collider::rectangle(float w, float h) {
  Collider c;
  c.type = ShapeType::Cuboid;
  c.values[0] = 0.5f * w;
  c.values[1] = 0.5f * h;
  return c;
}
Collider Collider::capsule(Vec2 a, Vec2 b, float r) {
  Collider c;
  c.type = ShapeType::Capsule;
  const float v[6] = {a.x, a.y, b.x, b.y, r, 0.0f};
  std::memcpy(c.values, v, sizeof(v));
  return c;
}
Collider Collider::triangle(Vec2 a, Vec2 b, Vec2 cc) {
  Collider c;
  c.type = ShapeType::Triangle;
  const float v[6] = {a.x, a.y, b.x, b.y, cc.x, cc.y};
  std::memcpy(c.values, v, sizeof(v));
  return c;
}

namespace {
constexpr double kPi = 3.14159265358979323846;

// Area and second moment about the origin of the triangle (0, p, q).
void tri_origin(double px, double py, double qx, double qy, double& a, double& j) {
  const double cr = px * qy - py * qx;  // twice the signed area
  a = 0.5 * cr;
  j = cr / 12.0 * (px * px + px * qx + qx * qx + py * py + py * qy + qy * qy);
}
}  // namespace

float Collider::area() const {
  const float* v = values;
  switch (type) {
    case ShapeType::Ball: return (float)(kPi * v[0] * v[0]);
    case ShapeType::Cuboid: return 4.0f * v[0] * v[1];
    case ShapeType::Capsule: {
      const double l = std::hypot((double)v[2] - v[0], (double)v[3] - v[1]);
      return (float)(2.0 * v[4] * l + kPi * v[4] * v[4]);
    }
    case ShapeType::Triangle: {
      double a = 0, at, jt;
      tri_origin(v[0], v[1], v[2], v[3], at, jt), a += at;
      tri_origin(v[2], v[3], v[4], v[5], at, jt), a += at;
      tri_origin(v[4], v[5], v[0], v[1], at, jt), a += at;
      return (float)std::fabs(a);
    }
    default: return 0.0f;
  }
}

float Collider::second_moment() const {
  const float* v = values;
  switch (type) {
    case ShapeType::Ball: return (float)(0.5 * kPi * std::pow((double)v[0], 4));
    case ShapeType::Cuboid: {
      const double a = 4.0 * v[0] * v[1];
      return (float)(a * ((double)v[0] * v[0] + (double)v[1] * v[1]) / 3.0);
    }
    case ShapeType::Capsule: {
      // Rectangle 2r x l about the segment centre plus two half discs, then
      // moved to the origin (parallel axis through the segment centre).
      const double ax = v[0], ay = v[1], bx = v[2], by = v[3], r = v[4];
      const double l = std::hypot(bx - ax, by - ay);
      const double ar = 2.0 * r * l, jr = ar * (l * l + 4.0 * r * r) / 12.0;
      const double ad = kPi * r * r;
      // half disc (radius r) centroid 4r/(3 pi) from its flat side, which sits
      // at l/2 from the centre: J about the centre of both = disc J about its
      // own centre + area * (l/2)^2 + cross term of the centroid offset.
      const double jd = 0.5 * kPi * r * r * r * r + ad * (0.25 * l * l) + 2.0 * (0.5 * ad) * (4.0 * r / (3.0 * kPi)) * l;
      const double cx = 0.5 * (ax + bx), cy = 0.5 * (ay + by);
      return (float)(jr + jd + (ar + ad) * (cx * cx + cy * cy));
    }
    case ShapeType::Triangle: {
      double a = 0, j = 0, at, jt;
      tri_origin(v[0], v[1], v[2], v[3], at, jt), a += at, j += jt;
      tri_origin(v[2], v[3], v[4], v[5], at, jt), a += at, j += jt;
      tri_origin(v[4], v[5], v[0], v[1], at, jt), a += at, j += jt;
      return (float)std::fabs(j);
    }
    default: return 0.0f;
  }
}

std::vector<SolidObstacle> construct_rigid_body_buffer_for_gpu(const std::vector<SolidBody>& bodies) {
  std::vector<SolidObstacle> out;
  for (size_t idx = 0; idx < bodies.size(); ++idx) {
    const SolidBody& b = bodies[idx];
    const float* v = b.collider.values;
    ShapeVariant shape;
    switch (b.collider.type) {
      case ShapeType::Ball: shape = ShapeVariant::from_ball(v[0]); break;
      case ShapeType::Cuboid: shape = ShapeVariant::from_rectangle({v[0], v[1]}); break;
      case ShapeType::Triangle: shape = ShapeVariant::from_triangle({v[0], v[1]}, {v[2], v[3]}, {v[4], v[5]}); break;
      case ShapeType::Capsule: shape = ShapeVariant::from_capsule({v[0], v[1]}, {v[2], v[3]}, v[4]); break;
      default: warn("Unsupported shape type for solid"); continue;
    }
    SolidObstacle o;
    o.entity_id = (uint32_t)idx;
    o.shape = shape;
    o.transform = b.transform.to_matrix();
    o.inverse_transform = b.transform.inverse_matrix();
    o.linear_velocity = b.linear_velocity;
    o.angular_velocity = b.angular_velocity;
    out.push_back(o);
  }
  return out;
}

BodyImpulse forces_to_solid_readback(const FluidToSolidForce& f, float delta_secs, float grid_length) {
  BodyImpulse r;
  r.force.x = f.force.x * delta_secs / grid_length;
  r.force.y = -(f.force.y * delta_secs / grid_length);
  r.torque = -(f.torque * delta_secs / grid_length);
  return r;
}

// --------------------------------------------------------------- FluidWorld

FluidWorld::FluidWorld(const FluidPlugin& plugin, int device) : plugin_(plugin) {
  const int rc = efl_ctx_create(device, &ctx_);
  if (rc != EFL_OK) throw FluidError("efl_ctx_create failed", rc);
}

FluidWorld::~FluidWorld() {
  for (Fluid& f : fluids_)
    if (f.alive && f.inst) efl_instance_destroy(f.inst);
  if (ctx_) efl_ctx_destroy(ctx_);
}

void FluidWorld::check(int rc, const char* what) {
  if (rc == EFL_OK) return;
  const char* m = ctx_ ? efl_last_error(ctx_) : nullptr;
  throw FluidError(std::string(what) + ": " + (m ? m : "error"), rc);
}

FluidWorld::Fluid& FluidWorld::fluid(Entity e) {
  if (e >= fluids_.size() || !fluids_[e].alive) throw FluidError("no such fluid entity");
  return fluids_[e];
}
const FluidWorld::Fluid& FluidWorld::fluid(Entity e) const {
  if (e >= fluids_.size() || !fluids_[e].alive) throw FluidError("no such fluid entity");
  return fluids_[e];
}
efl_instance* FluidWorld::instance(Entity e) { return fluid(e).inst; }

Entity FluidWorld::spawn_fluid(const FluidSettings& s, const Transform& transform, const ProjectionMethod& projection,
                               const ReinitializeMethod& reinit) {
  efl_settings cs{};
  cs.width = s.size.x;
  cs.height = s.size.y;
  cs.rho = s.rho;
  cs.gravity[0] = s.gravity.x;
  cs.gravity[1] = s.gravity.y;
  cs.length_unit = plugin_.length_unit();
  const Mat4 m = transform.to_matrix();
  std::memcpy(cs.transform, m.data(), sizeof(cs.transform));
  cs.projection = projection.to_c();
  cs.reinit = reinit.to_c();
  Fluid f;
  check(efl_instance_create(ctx_, &cs, &f.inst), "spawn_fluid");
  f.alive = true;
  f.settings = s;
  f.transform = transform;
  f.projection = projection;
  f.reinit = reinit;
  f.status = FluidStatus::Reset;
  fluids_.push_back(std::move(f));
  return (Entity)(fluids_.size() - 1);
}

void FluidWorld::despawn_fluid(Entity e) {
  Fluid& f = fluid(e);
  check(efl_instance_destroy(f.inst), "despawn_fluid");
  f.inst = nullptr;
  f.alive = false;
}

void FluidWorld::spawn_source(Entity e, const FluidSource& source, const Transform& transform,
                              const FluidSourceShape& shape, const FluidSourceVelocity& velocity, bool oneshot) {
  fluid(e).sources.push_back(Source{source, transform, shape, velocity, oneshot});
}

void FluidWorld::clear_sources(Entity e) { fluid(e).sources.clear(); }
ForcesToFluid& FluidWorld::forces(Entity e) { return fluid(e).forces; }

void FluidWorld::set_projection(Entity e, const ProjectionMethod& m) {
  Fluid& f = fluid(e);
  const efl_projection p = m.to_c();
  check(efl_set_projection(f.inst, &p), "set_projection");
  f.projection = m;
}

void FluidWorld::set_reinit(Entity e, const ReinitializeMethod& m) {
  Fluid& f = fluid(e);
  const efl_reinit r = m.to_c();
  check(efl_set_reinit(f.inst, &r), "set_reinit");
  f.reinit = m;
}

void FluidWorld::set_settings(Entity e, const FluidSettings& s, const Transform& transform) {
  Fluid& f = fluid(e);
  if (s.size.x != f.settings.size.x || s.size.y != f.settings.size.y)
    throw FluidError("FluidSettings.size cannot change after spawning");
  const float g[2] = {s.gravity.x, s.gravity.y};
  const Mat4 m = transform.to_matrix();
  check(efl_set_uniforms(f.inst, s.rho, g, m.data()), "set_settings");
  f.settings = s;
  f.transform = transform;
}

void FluidWorld::reset(Entity e) { fluid(e).status = FluidStatus::Reset; }
FluidStatus FluidWorld::status(Entity e) const { return fluid(e).status; }

void FluidWorld::set_bodies(const std::vector<SolidBody>& bodies, bool integrate) {
  bodies_ = bodies;
  integrate_ = integrate;
  bodies_dirty_ = true;
}

std::vector<SolidBody> FluidWorld::bodies() {
  if (integrate_ && !bodies_dirty_) {
    const std::vector<SolidObstacle> obs = construct_rigid_body_buffer_for_gpu(bodies_);
    std::vector<efl_obstacle> cur(obs.size());
    if (!cur.empty()) check(efl_get_obstacles(ctx_, cur.data(), (uint32_t)cur.size()), "bodies");
    for (const efl_obstacle& o : cur) {
      SolidBody& b = bodies_[o.entity_id];
      b.transform.translation = {o.transform[12], o.transform[13]};
      b.transform.rotation = std::atan2(o.transform[1], o.transform[0]);
      b.linear_velocity = {o.linear_velocity[0], o.linear_velocity[1]};
      b.angular_velocity = o.angular_velocity;
    }
  }
  return bodies_;
}

void FluidWorld::physics_step(float delta_secs) {
  step_number_ += 1;
  time_step_ = delta_secs;
}

// update_fluid_source_init_buffer / update_fluid_source_buffer
// (fluid_source_uniform.rs:77-158): inactive sources are skipped, at most
// MAX_FLUID_SOURCE are kept (with a warning), oneshot sources carry zero
// velocity.
void FluidWorld::upload_sources(Fluid& f) {
  for (int oneshot = 1; oneshot >= 0; --oneshot) {
    efl_source data[EFL_MAX_FLUID_SOURCES];
    uint32_t count = 0;
    for (const Source& s : f.sources) {
      if (s.oneshot != (oneshot == 1) || !s.source.active) continue;
      if (count >= EFL_MAX_FLUID_SOURCES) {
        warn("The maximum number of fluid source per fluid component is 16.");
        break;
      }
      efl_source& d = data[count++];
      d = efl_source{};
      d.center[0] = s.transform.translation.x;
      d.center[1] = s.transform.translation.y;
      const Vec2 v = s.shape.to_vec2();
      d.data[0] = v.x;
      d.data[1] = v.y;
      d.velocity[0] = oneshot ? 0.0f : s.velocity.value.x;
      d.velocity[1] = oneshot ? 0.0f : s.velocity.value.y;
      d.shape = s.shape.shape_type_digit();
      d.mode = (uint32_t)s.source.mode;
    }
    check(efl_set_sources(f.inst, data, count, oneshot), "update_fluid_source");
  }
}

bool FluidWorld::update() {
  // Update systems
  for (Fluid& f : fluids_) {
    if (!f.alive) continue;
    upload_sources(f);
    // apply_forces_and_clear: the list of this frame replaces the last one.
    const ForceToFluid* fp = f.forces.forces.empty() ? nullptr : f.forces.forces.data();
    static_assert(sizeof(ForceToFluid) == sizeof(efl_force), "ForceToFluid layout");
    check(efl_set_forces(f.inst, reinterpret_cast<const efl_force*>(fp), (uint32_t)f.forces.forces.size()),
          "apply_forces_and_clear");
    f.forces.forces.clear();
  }
  if (bodies_dirty_) {
    const std::vector<SolidObstacle> obs = construct_rigid_body_buffer_for_gpu(bodies_);
    std::vector<efl_obstacle> c(obs.size());
    std::vector<efl_body> b(obs.size());
    for (size_t k = 0; k < obs.size(); ++k) {
      c[k] = obs[k].to_c();
      const SolidBody& sb = bodies_[obs[k].entity_id];
      b[k].density = sb.density;
      b[k].area = sb.collider.area();
      b[k].inertia = sb.collider.second_moment();
      b[k].dynamic = sb.rigid_body == RigidBody::Dynamic ? 1u : 0u;
    }
    check(efl_set_obstacles(ctx_, c.empty() ? nullptr : c.data(), (uint32_t)c.size()), "construct_rigid_body_buffer_for_gpu");
    check(efl_set_bodies(ctx_, integrate_ && !b.empty() ? b.data() : nullptr, integrate_ ? (uint32_t)b.size() : 0u),
          "set_bodies");
    bodies_dirty_ = false;
  }
  // DiagnosticsNode, ahead of the fluid node every frame
  if (diagnostics_on_)
    for (Fluid& f : fluids_) {
      if (!f.alive) continue;
      efl_fluid_diagnostics d{};
      check(efl_get_fluid_diagnostics(f.inst, &d), "DiagnosticsNode");
      f.diag = FluidDiagnostics{d.volume, d.min_velocity, d.max_velocity};
    }
  // EulerFluidNode::update: idle unless a new physics step happened.
  if (step_number_ == rendered_step_) return false;
  rendered_step_ = step_number_;
  std::vector<efl_instance*> batch;
  for (Fluid& f : fluids_) {
    if (!f.alive) continue;
    if (f.status == FluidStatus::Reset) check(efl_reset(f.inst), "reset");
    batch.push_back(f.inst);
  }
  for (size_t k = 0; k < batch.size(); k += EFL_MAX_BATCH) {
    const uint32_t n = (uint32_t)std::min<size_t>(EFL_MAX_BATCH, batch.size() - k);
    check(efl_step(ctx_, batch.data() + k, n, time_step_), "EulerFluidNode::run");
  }
  for (Fluid& f : fluids_)
    if (f.alive) f.status = f.status == FluidStatus::Reset ? FluidStatus::Uninitialized : FluidStatus::Initialized;
  return true;
}

FluidTextures FluidWorld::textures(Entity e) {
  Fluid& f = fluid(e);
  FluidTextures t;
  check(efl_field_device_ptr(f.inst, EFL_FIELD_U0, &t.u), "FluidTextures");
  check(efl_field_device_ptr(f.inst, EFL_FIELD_V0, &t.v), "FluidTextures");
  check(efl_field_device_ptr(f.inst, EFL_FIELD_U_SOLID, &t.u_solid), "FluidTextures");
  check(efl_field_device_ptr(f.inst, EFL_FIELD_V_SOLID, &t.v_solid), "FluidTextures");
  check(efl_field_device_ptr(f.inst, EFL_FIELD_LS_AIR0, &t.levelset_air), "FluidTextures");
  check(efl_field_device_ptr(f.inst, EFL_FIELD_LS_SOLID, &t.levelset_solid), "FluidTextures");
  return t;
}

std::vector<FluidToSolidForce> FluidWorld::forces_to_solid(Entity e) {
  Fluid& f = fluid(e);
  std::vector<efl_solid_force> raw(EFL_MAX_SOLIDS);
  check(efl_get_solid_forces(f.inst, raw.data(), EFL_MAX_SOLIDS), "forces_to_solid");
  std::vector<FluidToSolidForce> out(std::min<size_t>(bodies_.size(), EFL_MAX_SOLIDS));
  for (size_t k = 0; k < out.size(); ++k) out[k] = FluidToSolidForce{{raw[k].force[0], raw[k].force[1]}, raw[k].torque};
  return out;
}

void FluidWorld::read_field(Entity e, efl_field field, std::vector<float>& out) {
  Fluid& f = fluid(e);
  size_t count = 0;
  uint32_t w, h, ch;
  check(efl_field_info(f.inst, field, &count, &w, &h, &ch), "read_field");
  out.resize(count);
  check(efl_get_field(f.inst, field, out.data(), count), "read_field");
}

FluidDiagnostics FluidWorld::diagnostics(Entity e) { return fluid(e).diag; }

efl_stats FluidWorld::stats(Entity e) {
  efl_stats s{};
  check(efl_get_stats(fluid(e).inst, &s), "stats");
  return s;
}

}  // namespace bevy_eulerian_fluid
