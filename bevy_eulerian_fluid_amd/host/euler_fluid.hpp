// euler_fluid.hpp - C++ host mirror of bevy_eulerian_fluid's main-world API
// on top of the C ABI in include/efl.h.
//
// The reference is a Bevy plugin: users spawn a FluidSettings entity with
// child FluidSource entities, push LocalForces (ForcesToFluid), read
// FluidTextures, and avian2d rigid bodies become SolidObstacle records.  This
// header keeps those names, fields, defaults and error behaviour, and replaces
// the ECS plumbing by a small FluidWorld that owns one efl_ctx and its
// instances:
//
//   FluidPlugin                    euler_fluid.rs:43-52 (panics on length_unit <= 0)
//   FluidSettings, FluidTextures   settings.rs:100-147
//   FluidSource*                   fluid_source.rs:31-88, conversion
//                                  fluid_source_uniform.rs:77-158
//   ForceToFluid, ForcesToFluid    apply_forces.rs:36-45, cleared after upload
//                                  (apply_forces.rs:157-167)
//   ProjectionMethod + configs     projection.rs:45-80, gauss_seidel.rs:33-42,
//                                  multi_grid.rs:56-71
//   ReinitializeMethod + config    reinitialize_levelset.rs:36-60,
//                                  fast_iterative_method.rs:14-23
//   FluidStatus                    fluid_status.rs:15-38
//   ShapeVariant, SolidObstacle    obstacle.rs:17-160
//   FluidToSolidForce + readback   fluid_to_solid.rs:60-64, 180-209
//   FluidTimeStep, PhysicsFrameInfo, gating   physics_time.rs:10-59,
//                                  render_node.rs:165-175
//
// A Rust panic becomes a thrown bevy_eulerian_fluid::FluidError (derived from
// std::runtime_error); a failed ABI call throws FluidError with the library's
// message; a Bevy warn! is written to stderr and the call continues, as in
// the reference.
#pragma once

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "efl.h"

namespace bevy_eulerian_fluid {

struct Vec2 {
  float x = 0.0f, y = 0.0f;
};
struct UVec2 {
  uint32_t x = 0, y = 0;
};
// Column-major 4x4, glam's Mat4 layout.
using Mat4 = std::array<float, 16>;

class FluidError : public std::runtime_error {
 public:
  explicit FluidError(const std::string& what, int status = EFL_E_INVALID)
      : std::runtime_error(what), status_(status) {}
  int status() const { return status_; }

 private:
  int status_;
};

// 2D Transform: translation, rotation about z (radians), scale.
struct Transform {
  Vec2 translation;
  float rotation = 0.0f;
  Vec2 scale{1.0f, 1.0f};
  static Transform from_translation(Vec2 t) { return Transform{t, 0.0f, {1.0f, 1.0f}}; }
  static Transform from_rotation_translation(float angle, Vec2 t) { return Transform{t, angle, {1.0f, 1.0f}}; }
  // T * R * S, as glam's Mat4::from_scale_rotation_translation.
  Mat4 to_matrix() const;
  // Inverse of to_matrix() (exact algebraic inverse of T * R * S).
  Mat4 inverse_matrix() const;
};

// FluidPlugin::new (euler_fluid.rs:43-52): length_unit <= 0 panics; the
// resulting FluidGridLength is 1 / length_unit (euler_fluid.rs:104).
class FluidPlugin {
 public:
  explicit FluidPlugin(float length_unit);
  float length_unit() const { return length_unit_; }
  float grid_length() const { return 1.0f / length_unit_; }

 private:
  float length_unit_;
};

// FluidSettings (settings.rs:110-114).
struct FluidSettings {
  float rho = 0.0f;
  Vec2 gravity;
  UVec2 size;
};

// ---- fluid sources (fluid_source.rs:31-88) ----
enum class FluidSourceMode : uint32_t { Source = 0, Sink = 1 };

struct FluidSource {
  bool active = false;
  FluidSourceMode mode = FluidSourceMode::Source;
};

class FluidSourceShape {
 public:
  static FluidSourceShape Circle(float radius) { return FluidSourceShape(0, {radius, 0.0f}); }
  static FluidSourceShape Aabb(Vec2 half_size) { return FluidSourceShape(1, half_size); }
  FluidSourceShape() : FluidSourceShape(0, {1.0f, 0.0f}) {}  // Circle { radius: 1.0 }
  Vec2 to_vec2() const { return data_; }
  uint32_t shape_type_digit() const { return digit_; }

 private:
  FluidSourceShape(uint32_t d, Vec2 v) : digit_(d), data_(v) {}
  uint32_t digit_;
  Vec2 data_;
};

struct FluidSourceVelocity {
  Vec2 value;
};

// ---- forces (apply_forces.rs:36-45) ----
struct ForceToFluid {
  Vec2 force;     // m/s^2
  Vec2 position;  // cells, x right, y down
};
struct ForcesToFluid {
  std::vector<ForceToFluid> forces;
};
using LocalForces = ForcesToFluid;

// ---- pressure projection (projection.rs:45-80) ----
struct GaussSeidelConfig {
  uint32_t num_iterations = 20;
};
struct MultiGridConfig {
  GaussSeidelConfig pre_smooth_config{2};
  GaussSeidelConfig post_smooth_config{2};
  GaussSeidelConfig coarsest_config{3};
};
// Residual target, not in the reference (it runs one V-cycle per step):
// max_cycles V-cycles (MultiGrid) or CG iterations (MgPcg) until
// ||r|| / ||b|| <= rel_tol, checked every check_every cycles.
struct SolveTarget {
  uint32_t max_cycles = 1;
  float rel_tol = 0.0f;
  uint32_t check_every = 1;
};

class ProjectionMethod {
 public:
  enum class Kind { Jacobi, GaussSeidel, MultiGrid, MgPcg };
  static ProjectionMethod Jacobi() { return ProjectionMethod(Kind::Jacobi); }
  static ProjectionMethod GaussSeidel(GaussSeidelConfig c = {}) {
    ProjectionMethod m(Kind::GaussSeidel);
    m.gauss_seidel_ = c;
    return m;
  }
  static ProjectionMethod MultiGrid(MultiGridConfig c = {}, SolveTarget t = {}) {
    ProjectionMethod m(Kind::MultiGrid);
    m.multi_grid_ = c;
    m.target_ = t;
    return m;
  }
  // CG preconditioned by one symmetric V-cycle (no reference counterpart).
  static ProjectionMethod MgPcg(MultiGridConfig c = {}, SolveTarget t = {100, 1e-5f, 1}) {
    ProjectionMethod m(Kind::MgPcg);
    m.multi_grid_ = c;
    m.target_ = t;
    return m;
  }
  ProjectionMethod() : ProjectionMethod(GaussSeidel()) {}  // projection.rs:54-58
  Kind kind() const { return kind_; }
  const GaussSeidelConfig& gauss_seidel() const { return gauss_seidel_; }
  const MultiGridConfig& multi_grid() const { return multi_grid_; }
  const SolveTarget& target() const { return target_; }
  // Display (projection.rs:60-80): "Jacobi", "Gauss-Seidel(20)", "Multigrid(2-3-2)".
  std::string to_string() const;
  efl_projection to_c() const;

 private:
  explicit ProjectionMethod(Kind k) : kind_(k) {}
  Kind kind_;
  GaussSeidelConfig gauss_seidel_;
  MultiGridConfig multi_grid_;
  SolveTarget target_;
};

// ---- redistancing (reinitialize_levelset.rs:36-60) ----
struct FastIterativeMethodConfig {
  uint32_t num_iterations = 10;
};

class ReinitializeMethod {
 public:
  static ReinitializeMethod JumpFlooding() { return ReinitializeMethod(true, {}); }
  static ReinitializeMethod FastIterative(FastIterativeMethodConfig c = {}) { return ReinitializeMethod(false, c); }
  ReinitializeMethod() : ReinitializeMethod(FastIterative()) {}
  bool is_jump_flooding() const { return jfa_; }
  const FastIterativeMethodConfig& fast_iterative() const { return fim_; }
  // Display: "Jump Flooding", "Fast Iterative(10)".
  std::string to_string() const;
  efl_reinit to_c() const;

 private:
  ReinitializeMethod(bool j, FastIterativeMethodConfig c) : jfa_(j), fim_(c) {}
  bool jfa_;
  FastIterativeMethodConfig fim_;
};

// FluidStatus (fluid_status.rs:15-20).
enum class FluidStatus { Reset, Uninitialized, Initialized };

// ---- solids (obstacle.rs:17-87) ----
// parry ShapeType ids.
enum class ShapeType : uint32_t { Ball = 0, Cuboid = 1, Capsule = 2, Triangle = 4, Unsupported = 0xffffffffu };

struct ShapeVariant {
  uint32_t shape = 0;
  float values[6] = {0, 0, 0, 0, 0, 0};
  static ShapeVariant from_ball(float radius);
  static ShapeVariant from_rectangle(Vec2 half_extents);
  static ShapeVariant from_triangle(Vec2 a, Vec2 b, Vec2 c);
  static ShapeVariant from_capsule(Vec2 a, Vec2 b, float radius);
};

struct SolidObstacle {
  uint32_t entity_id = 0;
  ShapeVariant shape;
  Mat4 transform{};
  Mat4 inverse_transform{};
  Vec2 linear_velocity;
  float angular_velocity = 0.0f;
  efl_obstacle to_c() const;
};

// The avian2d side of a solid: collider, pose, velocities and body type.
// density (relative to FluidSettings.rho) is only used when the library
// integrates Dynamic bodies itself (FluidWorld::set_integrate_bodies).
enum class RigidBody { Static, Kinematic, Dynamic };
struct Collider {
  ShapeType type = ShapeType::Ball;
  float values[6] = {0, 0, 0, 0, 0, 0};
  static Collider circle(float r);
  static Collider rectangle(float width, float height);  // avian's Collider::rectangle takes full sizes
  static Collider capsule(Vec2 a, Vec2 b, float radius);
  static Collider triangle(Vec2 a, Vec2 b, Vec2 c);
  // Area (cells^2) and second moment of area about the local origin
  // (cells^4), closed forms.
  float area() const;
  float second_moment() const;
};
struct SolidBody {
  Transform transform;  // GlobalTransform
  Collider collider;
  Vec2 linear_velocity;
  float angular_velocity = 0.0f;
  RigidBody rigid_body = RigidBody::Dynamic;
  float density = 1.0f;
};

// construct_rigid_body_buffer_for_gpu (obstacle.rs:89-160): entity_id is the
// body's index among all bodies; unsupported shapes are skipped with a
// warning (their index stays taken).
std::vector<SolidObstacle> construct_rigid_body_buffer_for_gpu(const std::vector<SolidBody>& bodies);

// FluidToSolidForce (fluid_to_solid.rs:60-64) and the conversion of
// forces_to_solid_readback (fluid_to_solid.rs:180-209): force * dt / dx with
// y flipped, and the torque to apply (already negated).
struct FluidToSolidForce {
  Vec2 force;
  float torque = 0.0f;
};
struct BodyImpulse {
  Vec2 force;
  float torque = 0.0f;
};
BodyImpulse forces_to_solid_readback(const FluidToSolidForce& f, float delta_secs, float grid_length);

// FluidTextures (settings.rs:139-147): device pointers of the fields (f32,
// row-major [y][x]); valid while the fluid exists.
struct FluidTextures {
  void* u = nullptr;             // (W+1) x H
  void* v = nullptr;             // W x (H+1)
  void* u_solid = nullptr;       // (W+1) x H
  void* v_solid = nullptr;       // W x (H+1)
  void* levelset_air = nullptr;  // W x H
  void* levelset_solid = nullptr;
};

// A fluid entity's handle inside a FluidWorld.
using Entity = uint32_t;

// One Bevy App with FluidPlugin: the fluid entities, their child sources,
// the solids, the physics clock and the render node, on one device.
// diagnostics/component.rs: FluidVolume, FluidMinVelocityMagnitude,
// FluidMaxVelocityMagnitude of a fluid
struct FluidDiagnostics {
  float volume = 0.0f;
  float min_velocity = 0.0f;
  float max_velocity = 0.0f;
};

class FluidWorld {
 public:
  FluidWorld(const FluidPlugin& plugin, int device = 0);
  ~FluidWorld();
  FluidWorld(const FluidWorld&) = delete;
  FluidWorld& operator=(const FluidWorld&) = delete;

  // Spawning FluidSettings (with its required components,
  // settings.rs:100-109) creates the grids (watch_fluid_component).
  Entity spawn_fluid(const FluidSettings& settings, const Transform& transform = {},
                     const ProjectionMethod& projection = ProjectionMethod(),
                     const ReinitializeMethod& reinit = ReinitializeMethod());
  void despawn_fluid(Entity e);
  // A child FluidSource; oneshot = FluidSourceOneshot.
  void spawn_source(Entity fluid, const FluidSource& source, const Transform& transform,
                    const FluidSourceShape& shape, const FluidSourceVelocity& velocity = {},
                    bool oneshot = false);
  void clear_sources(Entity fluid);
  ForcesToFluid& forces(Entity fluid);
  void set_projection(Entity fluid, const ProjectionMethod& m);
  void set_reinit(Entity fluid, const ReinitializeMethod& m);
  void set_settings(Entity fluid, const FluidSettings& s, const Transform& transform);
  // Insert FluidStatus::Reset: the next run re-initialises the fluid.
  void reset(Entity fluid);
  FluidStatus status(Entity fluid) const;

  // The rigid bodies (global, like SolidObstaclesBuffer).  With
  // integrate = true the library steps Dynamic bodies itself (its
  // semi-implicit integrator in place of avian2d); bodies() then returns the
  // integrated poses after each run.
  void set_bodies(const std::vector<SolidBody>& bodies, bool integrate = false);
  std::vector<SolidBody> bodies();

  // PhysicsSchedule end (update_physics_frame_info, physics_time.rs:50-59).
  void physics_step(float delta_secs);
  // One frame: the Update systems (source uniforms, obstacle buffer, forces
  // upload and clear), then the render node, which advances every fluid once
  // if a new physics step happened since the last run (render_node.rs:165-175)
  // and idles otherwise.  Returns true if the fluids were advanced.
  bool update();

  FluidTextures textures(Entity fluid);
  // The last forces_to_solid readback of a fluid, one per body index.
  std::vector<FluidToSolidForce> forces_to_solid(Entity fluid);
  // Copy a field to the host (efl_field ids); resizes out.
  void read_field(Entity fluid, efl_field field, std::vector<float>& out);
  efl_stats stats(Entity fluid);
  // FluidDiagnosticsPlugin (diagnostics/plugin.rs:20-48): with it on, every
  // update() runs the volume and min / max speed passes for each fluid before
  // the fluid node, as the reference's DiagnosticsNode does each frame
  // (render_graph edge DiagnosticsLabel -> FluidLabel); diagnostics() returns
  // the last values (the reference reads them back frames later).
  void enable_diagnostics(bool on) { diagnostics_on_ = on; }
  FluidDiagnostics diagnostics(Entity fluid);
  uint64_t physics_step_number() const { return step_number_; }
  float time_step() const { return time_step_; }
  efl_ctx* context() { return ctx_; }
  efl_instance* instance(Entity fluid);

 private:
  struct Source {
    FluidSource source;
    Transform transform;
    FluidSourceShape shape;
    FluidSourceVelocity velocity;
    bool oneshot;
  };
  struct Fluid {
    bool alive = false;
    efl_instance* inst = nullptr;
    FluidSettings settings;
    Transform transform;
    ProjectionMethod projection;
    ReinitializeMethod reinit;
    FluidStatus status = FluidStatus::Reset;
    std::vector<Source> sources;
    ForcesToFluid forces;
    FluidDiagnostics diag;
  };
  Fluid& fluid(Entity e);
  const Fluid& fluid(Entity e) const;
  void check(int rc, const char* what);
  void upload_sources(Fluid& f);

  FluidPlugin plugin_;
  efl_ctx* ctx_ = nullptr;
  std::vector<Fluid> fluids_;
  std::vector<SolidBody> bodies_;
  bool integrate_ = false;
  bool bodies_dirty_ = false;
  bool diagnostics_on_ = false;
  uint64_t step_number_ = 0;    // PhysicsFrameInfo.step_number
  uint64_t rendered_step_ = 0;  // CurrentPhysicsStepNumberRenderWorld
  float time_step_ = 0.0f;      // FluidTimeStep
};

}  // namespace bevy_eulerian_fluid
