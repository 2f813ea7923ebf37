// host_check.cpp - checks of the C++ host mirror (euler_fluid.hpp).
//
//   host_check cpu           host-only logic, no device call: FluidPlugin's
//                            panic, Display strings, defaults, the source and
//                            obstacle conversions, the readback conversion,
//                            body moments against numerical integration.
//   host_check scene <file>  a 128 x 96 fluid driven frame by frame through
//                            FluidWorld on device 0 (sources incl. an inactive
//                            one, bodies incl. an unsupported shape, local
//                            forces, idle frames, a final frame with the
//                            diagnostics plugin on); writes the obstacle
//                            records, the frame outcomes, every field, the
//                            solid forces and the diagnostics to <file>.
//
// Record format of <file>: 32-byte NUL-padded name, uint64 byte count, bytes.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "euler_fluid.hpp"

using namespace bevy_eulerian_fluid;

namespace {

int failures = 0;

void expect(bool ok, const std::string& what) {
  std::printf("%s %s\n", ok ? "ok  " : "FAIL", what.c_str());
  if (!ok) ++failures;
}

// Area and second moment about the origin by midpoint sampling on a fine
// lattice (independent of the closed forms).
void sample_moments(const Collider& c, double& area, double& j) {
  const float* v = c.values;
  double R = 0;
  switch (c.type) {
    case ShapeType::Ball: R = v[0]; break;
    case ShapeType::Cuboid: R = std::hypot(v[0], v[1]); break;
    case ShapeType::Capsule: R = std::fmax(std::hypot(v[0], v[1]), std::hypot(v[2], v[3])) + v[4]; break;
    default: R = std::fmax(std::fmax(std::hypot(v[0], v[1]), std::hypot(v[2], v[3])), std::hypot(v[4], v[5]));
  }
  const double h = R / 600.0;
  area = j = 0;
  for (double y = -R + 0.5 * h; y < R; y += h)
    for (double x = -R + 0.5 * h; x < R; x += h) {
      bool in = false;
      if (c.type == ShapeType::Ball) {
        in = x * x + y * y < (double)v[0] * v[0];
      } else if (c.type == ShapeType::Cuboid) {
        in = std::fabs(x) < v[0] && std::fabs(y) < v[1];
      } else if (c.type == ShapeType::Capsule) {
        const double abx = v[2] - v[0], aby = v[3] - v[1];
        double t = ((x - v[0]) * abx + (y - v[1]) * aby) / (abx * abx + aby * aby);
        t = std::fmin(1.0, std::fmax(0.0, t));
        in = std::hypot(x - (v[0] + t * abx), y - (v[1] + t * aby)) < v[4];
      } else {
        double s[3];
        for (int i = 0; i < 3; ++i) {
          const double px = v[2 * i], py = v[2 * i + 1], qx = v[2 * ((i + 1) % 3)], qy = v[2 * ((i + 1) % 3) + 1];
          s[i] = (qx - px) * (y - py) - (qy - py) * (x - px);
        }
        in = (s[0] > 0 && s[1] > 0 && s[2] > 0) || (s[0] < 0 && s[1] < 0 && s[2] < 0);
      }
      if (in) {
        area += h * h;
        j += (x * x + y * y) * h * h;
      }
    }
}

int run_cpu() {
  // FluidPlugin::new panics on length_unit <= 0 (euler_fluid.rs:47-49)
  bool threw = false;
  std::string msg;
  try {
    FluidPlugin p(0.0f);
  } catch (const FluidError& e) {
    threw = true;
    msg = e.what();
  }
  expect(threw && msg == "length_unit must be positive value.", "FluidPlugin::new(0) panics with the reference message");
  threw = false;
  try {
    FluidPlugin p(-1.0f);
  } catch (const FluidError&) {
    threw = true;
  }
  expect(threw, "FluidPlugin::new(-1) panics");
  expect(FluidPlugin(10.0f).grid_length() == 0.1f, "FluidGridLength = 1 / length_unit");

  // Display and defaults (projection.rs:54-80, gauss_seidel.rs:38-42,
  // multi_grid.rs:63-71, reinitialize_levelset.rs:42-60)
  expect(ProjectionMethod().to_string() == "Gauss-Seidel(20)", "default ProjectionMethod is Gauss-Seidel(20)");
  expect(ProjectionMethod::Jacobi().to_string() == "Jacobi", "Jacobi display");
  expect(ProjectionMethod::MultiGrid().to_string() == "Multigrid(2-3-2)", "Multigrid default display (pre-coarsest-post)");
  MultiGridConfig mg;
  mg.pre_smooth_config.num_iterations = 4;
  mg.coarsest_config.num_iterations = 7;
  mg.post_smooth_config.num_iterations = 5;
  expect(ProjectionMethod::MultiGrid(mg).to_string() == "Multigrid(4-7-5)", "Multigrid display order");
  expect(ReinitializeMethod().to_string() == "Fast Iterative(10)", "default ReinitializeMethod is Fast Iterative(10)");
  expect(ReinitializeMethod::JumpFlooding().to_string() == "Jump Flooding", "Jump Flooding display");
  const efl_projection pc = ProjectionMethod::MultiGrid(mg).to_c();
  expect(pc.kind == EFL_PROJECTION_MULTIGRID && pc.mg_pre_iterations == 4 && pc.mg_coarsest_iterations == 7 &&
             pc.mg_post_iterations == 5 && pc.max_cycles == 1 && pc.rel_tol == 0.0f,
         "MultiGrid -> efl_projection, one V-cycle per step like the reference");
  const efl_projection gc = ProjectionMethod().to_c();
  expect(gc.kind == EFL_PROJECTION_GAUSS_SEIDEL && gc.gs_iterations == 20, "GaussSeidel -> efl_projection");
  const efl_reinit rc = ReinitializeMethod::JumpFlooding().to_c();
  expect(rc.kind == EFL_REINIT_JUMP_FLOODING, "JumpFlooding -> efl_reinit");

  // FluidSourceShape (fluid_source.rs:58-80)
  expect(FluidSourceShape().shape_type_digit() == 0 && FluidSourceShape().to_vec2().x == 1.0f,
         "FluidSourceShape default is Circle { radius: 1.0 }");
  const FluidSourceShape aabb = FluidSourceShape::Aabb({3.0f, 4.0f});
  expect(aabb.shape_type_digit() == 1 && aabb.to_vec2().x == 3.0f && aabb.to_vec2().y == 4.0f, "Aabb digit and data");
  expect(FluidSourceShape::Circle(2.5f).to_vec2().y == 0.0f, "Circle data is (radius, 0)");

  // construct_rigid_body_buffer_for_gpu: ids are body indices, unsupported
  // shapes are skipped, transforms are T * R and their inverse
  std::vector<SolidBody> bodies(4);
  bodies[0].collider = Collider::circle(5.0f);
  bodies[0].transform = Transform::from_rotation_translation(0.7f, {10.0f, -3.0f});
  bodies[1].collider.type = ShapeType::Unsupported;
  bodies[2].collider = Collider::rectangle(8.0f, 6.0f);
  bodies[2].transform = Transform::from_rotation_translation(-1.2f, {-20.0f, 7.5f});
  bodies[2].linear_velocity = {1.5f, -2.0f};
  bodies[2].angular_velocity = 0.25f;
  bodies[3].collider = Collider::capsule({-4.0f, 0.0f}, {4.0f, 0.0f}, 2.0f);
  const std::vector<SolidObstacle> obs = construct_rigid_body_buffer_for_gpu(bodies);
  expect(obs.size() == 3 && obs[0].entity_id == 0 && obs[1].entity_id == 2 && obs[2].entity_id == 3,
         "unsupported shape skipped, entity ids keep body indices");
  expect(obs[1].shape.shape == 1 && obs[1].shape.values[0] == 4.0f && obs[1].shape.values[1] == 3.0f,
         "Collider::rectangle(8, 6) -> cuboid half extents (4, 3)");
  expect(obs[2].shape.shape == 2 && obs[2].shape.values[4] == 2.0f, "capsule -> (a, b, radius)");
  double worst = 0;
  for (const SolidObstacle& o : obs)
    for (int r = 0; r < 4; ++r)
      for (int c = 0; c < 4; ++c) {
        double s = 0;
        for (int k = 0; k < 4; ++k) s += (double)o.transform[k * 4 + r] * o.inverse_transform[c * 4 + k];
        worst = std::fmax(worst, std::fabs(s - (r == c ? 1.0 : 0.0)));
      }
  expect(worst < 1e-5, "inverse_transform * transform = I (max err " + std::to_string(worst) + ")");
  const efl_obstacle co = obs[1].to_c();
  expect(co.entity_id == 2 && co.linear_velocity[1] == -2.0f && co.angular_velocity == 0.25f && co.reserved == 0.0f,
         "SolidObstacle -> efl_obstacle");

  // forces_to_solid_readback (fluid_to_solid.rs:196-205)
  const BodyImpulse bi = forces_to_solid_readback(FluidToSolidForce{{2.0f, 3.0f}, 5.0f}, 0.5f, 0.25f);
  expect(bi.force.x == 4.0f && bi.force.y == -6.0f && bi.torque == -10.0f, "readback: x dt/dx, y flipped, torque negated");

  // closed-form body moments against sampling
  const Collider cs[4] = {Collider::circle(7.0f), Collider::rectangle(10.0f, 4.0f),
                          Collider::capsule({-6.0f, 1.0f}, {5.0f, 3.0f}, 2.5f),
                          Collider::triangle({-5.0f, -3.0f}, {6.0f, -2.0f}, {0.5f, 7.0f})};
  for (const Collider& c : cs) {
    double a, j;
    sample_moments(c, a, j);
    const double ea = std::fabs(c.area() - a) / a, ej = std::fabs(c.second_moment() - j) / j;
    expect(ea < 2e-3 && ej < 2e-3, "moments of shape " + std::to_string((int)c.type) + " (rel err area " +
                                       std::to_string(ea) + ", J " + std::to_string(ej) + ")");
  }
  std::printf("%s\n", failures ? "host checks FAILED" : "host checks passed");
  return failures ? 1 : 0;
}

struct Writer {
  FILE* f;
  void put(const char* name, const void* data, uint64_t n) {
    char hdr[32] = {0};
    std::memcpy(hdr, name, std::min<size_t>(31, std::strlen(name)));
    std::fwrite(hdr, 1, 32, f);
    std::fwrite(&n, 8, 1, f);
    if (n) std::fwrite(data, 1, n, f);
  }
};

int run_scene(const char* path) {
  FluidPlugin plugin(10.0f);
  FluidWorld world(plugin, 0);
  FluidSettings s{99.7f, {0.0f, 9.8f}, {128, 96}};
  const Entity fluid = world.spawn_fluid(s);  // default Transform, GaussSeidel(20), FastIterative(10)

  // oneshot box over the lower part, one continuous source, one inactive sink
  world.spawn_source(fluid, FluidSource{true, FluidSourceMode::Source}, Transform::from_translation({0.0f, -16.0f}),
                     FluidSourceShape::Aabb({64.0f, 32.0f}), FluidSourceVelocity{{7.0f, 7.0f}}, true);
  world.spawn_source(fluid, FluidSource{true, FluidSourceMode::Source}, Transform::from_translation({-48.0f, 28.0f}),
                     FluidSourceShape::Aabb({5.0f, 5.0f}), FluidSourceVelocity{{50.0f, 10.0f}});
  world.spawn_source(fluid, FluidSource{false, FluidSourceMode::Sink}, Transform::from_translation({30.0f, 0.0f}),
                     FluidSourceShape::Circle(6.0f));

  std::vector<SolidBody> bodies(5);
  bodies[0].collider = Collider::circle(6.0f);
  bodies[0].transform = Transform::from_rotation_translation(0.0f, {20.0f, -10.0f});
  bodies[0].linear_velocity = {3.0f, 1.0f};
  bodies[1].collider = Collider::rectangle(16.0f, 8.0f);
  bodies[1].transform = Transform::from_rotation_translation(0.5f, {-25.0f, -20.0f});
  bodies[1].angular_velocity = 0.8f;
  bodies[2].collider.type = ShapeType::Unsupported;
  bodies[3].collider = Collider::capsule({-5.0f, 0.0f}, {5.0f, 0.0f}, 3.0f);
  bodies[3].transform = Transform::from_rotation_translation(-0.3f, {35.0f, -25.0f});
  bodies[4].collider = Collider::triangle({-6.0f, -4.0f}, {6.0f, -4.0f}, {0.0f, 8.0f});
  bodies[4].transform = Transform::from_rotation_translation(1.1f, {0.0f, 10.0f});
  world.set_bodies(bodies);

  const float dt = 1.0f / 60.0f;
  std::vector<int32_t> frames;
  // frame 1: first physics step -> initialisation
  world.physics_step(dt);
  frames.push_back(world.update());
  // frame 2: no new physics step -> idle
  frames.push_back(world.update());
  // frame 3: local forces, step 1
  world.forces(fluid).forces.push_back(ForceToFluid{{0.0f, -40.0f}, {64.0f, 60.0f}});
  world.forces(fluid).forces.push_back(ForceToFluid{{25.0f, 0.0f}, {30.0f, 70.0f}});
  world.physics_step(dt);
  frames.push_back(world.update());
  // frame 4: forces pushed on an idle frame are replaced by the next frame's
  // (empty) list, as apply_forces_and_clear does every frame
  world.forces(fluid).forces.push_back(ForceToFluid{{0.0f, 500.0f}, {64.0f, 48.0f}});
  frames.push_back(world.update());
  // frames 5, 6: steps 2 and 3
  world.physics_step(dt);
  frames.push_back(world.update());
  world.physics_step(dt);
  frames.push_back(world.update());

  // a frame with the diagnostics plugin on and no physics step: the passes
  // run on the final state, the fluid node idles
  world.enable_diagnostics(true);
  const bool idle = !world.update();
  const FluidDiagnostics diag = world.diagnostics(fluid);

  // FluidTextures.u is the device field U0
  const FluidTextures tex = world.textures(fluid);
  std::vector<float> u;
  world.read_field(fluid, EFL_FIELD_U0, u);
  std::vector<float> ut(u.size());
  if (hipMemcpy(ut.data(), tex.u, u.size() * 4, hipMemcpyDeviceToHost) != hipSuccess) {
    std::fprintf(stderr, "hipMemcpy of FluidTextures.u failed\n");
    return 2;
  }
  const bool tex_ok = std::memcmp(ut.data(), u.data(), u.size() * 4) == 0 && tex.v && tex.u_solid && tex.v_solid &&
                      tex.levelset_air && tex.levelset_solid;

  FILE* f = std::fopen(path, "wb");
  if (!f) return 2;
  Writer w{f};
  w.put("frames", frames.data(), frames.size() * 4);
  const int32_t t = tex_ok;
  w.put("textures_ok", &t, 4);
  const std::vector<SolidObstacle> obs = construct_rigid_body_buffer_for_gpu(bodies);
  std::vector<efl_obstacle> dev(obs.size());
  if (efl_get_obstacles(world.context(), dev.data(), (uint32_t)dev.size()) != EFL_OK) return 2;
  w.put("obstacles", dev.data(), dev.size() * sizeof(efl_obstacle));
  const std::vector<FluidToSolidForce> fs = world.forces_to_solid(fluid);
  w.put("forces_to_solid", fs.data(), fs.size() * sizeof(FluidToSolidForce));
  const efl_stats st = world.stats(fluid);
  w.put("stats", &st, sizeof(st));
  const float dg[4] = {diag.volume, diag.min_velocity, diag.max_velocity, idle ? 1.0f : 0.0f};
  w.put("diagnostics", dg, sizeof(dg));
  for (int fld = 0; fld < EFL_FIELD_COUNT; ++fld) {
    std::vector<float> data;
    world.read_field(fluid, (efl_field)fld, data);
    char name[32];
    std::snprintf(name, sizeof(name), "field%d", fld);
    w.put(name, data.data(), data.size() * 4);
  }
  std::fclose(f);
  std::printf("scene written: %s (frames", path);
  for (int x : frames) std::printf(" %d", x);
  std::printf(")\n");
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc >= 2 && std::string(argv[1]) == "cpu") return run_cpu();
  if (argc >= 3 && std::string(argv[1]) == "scene") {
    try {
      return run_scene(argv[2]);
    } catch (const FluidError& e) {
      std::fprintf(stderr, "FluidError (%d): %s\n", e.status(), e.what());
      return 3;
    }
  }
  std::fprintf(stderr, "usage: host_check cpu | host_check scene <file>\n");
  return 64;
}
