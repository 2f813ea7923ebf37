/*
 * efl_oracle.h - CPU oracle of the per-step fluid pipeline (TEST INFRASTRUCTURE).
 *
 * A plain C++ restatement of bevy_eulerian_fluid's hot-path WGSL kernels and
 * of the dispatch order of EulerFluidNode::run (src/euler_fluid/render_node.rs
 * :193-476).  It is the checker for the HIP library and the timed CPU
 * baseline; only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg may load it.  Nothing in the product links or calls it.
 *
 * Field ids, stage ids and the scene structs are those of include/efl.h so
 * that a test can drive both implementations with the same inputs.
 */
#ifndef EFL_ORACLE_H
#define EFL_ORACLE_H

#include "../include/efl.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_inst orc_inst;

orc_inst* orc_create(const efl_settings* settings);
void orc_destroy(orc_inst* o);
int orc_set_projection(orc_inst* o, const efl_projection* p);
int orc_set_reinit(orc_inst* o, const efl_reinit* r);
int orc_set_uniforms(orc_inst* o, float rho, const float gravity[2], const float transform[16]);
int orc_set_sources(orc_inst* o, const efl_source* s, uint32_t n, int oneshot);
int orc_set_obstacles(orc_inst* o, const efl_obstacle* ob, uint32_t n);
int orc_set_forces(orc_inst* o, const efl_force* f, uint32_t n);
int orc_set_bodies(orc_inst* o, const efl_body* b, uint32_t n);
int orc_get_obstacles(orc_inst* o, efl_obstacle* out, uint32_t n);
int orc_init_levelset(orc_inst* o, const float* ls);
int orc_reset(orc_inst* o);
int orc_initialize(orc_inst* o);
int orc_step(orc_inst* o, float dt);
int orc_run_stage(orc_inst* o, int stage, float dt);
int orc_field_info(const orc_inst* o, int field, size_t* count, uint32_t* w, uint32_t* h, uint32_t* c);
int orc_get_field(const orc_inst* o, int field, void* out, size_t count);
int orc_set_field(orc_inst* o, int field, const void* in, size_t count);
int orc_get_stats(const orc_inst* o, efl_stats* out);
int orc_set_diagnostics(orc_inst* o, int enabled);
/* ||r||2/||b||2 and dt*max|D| of the current state (FP32 arithmetic). */
int orc_measure(orc_inst* o, float dt, float* rel_residual, float* max_norm_div);
/* Solve the projection's discrete Poisson system (current div, ls_air0,
 * area fractions) in FP64 to rel_tol, starting from zero; returns the
 * pressure (W x H doubles) and the achieved relative residual. */
int orc_solve_fp64(orc_inst* o, float dt, double rel_tol, double* p_out, double* rel_res_out, int* iters_out);
/* Apply the FP32 operator of residual.wgsl: r = b - A x on the finest level
 * for a caller-given x (W x H), b = the current div. */
int orc_residual_of(orc_inst* o, float dt, const float* x, float* r_out);
/* Single-threaded switch (1 = serial, otherwise OpenMP threads if built). */
/* volume, smallest and largest speed as the reference's diagnostics passes
 * compute them from u0 / v0 / ls0: out[3] = {volume, min, max} */
int orc_fluid_diagnostics(const orc_inst* o, float* out);
int orc_set_threads(int n);

/* Multigrid pyramid access for the pinning tests.  which: 0 x, 1 b, 2 level
 * set, 3 area fractions (4 per cell), 4 residual; level 0 is p0 / div /
 * levelset_air0 / area fractions / r0.  set = 1 copies `data` in, 0 out. */
int orc_mg_level_size(const orc_inst* o, int level, uint32_t* w, uint32_t* h);
int orc_mg_level(orc_inst* o, int level, int which, float* data, size_t count, int set);
/* restriction.wgsl on level -> level + 1; prolongation.wgsl level + 1 -> level */
int orc_mg_restrict(orc_inst* o, int level);
int orc_mg_prolong(orc_inst* o, int level);

#ifdef __cplusplus
}
#endif

#endif
