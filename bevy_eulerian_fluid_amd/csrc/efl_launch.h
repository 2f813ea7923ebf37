// efl_launch.h - host-side launchers of the HIP kernels (efl_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "efl_internal.h"

namespace efl {

// A launcher that finds an internal invariant broken skips its launch and
// records why; the entry point that queued it returns EFL_E_STATE with the
// message (nothing aborts across the C ABI).  One per host thread.
void set_launch_error(const char* msg);
const char* take_launch_error();  // nullptr when none; clears it

int tile_w();
int tile_h();
int cg_tile_h();
int sample_forces_tile_h();  // k_sample_forces: tile_w() x this per block

void launch_init_center(hipStream_t st, const Batch& b);
void launch_init_edge(hipStream_t st, const Batch& b);
void launch_fluid_source(hipStream_t st, const Batch& b, int oneshot);
void launch_update_solid(hipStream_t st, const Batch& b);
void launch_area_fraction(hipStream_t st, const Batch& b);
void launch_advect_u(hipStream_t st, const Batch& b);
void launch_advect_v(hipStream_t st, const Batch& b);
void launch_forces_u(hipStream_t st, const Batch& b);
void launch_forces_v(hipStream_t st, const Batch& b);
void launch_divergence(hipStream_t st, const Batch& b);
void launch_gs_color(hipStream_t st, const Batch& b, int parity, float weight, int pcg);
void launch_residual(hipStream_t st, const Batch& b, int pcg);
void launch_restrict(hipStream_t st, const Batch& b, int pcg);
void launch_prolong(hipStream_t st, const Batch& b, int pcg);
void launch_jacobi(hipStream_t st, const Batch& b, int reverse);
void launch_solve_u(hipStream_t st, const Batch& b);
void launch_solve_v(hipStream_t st, const Batch& b);
void launch_extrapolate_fused(hipStream_t st, const Batch& b);
int extrapolate_tile_w();
int extrapolate_tile_h();
int extrapolate_passes();
void launch_advect_levelset(hipStream_t st, const Batch& b);
void launch_jfa_init(hipStream_t st, const Batch& b);
void launch_jfa_step(hipStream_t st, const Batch& b, int step);  // Batch.level = source seed buffer
void launch_jfa_sdf(hipStream_t st, const Batch& b);  // Batch.level = final seed buffer
void launch_fim_init_active(hipStream_t st, const Batch& b);  // K28 + the first active labels
void launch_fim_fused(hipStream_t st, const Batch& b, int src);  // Batch::level = iterations
int fim_tile_w();
int fim_tile_h();
int fim_iterations_per_launch();
void launch_ls_gradient(hipStream_t st, const Batch& b);
void launch_sample_forces(hipStream_t st, const Batch& b);
void launch_accumulate_forces(hipStream_t st, const Batch& b);
void launch_residual_norm(hipStream_t st, const Batch& b);
void launch_pcg_init(hipStream_t st, const Batch& b);
void launch_pcg_rz(hipStream_t st, const Batch& b);
void launch_pcg_apply(hipStream_t st, const Batch& b);
void launch_pcg_axpy(hipStream_t st, const Batch& b);
void launch_cg_check(hipStream_t st, const Batch& b);
// kind 0: r . z after a preconditioner, 1: p . q after k_pcg_apply; Batch =
// the CG tiles of the instances with DevInst::prefold
void launch_prefold(hipStream_t st, const Batch& b, int kind);
void launch_mask_p(hipStream_t st, const Batch& b);
void launch_cg_tally(hipStream_t st, const Batch& b);  // one block per instance
// dst[16 k + j] = scal[j] of b.inst[k], j < 16 (one block per instance)
void launch_gather_scalars(hipStream_t st, const Batch& b, double* dst);
void launch_max_divergence(hipStream_t st, const Batch& b);
void launch_fluid_diagnostics(hipStream_t st, const FluidDiagArgs& a);
// advection then forces of one velocity component, one launch (the step's path)
void launch_advect_forces_u(hipStream_t st, const Batch& b);
void launch_advect_forces_v(hipStream_t st, const Batch& b);
// u and v faces in one launch: Batch::level 0 slots are u faces, level 1 v faces
void launch_advect_forces_uv(hipStream_t st, const Batch& b);
void launch_solve_uv(hipStream_t st, const Batch& b);
void launch_integrate_bodies(hipStream_t st, const BodyStep& a);
void launch_integrate_own_bodies(hipStream_t st, const BodyLists& a);  // one block per list

}  // namespace efl

namespace efl {
// fused multigrid (efl_mg.hip)
int mg_tile_w();
int mg_tile_h();
int bottom_cells_for_device(int device);
int bottom_cells_enable(int device, int cells);
int mg_setup_tile_w();
int mg_setup_tile_h();
void launch_mg_setup(hipStream_t st, const Batch& b);
// levels b.level[n] .. nlev - 1 of each instance in one workgroup (levels of at
// most mg_setup_coarse_cells() cells)
void launch_mg_setup_coarse(hipStream_t st, const Batch& b);
int mg_setup_coarse_cells();
// capacity in cells of the LDS buffer of the level `offset` levels below the
// first one of k_mg_setup_coarse
int mg_setup_coarse_lds_cells(int offset);
size_t bottom_lds_bytes(int cells);
// th: tile height, 16 (64 x 16 tiles; the only one with CG reductions) or 8
void launch_mg_down(hipStream_t st, const Batch& b, int pcg, int nsweeps, int first, float weight, int restrict_,
                    int cg_check = 0, int th = 16);
void launch_mg_up(hipStream_t st, const Batch& b, int pcg, int nsweeps, int first, float weight, int rz, int th = 16);
// levels lvl and lvl + 1 of the up pass in one launch (k_mg_up2: 2 sweeps on
// each; Batch::sel bits 8..10 carry level lvl + 1's buffer state)
void launch_mg_up2(hipStream_t st, const Batch& b, int pcg, int first, float weight, int rz, int th);
void launch_bottom(hipStream_t st, const Batch& b, int pcg, int symmetric, int max_cells);
// max_cells: the most cells any instance's bottom levels hold
void launch_mg_coarse(hipStream_t st, const CoarseBatch& b, int pcg, int symmetric, int max_cells);
}  // namespace efl
