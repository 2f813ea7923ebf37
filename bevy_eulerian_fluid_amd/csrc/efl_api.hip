// efl_api.hip - host scheduler and C ABI of libefl (include/efl.h).
//
// The scheduler replays EulerFluidNode::run (src/euler_fluid/render_node.rs
// :193-476) for a batch of instances: every stage is one (or a few) batched
// kernel launches covering all instances that take part in it, in the
// reference's order, on the context's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/efl.h"
#include "efl_internal.h"
#include "efl_launch.h"
#include "efl_nn.h"
#include "efl_reduce.h"

using efl::Batch;
using efl::DevInst;

static_assert(sizeof(efl_projection) == 32, "efl_projection layout");
static_assert(sizeof(efl_reinit) == 8, "efl_reinit layout");
static_assert(sizeof(efl_settings) == 128, "efl_settings layout");
static_assert(sizeof(efl_source) == 32, "efl_source layout");
static_assert(sizeof(efl_force) == 16, "efl_force layout");
static_assert(sizeof(efl_obstacle) == 176, "efl_obstacle layout (std430 SolidObstacle)");
static_assert(sizeof(efl_solid_force) == 16, "efl_solid_force layout");
static_assert(sizeof(efl_stats) == 40, "efl_stats layout");
static_assert(sizeof(efl_tuning) == 36, "efl_tuning layout");

// Scratch of efl_fluid_diagnostics_noalloc: the per-16x16-block partials
// (3 floats per block).  A file-scope device array, so that entry point
// allocates nothing; each GPU holds its own copy until the process ends.
// Host code never passes the symbol itself: efl_ctx_create looks up its
// address on the context's device (hipGetSymbolAddress) and keeps it in
// efl_ctx::scratch_ptr.  It needs no clearing: every launch writes each
// partial it reads.
static __device__ __align__(256) float scratch[3 * EFL_NOALLOC_MAX_BLOCKS];

struct efl_instance;
struct StepGraph {
  std::vector<efl_instance*> insts;
  uint32_t dt_bits = 0;
  uint64_t version = 0;
  // iter[0] / iter[1]: the odd / even CG iterations (a fused iteration's
  // residual buffers alternate)
  hipGraphExec_t pre = nullptr, iter[2] = {nullptr, nullptr}, post = nullptr;
};

struct efl_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  efl_obstacle* d_obs = nullptr;
  uint32_t nobs = 0;
  uint64_t obs_version = 0;
  bool timing = false;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  // replayed steps (HIP graphs), keyed by instance list, dt and gversion;
  // gversion changes whenever launch arguments would
  std::vector<StepGraph> graphs;
  uint64_t gversion = 1;
  bool graphs_on = false;  // eager launches measured faster than replay on this runtime
  int bottom_cells = 1024;  // the bottom solve takes the levels of at most this many cells
  double* h_scal = nullptr;  // pinned, 16 doubles per batch slot
  double* d_scal = nullptr;  // device staging of read_scalars: 16 doubles per batch slot
  std::vector<efl_instance*> insts;
  efl_body* d_bodies = nullptr;  // EFL_MAX_SOLIDS records
  uint32_t nbodies = 0;          // 0 = no integration of the context's list
  float* scratch_ptr = nullptr;  // device address of `scratch` on this context's GPU
  // efl_set_async: 0 = an MGPCG step reads its stopping state back (exact
  // iteration counts); > 0 = it queues the previous solve's iterations plus
  // this margin and reads nothing back (statistics resolve lazily)
  uint32_t async_margin = 0;
  uint64_t host_waits = 0;  // blocking waits for the device inside the step machinery (efl_host_waits)
  // the coarse levels of a V-cycle as one persistent launch (k_mg_coarse):
  // EFL_COARSE_FUSE=1 at context creation; off by default (measured slower)
  bool coarse_fuse = false;
  // MGPCG iterations fused into the preconditioner's launches wherever an
  // instance allows it (fused_axpy); EFL_CG_FUSE=0 at context creation runs
  // every iteration as k_pcg_apply + k_pcg_axpy (the textbook form of CG,
  // whose rounding differs)
  bool cg_fuse = true;
  // the up pass takes levels in pairs (k_mg_up2: level l + 1's post-
  // smoothing inside level l's launch, l even) wherever both are tiled
  // 2-sweep levels; the same bits, one launch fewer per pair.
  // EFL_UP_MERGE=0 at context creation launches every level on its own.
  bool up_merge = true;
};

struct efl_instance {
  efl_ctx* ctx = nullptr;
  efl_settings s{};
  DevInst h{};
  DevInst* d = nullptr;
  void* mem = nullptr;
  size_t bytes = 0;
  efl_force* d_forces = nullptr;
  uint32_t forces_cap = 0;
  bool dirty = true;
  uint64_t obs_version = ~0ull;
  bool initialized = false;
  bool diag = false;
  std::vector<float> custom_ls;
  float* d_custom_ls = nullptr;
  efl_stats stats{};
  int jfa_max_power = 0;
  // own obstacle list (efl_set_instance_obstacles)
  bool own_obs = false;
  efl_obstacle* d_obs_own = nullptr;  // EFL_MAX_SOLIDS records
  efl_body* d_bodies_own = nullptr;   // EFL_MAX_SOLIDS records
  uint32_t nobs_own = 0;
  bool own_bodies = false;
  // per multigrid level: 1 while the current x is in Level::xs (level 0
  // shares one pair state between the pressure and the CG z)
  uint8_t xsel[efl::kMaxLevels + 1] = {};
  // per multigrid level: 1 while x is known to be zero (not read)
  uint8_t xzero[efl::kMaxLevels + 1] = {};
  uint32_t cg_last = 0;  // CG iterations of the previous solve
  // fused CG iterations (k_mg_down's SEL_AXPY): the residual buffer that
  // holds the current r (0: pcg_r, 1: pcg_r2; in a fused solve 1 before the
  // even iterations), and the launch that carries the iteration next (1: the
  // level-0 k_mg_down, 2: the level-1 one, which runs the stopping test; 0:
  // none)
  uint8_t rpar = 0, fax = 0;
  // learned preconditioner (efl_set_learned): the network and this
  // instance's activations (padded to multiples of 8)
  efl_nn* nn = nullptr;
  efl::nn::Activations nn_act{};
  uint32_t nn_iters = 0, nn_fallback = 0;  // last learned solve
  bool nn_in_fallback = false;             // MGPCG is finishing a learned solve
  float nn_rel = 0.0f;                     // ||r|| / ||b|| when the network's part ended
  int bottom_cells = 1024;                 // the bottom solve takes levels of at most this many cells
  int mg_th = 0;  // smoother tile height (16 or 8; 0: by level size) where no reduction partials are left
  // asynchronous MGPCG steps: each solve's scalars are copied to one of two
  // pinned slots (h_stat, kScalars doubles each) behind that slot's event
  // and turned into stats when asked for or when the slot is needed again
  double* h_stat = nullptr;
  hipEvent_t stat_ev[2] = {nullptr, nullptr};
  bool stat_pend[2] = {false, false};
  uint64_t stat_seq[2] = {0, 0};
  uint64_t stat_next = 0;     // sequence number of the next asynchronous solve
  bool stat_pending = false;  // some slot is pending
  uint32_t recent[4] = {};   // iteration counts of the last resolved asynchronous solves
  uint32_t n_recent = 0;
  uint32_t missed_seen = 0;  // the device tally (SC_MISSED) already added to stats.unconverged_solves
};

struct efl_nn {
  efl_ctx* ctx = nullptr;
  efl::nn::Trainer tr;
  efl::nn::Packed pk{};
  float* d_tmp = nullptr;  // two sample planes (rhs, target) for efl_nn_add_sample
  size_t tmp_floats = 0;
  double* d_hist = nullptr;
  uint32_t hist_cap = 0;
};

namespace efl {
namespace {
thread_local const char* g_launch_error = nullptr;
}
void set_launch_error(const char* msg) {
  if (!g_launch_error) g_launch_error = msg;
}
const char* take_launch_error() {
  const char* m = g_launch_error;
  g_launch_error = nullptr;
  return m;
}
}  // namespace efl

namespace {

int fail(efl_ctx* c, int code, const char* fmt, ...) {
  if (c) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    c->err = buf;
  }
  return code;
}

// After a batch of launches: an invariant a launcher found broken, else the
// runtime's launch error.
int launch_status(efl_ctx* c, const char* what) {
  if (const char* m = efl::take_launch_error()) {
    (void)hipGetLastError();
    return fail(c, EFL_E_STATE, "%s: %s", what, m);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(c, EFL_E_HIP, "%s: %s", what, hipGetErrorString(e));
  return EFL_OK;
}

#define HIPCHK(ctx, call)                                                                   \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess) return fail(ctx, EFL_E_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

// A blocking wait for the context's stream inside the step machinery
// (counted: an asynchronous MGPCG step makes none).
int host_wait(efl_ctx* c) {
  ++c->host_waits;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return EFL_OK;
}

// ---- batches ---------------------------------------------------------------
enum Extent { EX_CENTER, EX_XMAJOR, EX_YMAJOR, EX_FULL, EX_LEVEL, EX_COARSE };

void extent_of(const efl_instance* in, Extent k, int lvl, int& ex, int& ey) {
  int W = (int)in->s.width, H = (int)in->s.height;
  switch (k) {
    case EX_CENTER: ex = (W / 8) * 8; ey = (H / 8) * 8; break;      // pipeline.rs:176-178
    case EX_XMAJOR: ex = W + 1; ey = (H / 64) * 64; break;          // pipeline.rs:180-182
    case EX_YMAJOR: ex = (W / 64) * 64; ey = H + 1; break;          // pipeline.rs:184-186
    case EX_FULL: ex = W; ey = H; break;
    case EX_LEVEL: ex = in->h.lev[lvl].w; ey = in->h.lev[lvl].h; break;      // multi_grid.rs:405
    case EX_COARSE: ex = in->h.lev[lvl + 1].w; ey = in->h.lev[lvl + 1].h; break;
  }
}

// levels: per-instance Batch.level values (or -1 = use `lvl`)
Batch make_batch(const std::vector<efl_instance*>& v, Extent k, int lvl, const std::vector<int>* levels, int tw,
                 int th);
// Full-grid batch of the CG and residual kernels (64 x 16 tiles, one
// reduction partial per tile).
Batch cg_batch(const std::vector<efl_instance*>& v) {
  return make_batch(v, EX_FULL, 0, nullptr, efl::tile_w(), efl::cg_tile_h());
}
Batch make_batch(const std::vector<efl_instance*>& v, Extent k, int lvl = 0, const std::vector<int>* levels = nullptr,
                 int tw = 0, int th = 0) {
  Batch b;
  std::memset(&b, 0, sizeof(b));
  if (tw <= 0) tw = efl::tile_w();
  if (th <= 0) th = efl::tile_h();
  int n = 0, start = 0;
  for (size_t i = 0; i < v.size(); ++i) {
    int l = levels ? (*levels)[i] : lvl;
    int ex, ey;
    extent_of(v[i], k, l, ex, ey);
    int tx = (ex + tw - 1) / tw, ty = (ey + th - 1) / th;
    if (ex <= 0 || ey <= 0) tx = ty = 0;
    b.start[n] = start;
    b.tiles_x[n] = tx > 0 ? tx : 1;
    b.ex[n] = ex;
    b.ey[n] = ey;
    b.level[n] = l;
    b.inst[n] = v[i]->d;
    start += tx * ty;
    ++n;
  }
  b.n = n;
  b.start[n] = start;
  b.nblocks = start;
  return b;
}

// ---- uploads ----------------------------------------------------------------
int upload(efl_instance* in) {
  efl_ctx* c = in->ctx;
  if (in->own_obs) {
    if (in->h.obs != in->d_obs_own || in->h.nobs != in->nobs_own) {
      in->h.obs = in->d_obs_own;
      in->h.nobs = in->nobs_own;
      in->dirty = true;
    }
  } else if (in->obs_version != c->obs_version || in->h.obs != c->d_obs) {
    in->h.obs = c->d_obs;
    in->h.nobs = c->nobs;
    in->obs_version = c->obs_version;
    in->dirty = true;
  }
  const efl_projection& pj = in->s.projection;
  const uint32_t maxc = pj.max_cycles ? pj.max_cycles : 1, every = pj.check_every ? pj.check_every : 1;
  const uint32_t flex = pj.kind == EFL_PROJECTION_LEARNED && !in->nn_in_fallback ? 1u : 0u;
  if (in->h.cg_tol != pj.rel_tol || in->h.cg_maxc != maxc || in->h.cg_every != every || in->h.cg_flex != flex) {
    in->h.cg_tol = pj.rel_tol;
    in->h.cg_maxc = maxc;
    in->h.cg_every = every;
    in->h.cg_flex = flex;
    in->dirty = true;
  }
  if (!in->dirty) return EFL_OK;
  HIPCHK(c, hipMemcpyAsync(in->d, &in->h, sizeof(DevInst), hipMemcpyHostToDevice, c->stream));
  in->dirty = false;
  return EFL_OK;
}

// ---- stages ------------------------------------------------------------------
using IV = std::vector<efl_instance*>;

template <class F>
IV filter(const IV& v, F f) {
  IV o;
  for (auto* i : v)
    if (f(i)) o.push_back(i);
  return o;
}

void stage_init(hipStream_t st, const IV& v) {  // render_node.rs:219-277
  efl::launch_init_center(st, make_batch(v, EX_CENTER));
  efl::launch_init_edge(st, make_batch(v, EX_XMAJOR));
  efl::launch_fluid_source(st, make_batch(v, EX_CENTER), 1);
}

void stage_reinit(hipStream_t st, const IV& v) {  // reinitialize_levelset.rs:162-243
  IV jfa = filter(v, [](efl_instance* i) { return i->s.reinit.kind == EFL_REINIT_JUMP_FLOODING; });
  IV fim = filter(v, [](efl_instance* i) { return i->s.reinit.kind != EFL_REINIT_JUMP_FLOODING; });
  if (!jfa.empty()) {
    efl::launch_jfa_init(st, make_batch(jfa, EX_CENTER));
    int M = 0;
    for (auto* i : jfa) M = std::max(M, i->jfa_max_power);
    for (int j = M; j >= 0; --j) {  // steps 2^M .. 1 (jump_flooding.rs:200-209)
      IV part;
      std::vector<int> src;
      for (auto* i : jfa)
        if (i->jfa_max_power >= j) {
          part.push_back(i);
          src.push_back((i->jfa_max_power - j) & 1);
        }
      efl::launch_jfa_step(st, make_batch(part, EX_CENTER, 0, &src), 1 << j);
    }
    std::vector<int> fin;
    for (auto* i : jfa) fin.push_back((i->jfa_max_power + 1) & 1);
    efl::launch_jfa_sdf(st, make_batch(jfa, EX_CENTER, 0, &fin));
  }
  if (!fim.empty()) {
    efl::launch_fim_init_active(st, make_batch(fim, EX_CENTER));
    uint32_t M = 0;
    for (auto* i : fim) M = std::max(M, i->s.reinit.fim_iterations);
    if (M == 0) return;
    // an even number of temporally blocked launches (the last one writes
    // ls0/lab), each instance's iterations spread over them
    const uint32_t per = (uint32_t)efl::fim_iterations_per_launch();
    const uint32_t L = 2 * ((M + 2 * per - 1) / (2 * per));
    for (uint32_t j = 0; j < L; ++j) {
      Batch b = make_batch(fim, EX_FULL, 0, nullptr, efl::fim_tile_w(), efl::fim_tile_h());
      for (int n = 0; n < b.n; ++n) {
        const uint32_t m = fim[n]->s.reinit.fim_iterations;
        b.level[n] = (int)(m / L + (j < m % L ? 1 : 0));
      }
      efl::launch_fim_fused(st, b, (int)(j & 1));
    }
  }
}

// Tiled smoothing of one level: count(i) red-black sweeps (black-red with
// reverse_colour), up to 3 per launch; each launch flips the level's buffer
// pair.
//   DOWN: the last launch also forms the residual and restricts it to level
//         lvl + 1 (runs even for 0 sweeps).
//   UP:   the first launch applies the prolongation from level lvl + 1
//         (runs even for 0 sweeps); with rz, instances whose last launch
//         is that first one get the fused z mask + r . z partials, and are
//         returned in *rz_done with the batch used (for the finalize).
//   PLAIN: sweeps only.
enum SweepMode { PLAIN, DOWN, UP };


// Batch of a tiled multigrid launch: the tiles cover the whole level (each
// launch writes a complete x into the other buffer of the pair), `ext` only
// bounds the updated cells; sel carries the instances' buffer state.  The
// CG bits (SEL_R2, SEL_AXPY) go only to launches of the preconditioner
// (pcg): a pressure launch's b is the divergence whatever state a finished
// CG solve left behind.
// The tile height of an instance's smoother launch: 64 x 16 for level 0 of
// the CG preconditioner (its launches leave reduction partials per 64 x 16
// tile), else the instance's tuning; by default 64 x 8 on levels of fewer
// than 2^20 cells (under four 64 x 16 tiles per CU of a 256-CU device, where
// twice the workgroups hide more latency: 5.9 against 7.2 us for two sweeps
// at 256^2 on MI355X) and 64 x 16 from there (12.9 against 14.3 us at
// 1024^2, where the taller tile's smaller halo share wins).
int mg_th_of(const efl_instance* i, int lvl, int pcg) {
  if (pcg && lvl == 0) return efl::mg_tile_h();
  if (i->mg_th) return i->mg_th;
  return (size_t)i->h.lev[lvl].w * i->h.lev[lvl].h >= ((size_t)1 << 20) ? efl::mg_tile_h() : 8;
}

Batch mg_batch(const IV& v, Extent ext, int lvl, bool with_coarse, int pcg) {
  Batch b = make_batch(v, EX_LEVEL, lvl, nullptr, efl::mg_tile_w(), mg_th_of(v[0], lvl, pcg));
  for (int n = 0; n < b.n; ++n) {
    extent_of(v[n], ext, lvl, b.ex[n], b.ey[n]);
    efl_instance* i = v[n];
    int sel = i->xsel[lvl] | (with_coarse ? i->xsel[lvl + 1] << 1 : 0) | (i->xzero[lvl] << 2);
    if ((size_t)i->h.lev[lvl].w * i->h.lev[lvl].h >= ((size_t)1 << 20)) sel |= efl::SEL_LAZY;
    if (pcg && lvl == 0 && i->rpar) sel |= efl::SEL_R2;
    if (pcg && !with_coarse && i->fax == lvl + 1) {  // a down launch that carries the CG iteration
      sel |= efl::SEL_AXPY;
      if (lvl == 0) i->rpar ^= 1;  // r is written to the other buffer
      // the stopping test in the level-1 launch; an instance with
      // pre-folded partials takes k_cg_check after this launch instead
      i->fax = lvl == 0 && !i->h.prefold ? 2 : 0;
    }
    b.sel[n] = sel;
  }
  return b;
}
void mg_flip(const IV& v, int lvl) {
  for (auto* i : v) {
    i->xsel[lvl] ^= 1;
    i->xzero[lvl] = 0;
  }
}

// Stencils of level 0 (all: of every level, with the level sets and area
// fractions of the coarse levels restricted on the way), one tiled launch per
// level down to the first level above level 0 of at most
// mg_setup_coarse_cells() cells, then one launch for every instance's
// remaining levels; the inputs are fixed for the rest of the projection.
int first_coarse_setup_level(const efl_instance* i) {
  const int nlev = i->h.nlev;
  auto cells = [&](int l) { return (int64_t)i->h.lev[l].w * i->h.lev[l].h; };
  for (int l0 = 1; l0 < nlev; ++l0) {
    if (cells(l0) > efl::mg_setup_coarse_cells()) continue;
    for (int l = l0 + 1; l < nlev; ++l)
      if (cells(l) > efl::mg_setup_coarse_lds_cells(l - l0)) return nlev;  // a level would not fit in LDS
    return l0;
  }
  return nlev;
}
void mg_setup(hipStream_t st, const IV& v, bool all) {
  int top = 0;
  for (auto* i : v) top = std::max(top, all ? first_coarse_setup_level(i) : 1);
  for (int l = 0; l < top; ++l) {
    IV part = filter(v, [&](efl_instance* i) { return l < (all ? first_coarse_setup_level(i) : 1); });
    Batch b = make_batch(part, EX_LEVEL, l, nullptr, efl::mg_setup_tile_w(), efl::mg_setup_tile_h());
    for (int n = 0; n < b.n; ++n) b.sel[n] = (all && l + 1 < part[n]->h.nlev) ? 1 : 0;
    efl::launch_mg_setup(st, b);
  }
  if (!all) return;
  Batch c;
  std::memset(&c, 0, sizeof(c));
  for (auto* i : v) {
    const int l0 = first_coarse_setup_level(i);
    if (l0 >= i->h.nlev) continue;
    c.inst[c.n] = i->d;
    c.level[c.n] = l0;
    ++c.n;
  }
  efl::launch_mg_setup_coarse(st, c);
}

// Bring level 0 back into its primary buffer (the pressure, or z for the CG
// preconditioner) where the last tiled launch left it in the second one.
void mg_settle(hipStream_t st, const IV& v, int pcg) {
  for (auto* i : v) {
    if (!i->xsel[0]) continue;
    const efl::Level& L = pcg ? i->h.lev_pcg0 : i->h.lev[0];
    (void)hipMemcpyAsync(L.x, L.xs, (size_t)L.w * L.h * sizeof(float), hipMemcpyDeviceToDevice, st);
    i->xsel[0] = 0;
    i->xzero[0] = 0;
  }
}

void gs_sweeps(hipStream_t st, const IV& v, int lvl, uint32_t (*count)(efl_instance*), float weight, int pcg,
               bool reverse_colour, Extent ext = EX_LEVEL, SweepMode mode = PLAIN, bool rz = false,
               IV* rz_done = nullptr, Batch* rz_batch = nullptr, bool cg_check = false) {
  const int first = reverse_colour ? 0 : 1;
  const int th = mg_th_of(v[0], lvl, pcg);
  for (auto* i : v)
    if (mg_th_of(i, lvl, pcg) != th) {  // instances of two tile heights: one pass for each
      for (int h : {efl::mg_tile_h(), 8}) {
        IV part = filter(v, [&](efl_instance* j) { return mg_th_of(j, lvl, pcg) == h; });
        if (!part.empty())
          gs_sweeps(st, part, lvl, count, weight, pcg, reverse_colour, ext, mode, rz, rz_done, rz_batch, cg_check);
      }
      return;
    }
  // sweeps per launch: 3, or 4 for plain sweeps on 64 x 8 tiles (a fixed
  // Gauss-Seidel count on a small grid is launch-bound: 25 launches for 100
  // sweeps instead of 34); the halo grows by two cells per sweep
  const uint32_t G = (mode == PLAIN && th == 8) ? 4 : 3;
  auto chunks = [&](efl_instance* i) {  // number of launches for this instance
    uint32_t c = count(i);
    uint32_t n = (c + G - 1) / G;
    return mode == PLAIN ? n : std::max<uint32_t>(n, 1);
  };
  uint32_t K = 0;
  for (auto* i : v) K = std::max(K, chunks(i));
  for (uint32_t k = 0; k < K; ++k) {
    for (uint32_t n = 0; n <= G; ++n) {
      IV part = filter(v, [&](efl_instance* i) {
        if (k >= chunks(i)) return false;
        uint32_t c = count(i);
        uint32_t left = c > G * k ? std::min<uint32_t>(G, c - G * k) : 0;
        return left == n;
      });
      if (part.empty()) continue;
      if (mode == UP && k == 0) {
        IV fuse = filter(part, [&](efl_instance* i) { return rz && chunks(i) == 1; });
        IV rest = filter(part, [&](efl_instance* i) { return !(rz && chunks(i) == 1); });
        if (!fuse.empty()) {
          Batch bf = mg_batch(fuse, ext, lvl, true, pcg);
          efl::launch_mg_up(st, bf, pcg, (int)n, first, weight, 1, th);
          mg_flip(fuse, lvl);
          if (rz_done) *rz_done = fuse;
          if (rz_batch) *rz_batch = bf;
        }
        if (!rest.empty()) {
          efl::launch_mg_up(st, mg_batch(rest, ext, lvl, true, pcg), pcg, (int)n, first, weight, 0, th);
          mg_flip(rest, lvl);
        }
      } else {
        bool last = false;
        for (auto* i : part) last = last || (k + 1 == chunks(i));
        if (mode == DOWN && last) {
          IV fin = filter(part, [&](efl_instance* i) { return k + 1 == chunks(i); });
          IV more = filter(part, [&](efl_instance* i) { return k + 1 != chunks(i); });
          // instances that carry the fused CG update get a launch of their own (the
          // update's instantiation is slower for the others: 332 vs 222 us at 4096^2)
          IV fin_ax = filter(fin, [&](efl_instance* i) { return pcg && lvl == 0 && i->fax == 1; });
          IV fin_pl = filter(fin, [&](efl_instance* i) { return !(pcg && lvl == 0 && i->fax == 1); });
          for (const IV* grp : {&fin_ax, &fin_pl})
            if (!grp->empty())
              efl::launch_mg_down(st, mg_batch(*grp, ext, lvl, false, pcg), pcg, (int)n, first, weight, 1,
                                  cg_check && k == 0, th);
          mg_flip(fin, lvl);
          for (auto* i : fin) {  // the coarse x starts at zero (not stored)
            i->xsel[lvl + 1] = 0;
            i->xzero[lvl + 1] = 1;
          }
          if (!more.empty()) {
            efl::launch_mg_down(st, mg_batch(more, ext, lvl, false, pcg), pcg, (int)n, first, weight, 0,
                                cg_check && k == 0, th);
            mg_flip(more, lvl);
          }
        } else {
          efl::launch_mg_down(st, mg_batch(part, ext, lvl, false, pcg), pcg, (int)n, first, weight, 0, cg_check && k == 0,
                              th);
          mg_flip(part, lvl);
        }
      }
    }
  }
}

uint32_t cnt_pre(efl_instance* i) { return i->s.projection.mg_pre_iterations; }
uint32_t cnt_post(efl_instance* i) { return i->s.projection.mg_post_iterations; }
uint32_t cnt_coarse(efl_instance* i) { return i->s.projection.mg_coarsest_iterations; }
uint32_t cnt_gs(efl_instance* i) { return i->s.projection.gs_iterations; }

// First level whose grid holds at most the instance's bottom_cells cells:
// from there the rest of the cycle runs in one workgroup.  -1 if none.
int bottom_level(const efl_instance* i) {
  for (int l = 0; l < i->h.nlev; ++l)
    if (i->h.lev[l].w * i->h.lev[l].h <= i->bottom_cells) return l;
  return -1;
}

// 64 x 16 tiles of level l
int mg_tiles(const efl_instance* i, int l) {
  const efl::Level& L = i->h.lev[l];
  return ((L.w + efl::mg_tile_w() - 1) / efl::mg_tile_w()) * ((L.h + efl::mg_tile_h() - 1) / efl::mg_tile_h());
}
int bottom_cells_of(const efl_instance* i, int lb) {
  int c = 0;
  for (int q = lb; q < i->h.nlev; ++q) c += i->h.lev[q].w * i->h.lev[q].h;
  return c;
}

// First level of the fused coarse cycle (k_mg_coarse): the first level
// l >= 1 above the bottom with at most kCoarseTiles tiles (the levels below
// it have fewer), when every tiled level takes one launch each way (2 pre-
// and 2 post-smoothing sweeps) and the bottom fits the launch's LDS; -1 if
// the cycle has none or the context does not fuse (then every level is a
// launch of its own).
int coarse_level(const efl_instance* i) {
  const int lb = bottom_level(i);
  const efl_projection& p = i->s.projection;
  if (!i->ctx->coarse_fuse || lb < 2 || p.mg_pre_iterations != 2 || p.mg_post_iterations != 2) return -1;
  if (efl::bottom_lds_bytes(bottom_cells_of(i, lb)) > 40 * 1024) return -1;
  for (int l = 1; l < lb; ++l)
    if (mg_tiles(i, l) <= efl::kCoarseTiles) return lb - l <= efl::kCoarseMaxLevels ? l : -1;
  return -1;
}

// The fused coarse cycle of the instances v from level l0: the Batch::sel bits
// of every phase, and the buffer-pair flags left behind, are those the
// separate down / bottom / up launches of v_cycle would use and leave.
void coarse_cycle(hipStream_t st, const IV& v, int l0, int pcg, bool symmetric) {
  efl::CoarseBatch cb;
  std::memset(&cb, 0, sizeof(cb));
  int start = 0, cells = 0;
  for (size_t k = 0; k < v.size(); ++k) {
    efl_instance* i = v[k];
    const int lb = bottom_level(i);
    cb.start[k] = start;
    start += mg_tiles(i, l0);
    cb.inst[k] = i->d;
    cb.l0[k] = (int8_t)l0;
    cb.lb[k] = (int8_t)lb;
    cb.coarse[k] = (int8_t)std::min<uint32_t>(i->s.projection.mg_coarsest_iterations, 127);
    int ph = 0;
    for (int q = l0; q < lb; ++q) {  // down: as mg_batch + mg_flip + the coarse x starting at zero
      int sel = i->xsel[q] | (i->xzero[q] << 2);
      if (pcg && i->fax == q + 1) {  // level 1 carries the iteration's stopping test
        sel |= efl::SEL_AXPY;
        i->fax = 0;
      }
      cb.sel[k][ph++] = (uint8_t)sel;
      i->xsel[q] ^= 1;
      i->xzero[q] = 0;
      i->xsel[q + 1] = 0;
      i->xzero[q + 1] = 1;
    }
    cb.sel[k][ph++] = (uint8_t)(i->xsel[lb] | (i->xzero[lb] << 2));  // bottom
    i->xzero[lb] = 0;
    for (int q = lb - 1; q >= l0; --q) {  // up
      cb.sel[k][ph++] = (uint8_t)(i->xsel[q] | (i->xsel[q + 1] << 1) | (i->xzero[q] << 2));
      i->xsel[q] ^= 1;
      i->xzero[q] = 0;
    }
    cells = std::max(cells, bottom_cells_of(i, lb));
  }
  cb.n = (int)v.size();
  cb.start[cb.n] = start;
  efl::launch_mg_coarse(st, cb, pcg, symmetric ? 1 : 0, cells);
}

// Level l of instance i takes level l + 1's up pass into its own launch
// (k_mg_up2): l even, both levels tiled (above the bottom, the coarsest and
// a fused coarse cycle), one 2-sweep post-smoothing launch on each.
bool up_pairs(const efl_instance* i, int l) {
  if (!i->ctx->up_merge || (l & 1) || cnt_post(const_cast<efl_instance*>(i)) != 2) return false;
  const int b = bottom_level(i), c = coarse_level(i);
  const auto tiled = [&](int q) { return (b < 0 || q < b) && q < i->h.nlev - 1 && (c < 0 || q < c); };
  return tiled(l) && tiled(l + 1);
}

// The up pass of levels l and l + 1 of the instances v (up_pairs) in one
// launch per tile height; the buffer-pair state moves as the two separate
// launches would move it.  rz (level 0 of the CG preconditioner): z masked,
// r . z and z . A z partials; the instances are appended to *rz_done.
void up_pair(hipStream_t st, const IV& v, int l, int pcg, bool symmetric, bool rz, IV* rz_done) {
  for (int h : {efl::mg_tile_h(), 8}) {
    IV part = filter(v, [&](efl_instance* i) { return mg_th_of(i, l, pcg) == h; });
    if (part.empty()) continue;
    std::vector<int> selc;
    for (auto* i : part) selc.push_back(i->xsel[l + 1] | (i->xsel[l + 2] << 1) | (i->xzero[l + 1] << 2));
    Batch b = mg_batch(part, EX_LEVEL, l, true, pcg);
    for (int n = 0; n < b.n; ++n) b.sel[n] |= selc[n] << 8;
    efl::launch_mg_up2(st, b, pcg, symmetric ? 0 : 1, 1.0f, rz ? 1 : 0, h);
    mg_flip(part, l + 1);
    mg_flip(part, l);
    if (rz && rz_done) rz_done->insert(rz_done->end(), part.begin(), part.end());
  }
}

// v_cycle, multi_grid.rs:390-475 (symmetric: post-smoothing black-red and
// the coarsest level also run back in reverse colour order, see the oracle)
// The bottom solves of one V-cycle over a batch.  Instances reach their
// bottom level at different depths; each one's bottom solve is queued when
// the cycle passes its level and all of them run in one k_bottom launch
// (one workgroup per instance, each at its own level) just before the first
// up launch, which is the earliest any of them is needed.  No instance's
// data is touched between its queueing and the launch, so the results are
// those of one launch per level.
struct BottomQueue {
  Batch b;
  int cells = 0;
  BottomQueue() { std::memset(&b, 0, sizeof(b)); }
};
void queue_bottom(BottomQueue& q, efl_instance* i, int l) {
  const int n = q.b.n++;
  q.b.inst[n] = i->d;
  q.b.level[n] = l;
  q.b.sel[n] = i->xsel[l] | (i->xzero[l] << 2);
  i->xzero[l] = 0;
  q.b.ex[n] = (int)i->s.projection.mg_pre_iterations;
  q.b.ey[n] = (int)(i->s.projection.mg_post_iterations & 0xffff) | (int)(i->s.projection.mg_coarsest_iterations << 16);
  int c = 0;
  for (int k = l; k < i->h.nlev; ++k) c += i->h.lev[k].w * i->h.lev[k].h;
  q.cells = std::max(q.cells, c);
}
void flush_bottoms(hipStream_t st, BottomQueue& q, int pcg, bool symmetric) {
  if (q.b.n > 0) efl::launch_bottom(st, q.b, pcg, symmetric ? 1 : 0, q.cells);
  q = BottomQueue();
}
void v_cycle_rec(hipStream_t st, const IV& v_in, int l, int pcg, bool symmetric, BottomQueue& q) {
  if (v_in.empty()) return;
  IV v = v_in;
  if (l >= 1) {  // instances whose remaining cycle is one fused launch
    IV fused = filter(v, [l](efl_instance* i) { return coarse_level(i) == l; });
    if (!fused.empty()) {
      coarse_cycle(st, fused, l, pcg, symmetric);
      v = filter(v, [l](efl_instance* i) { return coarse_level(i) != l; });
      if (v.empty()) return;
    }
  }
  IV bot = filter(v, [l](efl_instance* i) { return bottom_level(i) == l; });
  IV coarse = filter(v, [l](efl_instance* i) { return bottom_level(i) < 0 && l == i->h.nlev - 1; });
  IV rest = filter(v, [l](efl_instance* i) {
    int b = bottom_level(i);
    return (b < 0 || l < b) && l < i->h.nlev - 1;
  });
  for (auto* i : bot) queue_bottom(q, i, l);
  if (!coarse.empty()) {
    gs_sweeps(st, coarse, l, cnt_coarse, 1.0f, pcg, false);
    if (symmetric) gs_sweeps(st, coarse, l, cnt_coarse, 1.0f, pcg, true);
  }
  if (!rest.empty()) {
    gs_sweeps(st, rest, l, cnt_pre, 1.0f, pcg, false, EX_LEVEL, DOWN);
    v_cycle_rec(st, rest, l + 1, pcg, symmetric, q);
    flush_bottoms(st, q, pcg, symmetric);  // every bottom below this level is queued by now
    // level l's up: with level l + 1's (up_pairs), on its own, or in level
    // l - 1's launch
    IV pair = filter(rest, [l](efl_instance* i) { return up_pairs(i, l); });
    IV own = filter(rest, [l](efl_instance* i) { return !up_pairs(i, l) && !(l >= 1 && up_pairs(i, l - 1)); });
    if (!pair.empty()) up_pair(st, pair, l, pcg, symmetric, false, nullptr);
    if (!own.empty()) gs_sweeps(st, own, l, cnt_post, 1.0f, pcg, symmetric, EX_LEVEL, UP);
  }
}
void v_cycle(hipStream_t st, const IV& v_in, int l, int pcg, bool symmetric) {
  if (v_in.empty()) return;
  BottomQueue q;
  v_cycle_rec(st, v_in, l, pcg, symmetric, q);
  flush_bottoms(st, q, pcg, symmetric);
  if (l == 0) mg_settle(st, v_in, pcg);
}

// z = M r (one symmetric V-cycle from z = 0), z masked to the pressure
// unknowns, rz = r . z and beta.  Instances whose level-0 post-smoothing
// fits one launch get the mask and the r . z reduction fused into it; the
// others take the separate k_pcg_rz pass.
// Instances with DevInst::prefold take k_cg_check (one block folds ||r||^2)
// rather than have every block of the launch fold it.
bool check_in_smoother(const efl_instance* i) { return !i->h.prefold && bottom_level(i) != 0 && i->h.nlev > 1; }
bool prefolds(const efl_instance* i) { return i->h.prefold != 0; }
// the fused CG update (cg_iteration): level 0 pre-smooths in one 2-sweep
// launch and level 1 is a tiled down level too (not the bottom or the
// coarsest), post-smoothing in one launch (so r . z comes from k_mg_up)
// (with pre-folded partials too: the launch then reads gamma and delta from
// k_prefold's scalars, and one k_cg_check block after it runs the stopping
// test, the same folds in the same order as the inline ones)
bool fused_axpy(efl_instance* i) {
  const int b = bottom_level(i);
  return (check_in_smoother(i) || (i->h.prefold && b != 0 && i->h.nlev > 1)) && cnt_pre(i) == 2 &&
         cnt_post(i) <= 3 && (b > 1 || (b < 0 && i->h.nlev > 2));
}
void reset_fused(const IV& v) {  // r is in pcg_r after k_residual_norm(pcg_init)
  for (auto* i : v) i->rpar = i->fax = 0;
}
void prefold(hipStream_t st, const IV& v, int kind) {
  IV pf = filter(v, prefolds);
  if (!pf.empty()) efl::launch_prefold(st, cg_batch(pf), kind);
}

// with cg_check, the first tiled launch of level 0 runs the CG stopping test
// (instances that start in the bottom solver have run k_cg_check instead)
void precondition(hipStream_t st, const IV& v, bool cg_check = false) {
  for (auto* i : v) {  // z starts at zero: the first launch reads nothing
    i->xsel[0] = 0;
    i->xzero[0] = 1;
  }
  IV bot = filter(v, [](efl_instance* i) { return bottom_level(i) == 0; });
  IV tiled = filter(v, [](efl_instance* i) { return bottom_level(i) != 0; });
  if (!bot.empty()) v_cycle(st, bot, 0, 1, true);
  IV fused;
  if (!tiled.empty()) {
    // level 0 by hand so that its post-smoothing can carry r . z
    IV coarse = filter(tiled, [](efl_instance* i) { return bottom_level(i) < 0 && i->h.nlev == 1; });
    IV rest = filter(tiled, [](efl_instance* i) { return !(bottom_level(i) < 0 && i->h.nlev == 1); });
    if (!coarse.empty()) {
      gs_sweeps(st, coarse, 0, cnt_coarse, 1.0f, 1, false);
      gs_sweeps(st, coarse, 0, cnt_coarse, 1.0f, 1, true);
    }
    if (!rest.empty()) {
      // instances whose level-0 launch carries the CG iteration with
      // pre-folded partials: their stopping test follows that launch
      IV pfc = filter(rest, [](efl_instance* i) { return i->fax == 1 && i->h.prefold; });
      gs_sweeps(st, rest, 0, cnt_pre, 1.0f, 1, false, EX_LEVEL, DOWN, false, nullptr, nullptr, cg_check);
      if (!pfc.empty()) efl::launch_cg_check(st, cg_batch(pfc));
      v_cycle(st, rest, 1, 1, true);  // leaves level 1's up to level 0 where they pair
      IV pair = filter(rest, [](efl_instance* i) { return up_pairs(i, 0); });
      IV own = filter(rest, [](efl_instance* i) { return !up_pairs(i, 0); });
      if (!pair.empty()) up_pair(st, pair, 0, 1, true, true, &fused);
      if (!own.empty()) {
        IV f;
        gs_sweeps(st, own, 0, cnt_post, 1.0f, 1, true, EX_LEVEL, UP, true, &f);
        fused.insert(fused.end(), f.begin(), f.end());
      }
    }
  }
  mg_settle(st, v, 1);
  IV sep = filter(v, [&](efl_instance* i) { return std::find(fused.begin(), fused.end(), i) == fused.end(); });
  if (!sep.empty()) efl::launch_pcg_rz(st, cg_batch(sep));
  prefold(st, v, 0);
}

// The first 16 scalars of every instance of v into c->h_scal (16 per
// instance, in v's order), then a wait: one copy for a single instance, one
// gather launch and one copy for several (a copy per instance costs a copy
// launch each: 64 of them took ~0.4 ms).
int read_scalars(efl_ctx* c, const IV& v) {
  if (v.size() > (size_t)EFL_MAX_BATCH) return fail(c, EFL_E_RANGE, "%zu instances in one read", v.size());
  if (v.empty()) return EFL_OK;
  if (v.size() == 1) {
    HIPCHK(c, hipMemcpyAsync(c->h_scal, v[0]->h.scal, 16 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    return host_wait(c);
  }
  efl::Batch b;
  std::memset(&b, 0, sizeof(b));
  b.n = (int)v.size();
  for (size_t j = 0; j < v.size(); ++j) b.inst[j] = v[j]->d;
  efl::launch_gather_scalars(c->stream, b, c->d_scal);
  HIPCHK(c, hipMemcpyAsync(c->h_scal, c->d_scal, 16 * v.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  return host_wait(c);
}

double rel_of(const double* sc) {
  double rr = sc[2], bb = sc[3];
  return bb > 0.0 ? std::sqrt(rr / bb) : std::sqrt(rr);
}

// ---- projection (projection.rs:89-164) ----------------------------------------
bool mg_needs_host(const efl_instance* i) {  // multigrid repeated to a residual target
  return i->s.projection.kind == EFL_PROJECTION_MULTIGRID && i->s.projection.rel_tol > 0.0f;
}
bool is_cg(const efl_instance* i) { return i->s.projection.kind == EFL_PROJECTION_MGPCG; }
bool is_learned(const efl_instance* i) { return i->s.projection.kind == EFL_PROJECTION_LEARNED; }
uint32_t max_cycles(const efl_instance* i) { return i->s.projection.max_cycles ? i->s.projection.max_cycles : 1; }

// The solves that need no host decision: Gauss-Seidel (gauss_seidel.rs:
// 120-148, weight 1.9), Jacobi (projection.rs:135-143) and multigrid with a
// fixed cycle count.  Enqueue only.
void proj_fixed(hipStream_t st, const IV& v) {
  IV gs = filter(v, [](efl_instance* i) { return i->s.projection.kind == EFL_PROJECTION_GAUSS_SEIDEL; });
  IV jac = filter(v, [](efl_instance* i) { return i->s.projection.kind == EFL_PROJECTION_JACOBI; });
  IV mg = filter(v, [](efl_instance* i) {
    return i->s.projection.kind == EFL_PROJECTION_MULTIGRID && !mg_needs_host(i);
  });
  if (!gs.empty()) {
    mg_setup(st, gs, false);
    gs_sweeps(st, gs, 0, cnt_gs, 1.9f, 0, false, EX_CENTER);
    mg_settle(st, gs, 0);
  }
  if (!jac.empty()) {
    Batch b = make_batch(jac, EX_CENTER);
    for (int k = 0; k < 50; ++k) {
      efl::launch_jacobi(st, b, 0);
      efl::launch_jacobi(st, b, 1);
    }
  }
  if (!mg.empty()) {
    mg_setup(st, mg, true);
    for (uint32_t k = 0;; ++k) {
      IV act = filter(mg, [k](efl_instance* i) { return k < max_cycles(i); });
      if (act.empty()) break;
      v_cycle(st, act, 0, 0, false);
    }
  }
}
void proj_fixed_stats(const IV& v) {
  for (auto* i : v) {
    i->stats.cycles = 0;
    i->stats.converged = 1;
    i->stats.rel_residual = -1.0f;
    switch (i->s.projection.kind) {
      case EFL_PROJECTION_GAUSS_SEIDEL: i->stats.cycles = i->s.projection.gs_iterations; break;
      case EFL_PROJECTION_JACOBI: i->stats.cycles = 100; break;
      case EFL_PROJECTION_MULTIGRID: i->stats.cycles = max_cycles(i); break;
      default: break;
    }
  }
}

// Multigrid repeated until ||r|| / ||b|| <= rel_tol (checked every
// check_every cycles on the host).
int proj_mg_tol(efl_ctx* c, const IV& mg) {
  hipStream_t st = c->stream;
  mg_setup(st, mg, true);
  IV act = mg;
  uint32_t k = 0;
  while (!act.empty()) {
    v_cycle(st, act, 0, 0, false);
    ++k;
    for (auto* i : act) i->stats.cycles = k;
    IV keep, chk;
    for (auto* i : act) {
      uint32_t every = i->s.projection.check_every ? i->s.projection.check_every : 1;
      if (k % every == 0 || k >= max_cycles(i)) chk.push_back(i);
      else if (k < max_cycles(i)) keep.push_back(i);
    }
    if (!chk.empty()) {
      efl::launch_residual_norm(st, cg_batch(chk));
      int rc = read_scalars(c, chk);
      if (rc) return rc;
      for (size_t j = 0; j < chk.size(); ++j) {
        efl_instance* i = chk[j];
        double rel = rel_of(c->h_scal + 16 * j);
        i->stats.rel_residual = (float)rel;
        bool ok = rel <= i->s.projection.rel_tol;
        i->stats.converged = ok;
        if (!ok && k < max_cycles(i)) keep.push_back(i);
      }
    }
    act.swap(keep);
  }
  return EFL_OK;
}

// CG preconditioned by one symmetric V-cycle.  The stopping test runs on the
// device (k_cg_check after k_pcg_axpy) and every kernel of a stopped
// instance returns at once, so iterations are queued ahead and the host only
// reads the flags between batches.
void cg_begin(hipStream_t st, const IV& cg) {
  mg_setup(st, cg, true);
  Batch full = cg_batch(cg);
  efl::launch_mask_p(st, full);
  efl::launch_pcg_init(st, full);  // r = b - A p0, the initial test
  reset_fused(cg);
  precondition(st, cg);
}
// With fuse, instances whose first two down launches are tiled 2-sweep
// launches run the whole iteration (k_pcg_apply's and k_pcg_axpy's work, in
// the Chronopoulos-Gear form) in the level-0 one and the stopping test in the
// level-1 one.  Such an instance's residual changes buffer every iteration
// (the host tracks it), so a replayed graph captures an odd and an even
// iteration.
void cg_iteration(hipStream_t st, const IV& cg, bool fuse) {
  IV fx = fuse ? filter(cg, fused_axpy) : IV{};
  if (fx.size() < cg.size()) {
    IV plain = filter(cg, [&](efl_instance* i) { return std::find(fx.begin(), fx.end(), i) == fx.end(); });
    Batch b = cg_batch(plain);
    efl::launch_pcg_apply(st, b);  // p = z + beta p, q = A p
    prefold(st, plain, 1);
    efl::launch_pcg_axpy(st, b);  // x, r, ||r||^2
  }
  // the others run the whole iteration (p, A p, x, r) in the level-0 launch
  // of the preconditioner
  for (auto* i : fx) i->fax = 1;
  IV sep = filter(cg, [&](efl_instance* i) {
    return !check_in_smoother(i) && std::find(fx.begin(), fx.end(), i) == fx.end();
  });
  if (!sep.empty()) efl::launch_cg_check(st, cg_batch(sep));
  precondition(st, cg, true);  // the stopping test, z = M r, r . z
}
// Queue as many iterations as the previous solve took, then one at a time
// until every instance has stopped; `iter` replays one iteration (a graph)
// or is null (eager launches).
// The stopping state of the asynchronous solve in slot k, once its copy has
// landed (block: wait for it).  Sets the statistics and the next budget's
// estimate.
int resolve_slot(efl_instance* i, int k, bool block) {
  if (!i->stat_pend[k]) return EFL_OK;
  efl_ctx* c = i->ctx;
  if (block) {
    HIPCHK(c, hipEventSynchronize(i->stat_ev[k]));
  } else if (hipEventQuery(i->stat_ev[k]) != hipSuccess) {
    (void)hipGetLastError();  // hipErrorNotReady is not an error here
    return EFL_OK;
  }
  const double* sc = i->h_stat + (size_t)k * efl::kScalars;
  i->stats.cycles = (uint32_t)sc[efl::SC_ITER];
  i->stats.converged = sc[efl::SC_DONE] != 0.0 && sc[efl::SC_CONV] != 0.0;
  i->stats.rel_residual = (float)sc[efl::SC_REL];
  const uint32_t tally = (uint32_t)sc[efl::SC_MISSED];
  i->stats.unconverged_solves += tally - i->missed_seen;
  i->missed_seen = tally;
  i->cg_last = i->stats.cycles;
  i->recent[i->n_recent++ % 4] = i->stats.cycles;
  i->stat_pend[k] = false;
  i->stat_pending = i->stat_pend[0] || i->stat_pend[1];
  return EFL_OK;
}
// Both slots, the older solve first.
int resolve_stats(efl_instance* i, bool block) {
  const int first = i->stat_seq[0] <= i->stat_seq[1] ? 0 : 1;
  for (int k : {first, 1 - first}) {
    int rc = resolve_slot(i, k, block);
    if (rc) return rc;
  }
  return EFL_OK;
}

// Asynchronous MGPCG (efl_set_async): queue the largest of the last four
// resolved iteration counts plus the margin (at most max_cycles) and copy the
// stopping state to pinned memory behind an event.  The step never waits for
// its own solve: iterations after the done flag rose return at once, so
// whenever the budget covers the solve the result is the synchronous one; a
// solve that needs more stops at the budget, unconverged (counted on the
// device, k_cg_tally).  Each instance keeps at most two solves in flight: the
// slot this solve reuses holds the solve before last, whose statistics are
// taken first (a wait only if the device has not finished that step), which
// keeps the estimate at most two steps old.
// iter: null (eager launches) or the graphs of the odd and even iterations.
int cg_run_async(efl_ctx* c, const IV& cg, const hipGraphExec_t* iter) {
  hipStream_t st = c->stream;
  uint32_t n = 0;
  std::vector<uint32_t> want;
  for (auto* i : cg) {
    if (!i->h_stat) {
      HIPCHK(c, hipHostMalloc((void**)&i->h_stat, 2 * efl::kScalars * sizeof(double), hipHostMallocDefault));
      for (hipEvent_t& e : i->stat_ev) HIPCHK(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    int rc = resolve_stats(i, false);
    if (rc) return rc;
    const int k = (int)(i->stat_next & 1);
    if (i->stat_pend[k]) {  // the solve before last is still running: frames in flight
      ++c->host_waits;
      rc = resolve_slot(i, k, true);
      if (rc) return rc;
    }
    uint32_t base = i->cg_last ? i->cg_last : 4u;
    for (uint32_t q = 0; q < std::min(i->n_recent, 4u); ++q) base = std::max(base, i->recent[q]);
    want.push_back(std::min(base + c->async_margin, i->h.cg_maxc));
    n = std::max(n, want.back());
  }
  // each instance takes part in its own budget of iterations (eager path;
  // a replayed graph advances the whole set), as in cg_run
  for (uint32_t k = 0; k < n; ++k) {
    if (iter) {
      HIPCHK(c, hipGraphLaunch(iter[k & 1], st));
    } else {
      IV sub;
      for (size_t j = 0; j < cg.size(); ++j)
        if (want[j] > k) sub.push_back(cg[j]);
      cg_iteration(st, sub, c->cg_fuse);
    }
  }
  efl::launch_cg_tally(st, cg_batch(cg));
  for (auto* i : cg) {
    const int k = (int)(i->stat_next & 1);
    HIPCHK(c, hipMemcpyAsync(i->h_stat + (size_t)k * efl::kScalars, i->h.scal, efl::kScalars * sizeof(double),
                             hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipEventRecord(i->stat_ev[k], st));
    i->stat_pend[k] = true;
    i->stat_seq[k] = i->stat_next++;
    i->stat_pending = true;
    i->rpar = i->fax = 0;
  }
  return EFL_OK;
}

// Each instance is queued as many iterations as its previous solve took (4
// the first time); the eager path leaves an instance out of the queued
// iterations beyond its own count, since a stopped instance's kernels return
// at once but its workgroups are still dispatched (in a batch whose solves
// end at different iterations that was most of the tail's time).  Every
// instance keeps its own iteration sequence (rpar / fax change only in the
// launches it takes part in), so leaving it out changes none of its results.
// Then the flags are read and the instances that have not stopped continue
// one iteration at a time.  A replayed graph advances the whole set.
int cg_run(efl_ctx* c, const IV& cg, const hipGraphExec_t* iter, bool sync = false) {
  if (c->async_margin && !sync) return cg_run_async(c, cg, iter);
  hipStream_t st = c->stream;
  uint32_t n = 0, queued = 0;
  std::vector<uint32_t> want(cg.size());
  for (size_t j = 0; j < cg.size(); ++j) {
    efl_instance* i = cg[j];
    want[j] = std::min(i->cg_last ? i->cg_last : 4u, i->h.cg_maxc);
    n = std::max(n, want[j]);
  }
  IV act = cg;
  bool first = true;
  while (!act.empty()) {
    for (uint32_t k = 0; k < n; ++k, ++queued) {
      if (iter) {
        HIPCHK(c, hipGraphLaunch(iter[queued & 1], st));
      } else if (first) {
        IV sub;
        for (size_t j = 0; j < cg.size(); ++j)
          if (want[j] > k) sub.push_back(cg[j]);
        cg_iteration(st, sub, c->cg_fuse);
      } else {
        cg_iteration(st, act, c->cg_fuse);
      }
    }
    first = false;
    n = 1;
    int rc = read_scalars(c, act);
    if (rc) return rc;
    IV keep;
    for (size_t j = 0; j < act.size(); ++j) {
      const double* sc = c->h_scal + 16 * j;
      efl_instance* i = act[j];
      if (sc[efl::SC_DONE] == 0.0) {
        keep.push_back(i);
        continue;
      }
      // once the done flag is up every kernel of the iteration returns at
      // once, so these are the solve's final values
      i->stats.cycles = (uint32_t)sc[efl::SC_ITER];
      i->stats.converged = sc[efl::SC_CONV] != 0.0;
      i->stats.rel_residual = (float)sc[efl::SC_REL];
      i->cg_last = i->stats.cycles;
      i->rpar = i->fax = 0;  // the next solve starts from k_pcg_init's pcg_r
    }
    if (iter) keep = keep.empty() ? keep : act;  // the graph advances the whole set
    act.swap(keep);
  }
  return EFL_OK;
}

// ---- learned preconditioner ---------------------------------------------------
void nn_free_activations(efl::nn::Activations& a) {
  for (int k = 0; k < efl::nn::kTensors; ++k) {
    if (a.t[k]) (void)hipFree(a.t[k]);
    if (a.g[k]) (void)hipFree(a.g[k]);
    a.t[k] = a.g[k] = nullptr;
  }
  if (a.scratch) (void)hipFree(a.scratch);
  a.scratch = nullptr;
}

// forced = true runs the network even when the solve has stopped (timing)
void nn_precondition(hipStream_t st, efl_instance* i, bool forced = false) {
  efl::nn::precondition(st, i->nn->pk, i->nn_act, i->d, i->s.width, i->s.height,
                        forced ? nullptr : i->h.scal + efl::SC_DONE);
  if (!forced) prefold(st, IV{i}, 0);
}

// Flexible CG preconditioned by the network.  The host reads the residual
// every few iterations; when it has fallen by less than a factor 0.7 over the
// last 8 iterations (stall), or max_cycles pass without convergence, the
// solve continues as MGPCG from the current iterate.
int proj_learned(efl_ctx* c, efl_instance* i) {
  if (!i->nn) return fail(c, EFL_E_STATE, "learned projection without a network (efl_set_learned)");
  hipStream_t st = c->stream;
  IV v{i};
  mg_setup(st, v, true);  // stencil and masks (and the fallback's pyramid)
  Batch full = cg_batch(v);
  efl::launch_mask_p(st, full);
  efl::launch_pcg_init(st, full);  // r = b - A p0, ||r||, the initial test
  reset_fused(v);
  nn_precondition(st, i);
  const uint32_t cap = max_cycles(i);
  std::vector<std::pair<uint32_t, double>> seen;  // (iteration, ||r|| / ||b||) at each read
  // queue what the last solve needed when it converged without help, else 4
  uint32_t n = std::min(std::max(i->nn_fallback == 0 ? i->nn_iters : 0u, 4u), cap);
  bool converged = false, stalled = false;
  uint32_t k = 0;
  double rel = 0.0;
  for (;;) {
    for (uint32_t j = 0; j < n; ++j) {
      efl::launch_pcg_apply(st, full);
      prefold(st, v, 1);
      efl::launch_pcg_axpy(st, full);
      efl::launch_cg_check(st, full);
      nn_precondition(st, i);
    }
    int rc = read_scalars(c, v);
    if (rc) return rc;
    const double* sc = c->h_scal;
    k = (uint32_t)sc[efl::SC_ITER];
    rel = rel_of(sc);
    const bool done = sc[efl::SC_DONE] != 0.0;
    converged = done && sc[efl::SC_CONV] != 0.0 && sc[efl::SC_REL] <= (double)i->s.projection.rel_tol;
    if (done) break;
    stalled = !std::isfinite(rel);
    for (auto it = seen.rbegin(); !stalled && it != seen.rend(); ++it)
      if (k >= it->first + 8) {
        stalled = rel > 0.7 * it->second;
        break;
      }
    if (stalled) break;
    seen.emplace_back(k, rel);
    n = std::min(4u, cap > k ? cap - k : 1u);
  }
  i->nn_iters = k;
  i->nn_fallback = 0;
  i->nn_rel = (float)rel;
  i->stats.cycles = k;
  i->stats.converged = converged;
  i->stats.rel_residual = (float)rel;
  if (converged) return EFL_OK;
  // fallback: MGPCG from the current iterate (from zero if it is not finite)
  i->nn_in_fallback = true;
  int rc = upload(i);
  if (rc) return rc;
  if (!std::isfinite(rel)) HIPCHK(c, hipMemsetAsync(i->h.p0, 0, (size_t)i->s.width * i->s.height * 4, st));
  efl::launch_pcg_init(st, full);
  reset_fused(v);
  precondition(st, v);
  const uint32_t keep = i->cg_last;
  rc = cg_run(c, v, nullptr, true);
  i->nn_in_fallback = false;
  if (rc) return rc;
  rc = upload(i);
  if (rc) return rc;
  i->nn_fallback = i->stats.cycles;
  i->cg_last = keep;
  i->stats.cycles += k;
  return EFL_OK;
}

// A fluid region with no Dirichlet term (no open face to air, to a fully
// solid fluid cell, or out of the grid) makes the pressure system singular,
// and inconsistent unless its divergence sums to zero; fluid trapped between
// bodies that close on each other breaks that, and CG then diverges.  After
// a solve that did not converge, the host labels the active cells (linked
// across open faces) of the instance, subtracts the mean divergence of every
// such sealed component (double sum in raster order, one float subtraction
// per cell) and returns the number of components changed.  A rare path: it
// reads the level set, area fractions and divergence back.
int repair_sealed(efl_ctx* c, efl_instance* i, int* changed) {
  *changed = 0;
  const int W = (int)i->s.width, H = (int)i->s.height;
  const size_t n = (size_t)W * H;
  std::vector<float> ls(n), div(n), af(4 * n);
  hipStream_t st = c->stream;
  HIPCHK(c, hipMemcpyAsync(ls.data(), i->h.ls0, n * 4, hipMemcpyDeviceToHost, st));
  HIPCHK(c, hipMemcpyAsync(div.data(), i->h.div, n * 4, hipMemcpyDeviceToHost, st));
  HIPCHK(c, hipMemcpyAsync(af.data(), i->h.afs, n * 16, hipMemcpyDeviceToHost, st));
  int rc = host_wait(c);
  if (rc) return rc;
  auto active = [&](size_t k) {
    return ls[k] < 0.0f && !(af[4 * k] == 0.0f && af[4 * k + 1] == 0.0f && af[4 * k + 2] == 0.0f && af[4 * k + 3] == 0.0f);
  };
  std::vector<uint8_t> seen(n, 0);
  std::vector<size_t> todo, cells;
  const int dxs[4] = {-1, 1, 0, 0}, dys[4] = {0, 0, -1, 1};
  for (size_t s0 = 0; s0 < n; ++s0) {
    if (seen[s0] || !active(s0)) continue;
    bool sealed = true;
    cells.clear();
    todo.assign(1, s0);
    seen[s0] = 1;
    while (!todo.empty()) {
      const size_t k = todo.back();
      todo.pop_back();
      cells.push_back(k);
      const int x = (int)(k % W), y = (int)(k / W);
      for (int f = 0; f < 4; ++f) {
        if (!(af[4 * k + f] > 0.0f)) continue;
        const int jx = x + dxs[f], jy = y + dys[f];
        if (jx < 0 || jy < 0 || jx >= W || jy >= H) continue;
        const size_t j = (size_t)jy * W + jx;
        if (!active(j)) {
          sealed = false;
          continue;
        }
        if (!seen[j]) {
          seen[j] = 1;
          todo.push_back(j);
        }
      }
    }
    if (!sealed) continue;
    std::sort(cells.begin(), cells.end());
    double sum = 0.0;
    for (size_t k : cells) sum += (double)div[k];
    const float mean = (float)(sum / (double)cells.size());
    if (mean == 0.0f) continue;
    for (size_t k : cells) div[k] = div[k] - mean;
    ++*changed;
  }
  if (*changed) HIPCHK(c, hipMemcpyAsync(i->h.div, div.data(), n * 4, hipMemcpyHostToDevice, st));
  return EFL_OK;
}

// MGPCG instances whose solve did not converge: repair sealed regions and
// solve again from zero (the iteration counts add up).
int resolve_sealed(efl_ctx* c, const IV& cg) {
  IV again;
  std::vector<uint32_t> first;
  for (auto* i : cg) {
    if (i->stats.converged) continue;
    int changed = 0;
    int rc = repair_sealed(c, i, &changed);
    if (rc) return rc;
    if (!changed) continue;
    HIPCHK(c, hipMemsetAsync(i->h.p0, 0, (size_t)i->s.width * i->s.height * 4, c->stream));
    again.push_back(i);
    first.push_back(i->stats.cycles);
  }
  if (again.empty()) return EFL_OK;
  cg_begin(c->stream, again);
  int rc = cg_run(c, again, nullptr, true);
  if (rc) return rc;
  for (size_t k = 0; k < again.size(); ++k) again[k]->stats.cycles += first[k];
  return EFL_OK;
}

int stage_projection(efl_ctx* c, const IV& v) {
  hipStream_t st = c->stream;
  IV fixed = filter(v, [](efl_instance* i) { return !is_cg(i) && !mg_needs_host(i) && !is_learned(i); });
  IV mgt = filter(v, [](efl_instance* i) { return mg_needs_host(i); });
  IV cg = filter(v, [](efl_instance* i) { return is_cg(i); });
  proj_fixed_stats(v);
  proj_fixed(st, fixed);
  if (!mgt.empty()) {
    int rc = proj_mg_tol(c, mgt);
    if (rc) return rc;
  }
  if (!cg.empty()) {
    cg_begin(st, cg);
    int rc = cg_run(c, cg, nullptr);
    if (rc) return rc;
    if (!c->async_margin) {  // the repair needs the outcome on the host
      rc = resolve_sealed(c, cg);
      if (rc) return rc;
    }
  }
  for (auto* i : v)
    if (is_learned(i)) {
      int rc = proj_learned(c, i);
      if (rc) return rc;
    }
  return EFL_OK;
}

// render_node.rs:540-608: the validity masks, then 10 passes per component.
// u and v are independent: one launch covers both (Batch::level = the
// component), two when the instance list fills more than half a batch.
// The u-face batch and the v-face batch of v (Batch::level 0 and 1).  When
// both fit one batch, *uv gets them appended (u slots first) and the result
// is true; kernels that take both components then run in one launch.
bool uv_batches(const IV& v, int tw, int th, Batch& bu, Batch& bv, Batch& uv) {
  bu = make_batch(v, EX_XMAJOR, 0, nullptr, tw, th);
  bv = make_batch(v, EX_YMAJOR, 0, nullptr, tw, th);
  for (int n = 0; n < bv.n; ++n) bv.level[n] = 1;
  if (bu.n + bv.n > efl::kMaxBatch) return false;
  uv = bu;  // append the v slots
  const int base = bu.start[bu.n];
  for (int n = 0; n < bv.n; ++n) {
    const int k = bu.n + n;
    uv.start[k] = base + bv.start[n];
    uv.tiles_x[k] = bv.tiles_x[n];
    uv.ex[k] = bv.ex[n];
    uv.ey[k] = bv.ey[n];
    uv.level[k] = 1;
    uv.sel[k] = 0;
    uv.inst[k] = bv.inst[n];
  }
  uv.n = bu.n + bv.n;
  uv.start[uv.n] = base + bv.start[bv.n];
  uv.nblocks = uv.start[uv.n];
  return true;
}

void stage_extrapolate(hipStream_t st, const IV& v) {
  Batch bu, bv, b;
  if (uv_batches(v, efl::extrapolate_tile_w(), efl::extrapolate_tile_h(), bu, bv, b)) {
    efl::launch_extrapolate_fused(st, b);
  } else {
    efl::launch_extrapolate_fused(st, bu);
    efl::launch_extrapolate_fused(st, bv);
  }
}

void stage_fluid_to_solid(hipStream_t st, const IV& v) {  // render_node.rs:610-641
  const Batch per = make_batch(v, EX_CENTER, 0, nullptr, efl::tile_w(), efl::sample_forces_tile_h());
  efl::launch_sample_forces(st, per);
  efl::launch_accumulate_forces(st, per);
}

int measure_after_projection(efl_ctx* c, const IV& v) {
  IV d = filter(v, [](efl_instance* i) { return i->diag && i->stats.rel_residual < 0.0f; });
  if (d.empty()) return EFL_OK;
  efl::launch_residual_norm(c->stream, cg_batch(d));
  int rc = read_scalars(c, d);
  if (rc) return rc;
  for (size_t j = 0; j < d.size(); ++j) d[j]->stats.rel_residual = (float)rel_of(c->h_scal + 16 * j);
  return EFL_OK;
}

int measure_divergence(efl_ctx* c, const IV& v) {
  IV d = filter(v, [](efl_instance* i) { return i->diag; });
  if (d.empty()) return EFL_OK;
  Batch b = make_batch(d, EX_FULL);
  efl::launch_max_divergence(c->stream, b);
  std::vector<float> host(4 * d.size());
  for (size_t j = 0; j < d.size(); ++j)
    HIPCHK(c, hipMemcpyAsync(&host[4 * j], d[j]->h.diag, 4 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  int rc = host_wait(c);
  if (rc) return rc;
  for (size_t j = 0; j < d.size(); ++j) d[j]->stats.max_norm_divergence = d[j]->h.dt * host[4 * j];
  return EFL_OK;
}

int run_init(efl_ctx* c, const IV& v) {
  hipStream_t st = c->stream;
  stage_init(st, v);
  for (auto* i : v) {
    if (!i->custom_ls.empty()) {
      // min-combine the staged level set into levelset_air1 (host round trip
      // at initialisation only)
      size_t n = (size_t)i->s.width * i->s.height;
      std::vector<float> cur(n);
      HIPCHK(c, hipMemcpyAsync(cur.data(), i->h.ls1, n * 4, hipMemcpyDeviceToHost, st));
      int rc = host_wait(c);
      if (rc) return rc;
      for (size_t k = 0; k < n; ++k) cur[k] = std::fmin(cur[k], i->custom_ls[k]);
      HIPCHK(c, hipMemcpyAsync(i->h.ls1, cur.data(), n * 4, hipMemcpyHostToDevice, st));
      rc = host_wait(c);
      if (rc) return rc;
    }
  }
  stage_reinit(st, v);
  for (auto* i : v) {
    i->initialized = true;
    i->stats.steps = 0;
  }
  return EFL_OK;
}

// ---- the step (render_node.rs:278-466) ---------------------------------------
// solids .. divergence (render_node.rs:278-330)
void phase_pre(hipStream_t st, const IV& v) {
  Batch center = make_batch(v, EX_CENTER), bx, by, uv;
  const bool one = uv_batches(v, 0, 0, bx, by, uv);
  efl::launch_update_solid(st, center);
  efl::launch_area_fraction(st, center);
  if (one) {
    efl::launch_advect_forces_uv(st, uv);  // K6 + K8, K7 + K9
  } else {
    efl::launch_advect_forces_u(st, bx);  // K6 + K8
    efl::launch_advect_forces_v(st, by);  // K7 + K9
  }
  efl::launch_divergence(st, center);
}
// velocity update (render_node.rs:331-350)
void phase_velocity(hipStream_t st, const IV& v) {
  Batch bx, by, uv;
  if (uv_batches(v, 0, 0, bx, by, uv)) {
    efl::launch_solve_uv(st, uv);
  } else {
    efl::launch_solve_u(st, bx);
    efl::launch_solve_v(st, by);
  }
}
// extrapolation .. fluid -> solid (render_node.rs:351-466)
void phase_post(hipStream_t st, const IV& v) {
  Batch center = make_batch(v, EX_CENTER);
  stage_extrapolate(st, v);
  efl::launch_advect_levelset(st, center);
  efl::launch_fluid_source(st, center, 0);
  stage_reinit(st, v);
  efl::launch_ls_gradient(st, center);
  stage_fluid_to_solid(st, v);
}
void step_done(const IV& v) {
  for (auto* i : v) {
    i->stats.steps += 1;
    // a synchronous solve's outcome is known here (asynchronous ones are
    // counted on the device and added when their statistics resolve)
    if (!i->stat_pending && !i->stats.converged) i->stats.unconverged_solves += 1;
    if (i->h.nforces) {  // apply_forces_and_clear (apply_forces.rs:157-167)
      i->h.nforces = 0;
      i->dirty = true;
    }
  }
}

int run_step(efl_ctx* c, const IV& v) {
  hipStream_t st = c->stream;
  bool timed = c->timing && v.size() == 1;
  if (timed) HIPCHK(c, hipEventRecord(c->ev[0], st));
  phase_pre(st, v);
  if (timed) HIPCHK(c, hipEventRecord(c->ev[1], st));
  int rc = stage_projection(c, v);
  if (rc) return rc;
  if (timed) HIPCHK(c, hipEventRecord(c->ev[2], st));
  rc = measure_after_projection(c, v);
  if (rc) return rc;
  phase_velocity(st, v);
  rc = measure_divergence(c, v);
  if (rc) return rc;
  phase_post(st, v);
  if (timed) HIPCHK(c, hipEventRecord(c->ev[3], st));
  step_done(v);
  if (timed) {
    ++c->host_waits;
    HIPCHK(c, hipEventSynchronize(c->ev[3]));
    float a = 0, b = 0;
    hipEventElapsedTime(&a, c->ev[1], c->ev[2]);
    hipEventElapsedTime(&b, c->ev[0], c->ev[3]);
    v[0]->stats.projection_ms = a;
    v[0]->stats.step_ms = b;
  }
  return EFL_OK;
}

// Rigid-body step after the fluid step: one launch for the context's list
// (forces of every stepped instance that uses it) and one per instance with
// its own dynamic list.
void integrate_bodies(efl_ctx* c, const IV& stepped, float dt) {
  auto fill = [&](efl::BodyStep& a, const IV& from) {
    a.ninst = 0;
    for (auto* i : from) {
      if (a.ninst >= efl::kMaxBatch) break;
      a.fts[a.ninst] = i->h.fts;
      a.dx[a.ninst] = i->h.dx;
      a.inv_dx[a.ninst] = 1.0f / i->h.dx;
      ++a.ninst;
    }
    efl_instance* f = from[0];
    a.dt = dt;
    a.gwx = f->h.gx;   // grid frame (y down) -> world frame (y up)
    a.gwy = -f->h.gy;
    a.half_w = 0.5f * (float)f->s.width;
    a.half_h = 0.5f * (float)f->s.height;
    a.rho_fluid = f->h.rho;
  };
  if (c->nbodies > 0 && c->nbodies == c->nobs) {
    IV from = filter(stepped, [](efl_instance* i) { return !i->own_obs; });
    if (!from.empty()) {
      efl::BodyStep a;
      std::memset(&a, 0, sizeof(a));
      a.obs = c->d_obs;
      a.bodies = c->d_bodies;
      a.state = (float4*)(c->d_bodies + EFL_MAX_SOLIDS);
      a.n = (int)c->nbodies;
      fill(a, from);
      efl::launch_integrate_bodies(c->stream, a);
    }
  }
  // instances with their own lists: up to kMaxBodyLists per launch
  efl::BodyLists L;
  std::memset(&L, 0, sizeof(L));
  L.dt = dt;
  for (auto* i : stepped) {
    if (!i->own_obs || !i->own_bodies || i->nobs_own == 0) continue;
    efl::BodyList& b = L.list[L.n++];
    b.obs = i->d_obs_own;
    b.bodies = i->d_bodies_own;
    b.state = (float4*)(i->d_bodies_own + EFL_MAX_SOLIDS);
    b.fts = i->h.fts;
    b.n = (int)i->nobs_own;
    b.dx = i->h.dx;
    b.gwx = i->h.gx;  // grid frame (y down) -> world frame (y up)
    b.gwy = -i->h.gy;
    b.half_w = 0.5f * (float)i->s.width;
    b.half_h = 0.5f * (float)i->s.height;
    b.rho_fluid = i->h.rho;
    if (L.n == efl::kMaxBodyLists) {
      efl::launch_integrate_own_bodies(c->stream, L);
      L.n = 0;
    }
  }
  if (L.n > 0) efl::launch_integrate_own_bodies(c->stream, L);
}

// ---- replayed steps -------------------------------------------------------------
void drop_graphs(efl_ctx* c) {
  for (auto& g : c->graphs)
    for (hipGraphExec_t e : {g.pre, g.iter[0], g.iter[1], g.post})
      if (e) hipGraphExecDestroy(e);
  c->graphs.clear();
}

template <class F>
hipGraphExec_t capture(efl_ctx* c, F&& enqueue) {
  hipGraph_t g = nullptr;
  if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed) != hipSuccess) return nullptr;
  enqueue();
  if (hipStreamEndCapture(c->stream, &g) != hipSuccess || !g) return nullptr;
  hipGraphExec_t e = nullptr;
  if (hipGraphInstantiate(&e, g, nullptr, nullptr, 0) != hipSuccess) e = nullptr;
  hipGraphDestroy(g);
  return e;
}

// The step as up to three replayed graphs: everything before the CG
// iterations (solids .. divergence, the fixed solves, the CG start), one CG
// iteration (replayed as often as the host loop asks), everything after the
// solve including the rigid bodies.  The launch arguments of a step depend
// only on the instance list, dt and what bumps gversion; the multigrid
// buffer-pair state the launches were recorded with is the state every step
// starts and ends in (checked at capture).  *used = false when the step must
// run eagerly (timing, diagnostics, multigrid to a residual target).
int run_step_graph(efl_ctx* c, const IV& v, float dt, bool* used) {
  *used = false;
  if (!c->graphs_on || c->timing) return EFL_OK;
  for (auto* i : v)
    if (i->diag || mg_needs_host(i) || is_learned(i)) return EFL_OK;
  uint32_t db = 0;
  std::memcpy(&db, &dt, 4);
  StepGraph* g = nullptr;
  for (auto& e : c->graphs)
    if (e.insts == v && e.dt_bits == db && e.version == c->gversion) g = &e;
  IV cg = filter(v, [](efl_instance* i) { return is_cg(i); });
  IV fixed = filter(v, [](efl_instance* i) { return !is_cg(i); });
  hipStream_t st = c->stream;
  if (!g) {
    if (!c->graphs.empty() && c->graphs[0].version != c->gversion) drop_graphs(c);
    std::vector<std::vector<uint8_t>> snap;
    for (auto* i : v) {
      snap.emplace_back(i->xsel, i->xsel + efl::kMaxLevels + 1);
      snap.back().insert(snap.back().end(), i->xzero, i->xzero + efl::kMaxLevels + 1);
    }
    StepGraph n;
    n.insts = v;
    n.dt_bits = db;
    n.version = c->gversion;
    n.pre = capture(c, [&] {
      phase_pre(st, v);
      proj_fixed(st, fixed);
      if (!cg.empty()) cg_begin(st, cg);
    });
    for (int k = 0; k < 2 && !cg.empty(); ++k) n.iter[k] = capture(c, [&] { cg_iteration(st, cg, c->cg_fuse); });
    n.post = capture(c, [&] {
      phase_velocity(st, v);
      phase_post(st, v);
      integrate_bodies(c, v, dt);
    });
    bool ok = n.pre && n.post && (cg.empty() || (n.iter[0] && n.iter[1]));
    for (auto* i : cg) ok = ok && i->rpar == 0;  // the odd iteration's buffers again
    for (size_t k = 0; k < v.size() && ok; ++k) {
      std::vector<uint8_t> now(v[k]->xsel, v[k]->xsel + efl::kMaxLevels + 1);
      now.insert(now.end(), v[k]->xzero, v[k]->xzero + efl::kMaxLevels + 1);
      ok = now == snap[k];
    }
    if (!ok) {  // run this configuration eagerly from now on
      for (hipGraphExec_t e : {n.pre, n.iter[0], n.iter[1], n.post})
        if (e) hipGraphExecDestroy(e);
      (void)hipGetLastError();
      c->graphs_on = false;
      for (size_t k = 0; k < v.size(); ++k) {  // back to the recorded start state
        std::copy(snap[k].begin(), snap[k].begin() + efl::kMaxLevels + 1, v[k]->xsel);
        std::copy(snap[k].begin() + efl::kMaxLevels + 1, snap[k].end(), v[k]->xzero);
      }
      return EFL_OK;
    }
    if (c->graphs.size() >= 8) {
      for (hipGraphExec_t e : {c->graphs[0].pre, c->graphs[0].iter[0], c->graphs[0].iter[1], c->graphs[0].post})
        if (e) hipGraphExecDestroy(e);
      c->graphs.erase(c->graphs.begin());
    }
    c->graphs.push_back(n);
    g = &c->graphs.back();
  }
  *used = true;
  proj_fixed_stats(v);
  HIPCHK(c, hipGraphLaunch(g->pre, st));
  if (!cg.empty()) {
    int rc = cg_run(c, cg, g->iter);  // the odd and even iterations in turn
    if (rc) return rc;
    if (!c->async_margin) {
      rc = resolve_sealed(c, cg);  // eager launches on this rare path
      if (rc) return rc;
    }
  }
  HIPCHK(c, hipGraphLaunch(g->post, st));
  step_done(v);
  return EFL_OK;
}

size_t al(size_t x) { return (x + 255) & ~(size_t)255; }

struct FieldDesc { void* p; size_t n; uint32_t w, h, c; };

FieldDesc field_desc(efl_instance* in, int f) {
  DevInst& h = in->h;
  uint32_t W = in->s.width, H = in->s.height;
  switch (f) {
    case EFL_FIELD_U0: return {h.u0, (size_t)(W + 1) * H, W + 1, H, 1};
    case EFL_FIELD_U1: return {h.u1, (size_t)(W + 1) * H, W + 1, H, 1};
    case EFL_FIELD_V0: return {h.v0, (size_t)W * (H + 1), W, H + 1, 1};
    case EFL_FIELD_V1: return {h.v1, (size_t)W * (H + 1), W, H + 1, 1};
    case EFL_FIELD_U_SOLID: return {h.us, (size_t)(W + 1) * H, W + 1, H, 1};
    case EFL_FIELD_V_SOLID: return {h.vs, (size_t)W * (H + 1), W, H + 1, 1};
    case EFL_FIELD_SOLID_ID: return {h.sid, (size_t)W * H, W, H, 1};
    case EFL_FIELD_U_VALID0: return {h.uval0, (size_t)(W + 1) * H, W + 1, H, 1};
    case EFL_FIELD_U_VALID1: return {h.uval1, (size_t)(W + 1) * H, W + 1, H, 1};
    case EFL_FIELD_V_VALID0: return {h.vval0, (size_t)W * (H + 1), W, H + 1, 1};
    case EFL_FIELD_V_VALID1: return {h.vval1, (size_t)W * (H + 1), W, H + 1, 1};
    case EFL_FIELD_DIV: return {h.div, (size_t)W * H, W, H, 1};
    case EFL_FIELD_P0: return {h.p0, (size_t)W * H, W, H, 1};
    case EFL_FIELD_P1: return {h.p1, (size_t)W * H, W, H, 1};
    case EFL_FIELD_LS_AIR0: return {h.ls0, (size_t)W * H, W, H, 1};
    case EFL_FIELD_LS_AIR1: return {h.ls1, (size_t)W * H, W, H, 1};
    case EFL_FIELD_GRAD_LS: return {h.grad, (size_t)W * H * 2, W, H, 2};
    case EFL_FIELD_LS_SOLID: return {h.lss, (size_t)W * H, W, H, 1};
    case EFL_FIELD_AREA_FRACTION: return {h.afs, (size_t)W * H * 4, W, H, 4};
    case EFL_FIELD_FIM_LABELS0: return {h.lab0, (size_t)W * H, W, H, 1};
    case EFL_FIELD_FIM_LABELS: return {h.lab, (size_t)W * H, W, H, 1};
    case EFL_FIELD_JFA_SEEDS0: return {h.seed0, (size_t)W * H * 2, W, H, 2};
    case EFL_FIELD_JFA_SEEDS1: return {h.seed1, (size_t)W * H * 2, W, H, 2};
    case EFL_FIELD_SOLID_BINS: return {h.bins_last, (size_t)4 * EFL_MAX_SOLIDS, EFL_MAX_SOLIDS, 4, 1};
    case EFL_FIELD_SOLID_FORCES: return {h.fts, (size_t)4 * EFL_MAX_SOLIDS, EFL_MAX_SOLIDS, 1, 4};
    case EFL_FIELD_MG_R0: return {h.lev[0].r, (size_t)W * H, W, H, 1};
    default: return {nullptr, 0, 0, 0, 0};
  }
}

bool valid_projection(const efl_projection& p) {
  return p.kind >= EFL_PROJECTION_JACOBI && p.kind <= EFL_PROJECTION_LEARNED && p.rel_tol >= 0.0f;
}
bool valid_reinit(const efl_reinit& r) {
  return r.kind == EFL_REINIT_JUMP_FLOODING || r.kind == EFL_REINIT_FAST_ITERATIVE;
}

}  // namespace

// ============================================================================
// C ABI
// ============================================================================
extern "C" {

int efl_ctx_create(int device, efl_ctx** out) {
  if (!out) return EFL_E_INVALID;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return EFL_E_HIP;
  if (device < 0 || device >= n) return EFL_E_INVALID;
  efl_ctx* c = new (std::nothrow) efl_ctx();
  if (!c) return EFL_E_OOM;
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return EFL_E_HIP;
  }
  if (hipMalloc(&c->d_obs, sizeof(efl_obstacle) * EFL_MAX_SOLIDS) != hipSuccess ||
      hipMalloc(&c->d_scal, sizeof(double) * 16 * EFL_MAX_BATCH) != hipSuccess ||
      hipHostMalloc((void**)&c->h_scal, sizeof(double) * 16 * EFL_MAX_BATCH, hipHostMallocDefault) != hipSuccess) {
    if (c->d_obs) hipFree(c->d_obs);
    if (c->d_scal) hipFree(c->d_scal);
    hipStreamDestroy(c->stream);
    delete c;
    return EFL_E_OOM;
  }
  // the no-alloc entry point's scratch on this device (its own copy per GPU)
  if (hipGetSymbolAddress((void**)&c->scratch_ptr, HIP_SYMBOL(scratch)) != hipSuccess || !c->scratch_ptr) {
    hipFree(c->d_obs);
    hipFree(c->d_scal);
    hipHostFree(c->h_scal);
    hipStreamDestroy(c->stream);
    delete c;
    return EFL_E_HIP;
  }
  for (auto& e : c->ev) hipEventCreate(&e);
  c->bottom_cells = efl::bottom_cells_for_device(device);
  if (const char* e = getenv("EFL_COARSE_FUSE")) c->coarse_fuse = atoi(e) == 1;
  if (const char* e = getenv("EFL_CG_FUSE")) c->cg_fuse = atoi(e) != 0;
  if (const char* e = getenv("EFL_UP_MERGE")) c->up_merge = atoi(e) != 0;
  if (c->coarse_fuse) c->bottom_cells = 1024;  // the fused coarse cycle holds its bottom in a 256-thread launch
  *out = c;
  return EFL_OK;
}

int efl_ctx_destroy(efl_ctx* c) {
  if (!c) return EFL_E_INVALID;
  hipSetDevice(c->device);
  hipStreamSynchronize(c->stream);
  std::vector<efl_instance*> left = c->insts;
  for (auto* i : left) efl_instance_destroy(i);
  drop_graphs(c);
  for (auto& e : c->ev)
    if (e) hipEventDestroy(e);
  if (c->d_obs) hipFree(c->d_obs);
  if (c->d_bodies) hipFree(c->d_bodies);
  if (c->d_scal) hipFree(c->d_scal);
  if (c->h_scal) hipHostFree(c->h_scal);
  hipStreamDestroy(c->stream);
  delete c;
  return EFL_OK;
}

const char* efl_last_error(const efl_ctx* c) { return c ? c->err.c_str() : "null context"; }

int efl_sync(efl_ctx* c) {
  if (!c) return EFL_E_INVALID;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (auto* i : c->insts) {
    int rc = resolve_stats(i, true);
    if (rc) return rc;
    uint32_t failed = 0;  // a group barrier of the fused coarse cycle gave up (its fields are not valid)
    HIPCHK(c, hipMemcpy(&failed, i->h.counter + 2, sizeof(failed), hipMemcpyDeviceToHost));
    if (failed) return fail(c, EFL_E_HIP, "a fused coarse-cycle barrier timed out on an instance");
  }
  return EFL_OK;
}

int efl_set_async(efl_ctx* c, uint32_t margin) {
  if (!c) return EFL_E_INVALID;
  if (margin > 1000) return fail(c, EFL_E_RANGE, "margin %u > 1000", margin);
  c->async_margin = margin;
  return EFL_OK;
}

int efl_host_waits(const efl_ctx* c, uint64_t* out) {
  if (!c || !out) return EFL_E_INVALID;
  *out = c->host_waits;
  return EFL_OK;
}

int efl_set_timing(efl_ctx* c, int enabled) {
  if (!c) return EFL_E_INVALID;
  c->timing = enabled != 0;
  return EFL_OK;
}

int efl_set_graphs(efl_ctx* c, int enabled) {
  if (!c) return EFL_E_INVALID;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  drop_graphs(c);
  c->graphs_on = enabled != 0;
  return EFL_OK;
}

int efl_set_obstacles(efl_ctx* c, const efl_obstacle* obs, uint32_t n) {
  if (!c || (n && !obs)) return EFL_E_INVALID;
  if (n > EFL_MAX_SOLIDS) return fail(c, EFL_E_RANGE, "%u obstacles > EFL_MAX_SOLIDS", n);
  HIPCHK(c, hipSetDevice(c->device));
  if (n) HIPCHK(c, hipMemcpyAsync(c->d_obs, obs, sizeof(efl_obstacle) * n, hipMemcpyHostToDevice, c->stream));
  if (n != c->nobs) c->gversion++;  // the body integrator's launch depends on the count
  c->nobs = n;
  c->obs_version++;
  return EFL_OK;
}

int efl_set_bodies(efl_ctx* c, const efl_body* bodies, uint32_t n) {
  if (!c || (n && !bodies)) return EFL_E_INVALID;
  if (n > EFL_MAX_SOLIDS) return fail(c, EFL_E_RANGE, "%u bodies > EFL_MAX_SOLIDS", n);
  if (n && n != c->nobs) return fail(c, EFL_E_INVALID, "%u bodies for %u obstacles", n, c->nobs);
  HIPCHK(c, hipSetDevice(c->device));
  // records followed by the integrator state (previous accelerations)
  if (n && !c->d_bodies) HIPCHK(c, hipMalloc(&c->d_bodies, (sizeof(efl_body) + 16) * EFL_MAX_SOLIDS));
  if (n) HIPCHK(c, hipMemcpyAsync(c->d_bodies, bodies, sizeof(efl_body) * n, hipMemcpyHostToDevice, c->stream));
  if (n) HIPCHK(c, hipMemsetAsync(c->d_bodies + EFL_MAX_SOLIDS, 0, 16 * EFL_MAX_SOLIDS, c->stream));
  c->nbodies = n;
  c->gversion++;
  return EFL_OK;
}

int efl_get_obstacles(efl_ctx* c, efl_obstacle* out, uint32_t n) {
  if (!c || (n && !out)) return EFL_E_INVALID;
  if (n > c->nobs) return fail(c, EFL_E_RANGE, "%u > %u obstacles", n, c->nobs);
  HIPCHK(c, hipSetDevice(c->device));
  if (n) HIPCHK(c, hipMemcpyAsync(out, c->d_obs, sizeof(efl_obstacle) * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return EFL_OK;
}

int efl_set_instance_obstacles(efl_instance* in, const efl_obstacle* obs, const efl_body* bodies, uint32_t n) {
  if (!in || (n && !obs)) return EFL_E_INVALID;
  efl_ctx* c = in->ctx;
  if (n > EFL_MAX_SOLIDS) return fail(c, EFL_E_RANGE, "%u obstacles > EFL_MAX_SOLIDS", n);
  HIPCHK(c, hipSetDevice(c->device));
  if (!obs) {
    if (in->own_obs) c->gversion++;
    in->own_obs = false;
    in->own_bodies = false;
    in->nobs_own = 0;
    in->dirty = true;
    return EFL_OK;
  }
  if (!in->d_obs_own) HIPCHK(c, hipMalloc(&in->d_obs_own, sizeof(efl_obstacle) * EFL_MAX_SOLIDS));
  if (bodies && !in->d_bodies_own) HIPCHK(c, hipMalloc(&in->d_bodies_own, (sizeof(efl_body) + 16) * EFL_MAX_SOLIDS));
  if (n) HIPCHK(c, hipMemcpyAsync(in->d_obs_own, obs, sizeof(efl_obstacle) * n, hipMemcpyHostToDevice, c->stream));
  if (n && bodies) {
    HIPCHK(c, hipMemcpyAsync(in->d_bodies_own, bodies, sizeof(efl_body) * n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemsetAsync(in->d_bodies_own + EFL_MAX_SOLIDS, 0, 16 * EFL_MAX_SOLIDS, c->stream));
  }
  if (!in->own_obs || in->own_bodies != (bodies != nullptr) || in->nobs_own != n) c->gversion++;
  in->own_obs = true;
  in->own_bodies = bodies != nullptr;
  in->nobs_own = n;
  in->dirty = true;
  return EFL_OK;
}

int efl_get_instance_obstacles(efl_instance* in, efl_obstacle* out, uint32_t n) {
  if (!in || (n && !out)) return EFL_E_INVALID;
  efl_ctx* c = in->ctx;
  if (!in->own_obs) return efl_get_obstacles(c, out, n);
  if (n > in->nobs_own) return fail(c, EFL_E_RANGE, "%u > %u obstacles", n, in->nobs_own);
  HIPCHK(c, hipSetDevice(c->device));
  if (n) HIPCHK(c, hipMemcpyAsync(out, in->d_obs_own, sizeof(efl_obstacle) * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return EFL_OK;
}

int efl_instance_create(efl_ctx* c, const efl_settings* s, efl_instance** out) {
  if (!c || !s || !out) return EFL_E_INVALID;
  *out = nullptr;
  if (s->width < 8 || s->height < 8 || s->width > 16384 || s->height > 16384)
    return fail(c, EFL_E_INVALID, "grid size %ux%u outside [8, 16384]", s->width, s->height);
  if (!(s->length_unit > 0.0f)) return fail(c, EFL_E_INVALID, "length_unit must be positive");  // euler_fluid.rs:49-51
  if (!valid_projection(s->projection) || !valid_reinit(s->reinit))
    return fail(c, EFL_E_INVALID, "invalid projection or reinitialisation method");
  HIPCHK(c, hipSetDevice(c->device));
  efl_instance* in = new (std::nothrow) efl_instance();
  if (!in) return EFL_E_OOM;
  in->ctx = c;
  in->s = *s;
  in->bottom_cells = c->bottom_cells;
  size_t W = s->width, H = s->height, n = W * H, nu = (W + 1) * H, nv = W * (H + 1);
  // multigrid levels, multi_grid.rs:95-97
  int mn = (int)std::min(W, H);
  int nl = std::max(1, (int)std::log2((float)mn) - 2);
  nl = std::min(nl, efl::kMaxLevels);
  std::vector<std::pair<int, int>> lv(nl);
  lv[0] = {(int)W, (int)H};
  for (int l = 1; l < nl; ++l) lv[l] = {lv[l - 1].first / 2, lv[l - 1].second / 2};
  int nblk = (int)(((W + efl::tile_w() - 1) / efl::tile_w()) * ((H + efl::tile_h() - 1) / efl::tile_h()));
  // byte layout
  size_t off = 0;
  auto take = [&](size_t bytes) { size_t o = off; off += al(bytes); return o; };
  size_t o_u0 = take(nu * 4), o_u1 = take(nu * 4), o_us = take(nu * 4);
  size_t o_v0 = take(nv * 4), o_v1 = take(nv * 4), o_vs = take(nv * 4);
  size_t o_uv0 = take(nu * 4), o_uv1 = take(nu * 4), o_vv0 = take(nv * 4), o_vv1 = take(nv * 4);
  size_t o_sid = take(n * 4), o_div = take(n * 4), o_p0 = take(n * 4), o_p1 = take(n * 4);
  size_t o_ls0 = take(n * 4), o_ls1 = take(n * 4), o_lss = take(n * 4), o_phit = take(n * 4);
  size_t o_lab0 = take(n * 4), o_lab = take(n * 4), o_labt = take(n * 4);
  size_t o_grad = take(n * 8), o_s0 = take(n * 8), o_s1 = take(n * 8), o_afs = take(n * 16);
  size_t o_bins = take(4 * EFL_MAX_SOLIDS * 4), o_binsl = take(4 * EFL_MAX_SOLIDS * 4);
  size_t o_fts = take(EFL_MAX_SOLIDS * 16);
  size_t o_r0 = take(n * 4), o_xs0 = take(n * 4);
  std::vector<size_t> o_lx(nl), o_lb(nl), o_lls(nl), o_laf(nl), o_lr(nl), o_lxs(nl), o_lcf(nl), o_lfl(nl);
  for (int l = 0; l < nl; ++l) {
    size_t m = (size_t)lv[l].first * lv[l].second;
    if (l > 0) {
      o_lx[l] = take(m * 4); o_lb[l] = take(m * 4); o_lls[l] = take(m * 4); o_laf[l] = take(m * 16);
      o_lr[l] = take(m * 4); o_lxs[l] = take(m * 4);
    }
    o_lcf[l] = take(m * 16); o_lfl[l] = take(m * 4);
  }
  size_t o_pr2 = take(n * 4);
  size_t o_pr = take(n * 4), o_pz = take(n * 4), o_pp = take(n * 4), o_pp1 = take(n * 4), o_pq = take(n * 4);
  size_t o_red = take((size_t)4 * nblk * 8), o_scal = take(efl::kScalars * 8), o_cnt = take(16), o_diag = take(16);
  size_t o_fdp = take((size_t)3 * (W / 16) * (H / 16) * 4), o_fdr = take(16);
  size_t o_src0 = take(sizeof(efl_source) * EFL_MAX_FLUID_SOURCES), o_src = take(sizeof(efl_source) * EFL_MAX_FLUID_SOURCES);
  size_t o_desc = take(sizeof(DevInst));
  in->bytes = off;
  if (hipMalloc(&in->mem, off) != hipSuccess) {
    delete in;
    return fail(c, EFL_E_OOM, "hipMalloc(%zu) failed", off);
  }
  if (hipMemsetAsync(in->mem, 0, off, c->stream) != hipSuccess) {
    hipFree(in->mem);
    delete in;
    return fail(c, EFL_E_HIP, "hipMemsetAsync failed");
  }
  char* base = (char*)in->mem;
  DevInst& h = in->h;
  std::memset(&h, 0, sizeof(h));
  h.W = (int)W;
  h.H = (int)H;
  h.dx = 1.0f / s->length_unit;
  h.dt = 0.0f;
  h.rho = s->rho;
  h.gx = s->gravity[0];
  h.gy = s->gravity[1];
  bool zero = true;
  for (int k = 0; k < 16; ++k) zero = zero && s->transform[k] == 0.0f;
  for (int k = 0; k < 16; ++k) h.T[k] = zero ? ((k % 5) == 0 ? 1.0f : 0.0f) : s->transform[k];
  h.sizex = (float)W;
  h.sizey = (float)H;
  h.u0 = (float*)(base + o_u0); h.u1 = (float*)(base + o_u1); h.us = (float*)(base + o_us);
  h.v0 = (float*)(base + o_v0); h.v1 = (float*)(base + o_v1); h.vs = (float*)(base + o_vs);
  h.uval0 = (int32_t*)(base + o_uv0); h.uval1 = (int32_t*)(base + o_uv1);
  h.vval0 = (int32_t*)(base + o_vv0); h.vval1 = (int32_t*)(base + o_vv1);
  h.sid = (int32_t*)(base + o_sid); h.div = (float*)(base + o_div); h.p0 = (float*)(base + o_p0); h.p1 = (float*)(base + o_p1);
  h.ls0 = (float*)(base + o_ls0); h.ls1 = (float*)(base + o_ls1); h.lss = (float*)(base + o_lss);
  h.phi_tmp = (float*)(base + o_phit);
  h.lab0 = (uint32_t*)(base + o_lab0); h.lab = (uint32_t*)(base + o_lab); h.lab_tmp = (uint32_t*)(base + o_labt);
  h.grad = (float2*)(base + o_grad); h.seed0 = (float2*)(base + o_s0); h.seed1 = (float2*)(base + o_s1);
  h.afs = (float4*)(base + o_afs);
  h.bins = (int32_t*)(base + o_bins); h.bins_last = (int32_t*)(base + o_binsl);
  h.fts = (float4*)(base + o_fts);
  h.nlev = nl;
  h.lev[0] = {(int)W, (int)H, h.p0, h.div, h.ls0, h.afs, (float*)(base + o_r0), (float*)(base + o_xs0),
              (float4*)(base + o_lcf[0]), (int32_t*)(base + o_lfl[0])};
  for (int l = 1; l < nl; ++l)
    h.lev[l] = {lv[l].first, lv[l].second, (float*)(base + o_lx[l]), (float*)(base + o_lb[l]), (float*)(base + o_lls[l]),
                (float4*)(base + o_laf[l]), (float*)(base + o_lr[l]), (float*)(base + o_lxs[l]),
                (float4*)(base + o_lcf[l]), (int32_t*)(base + o_lfl[l])};
  h.pcg_r = (float*)(base + o_pr); h.pcg_z = (float*)(base + o_pz); h.pcg_p[0] = (float*)(base + o_pp); h.pcg_p[1] = (float*)(base + o_pp1); h.pcg_q = (float*)(base + o_pq);
  h.pcg_r2 = (float*)(base + o_pr2);
  h.lev_pcg0 = h.lev[0];  // z over r, same level set, area fractions and stencil
  h.lev_pcg0.x = h.pcg_z;
  h.lev_pcg0.b = h.pcg_r;
  h.red = (double*)(base + o_red);
  h.red_stride = nblk;
  {  // CG tiles (64 x 16), multigrid tiles and network blocks all cover ~1024 cells.
     // EFL_INLINE_FOLD_MAX lowers the threshold (0: every instance takes k_prefold);
     // the scalars are the same either way.
    const long cg_tiles = (long)((W + efl::tile_w() - 1) / efl::tile_w()) * ((H + efl::cg_tile_h() - 1) / efl::cg_tile_h());
    const long mg_tiles = (long)((W + efl::mg_tile_w() - 1) / efl::mg_tile_w()) * ((H + efl::mg_tile_h() - 1) / efl::mg_tile_h());
    const long nn_blocks = ((long)W * H + 1023) / 1024;
    long limit = efl::kInlineFold;
    if (const char* e = getenv("EFL_INLINE_FOLD_MAX")) limit = std::min<long>(limit, std::max(0L, atol(e)));
    h.prefold = std::max(cg_tiles, std::max(mg_tiles, nn_blocks)) > limit;
    h.cg_tiles = (int)cg_tiles;
  }
  h.scal = (double*)(base + o_scal);
  h.counter = (uint32_t*)(base + o_cnt);
  h.diag = (float*)(base + o_diag);
  h.fdp = (float*)(base + o_fdp);
  h.fdr = (float*)(base + o_fdr);
  h.src_init = (efl_source*)(base + o_src0);
  h.src = (efl_source*)(base + o_src);
  h.obs = c->d_obs;
  h.nobs = c->nobs;
  in->obs_version = c->obs_version;
  in->d = (DevInst*)(base + o_desc);
  in->jfa_max_power = (int)std::floor(std::log2((float)std::max(W, H)) - 1.0f);  // jump_flooding.rs:202
  in->dirty = true;
  int rc = upload(in);
  if (rc) {
    hipFree(in->mem);
    delete in;
    return rc;
  }
  c->insts.push_back(in);
  *out = in;
  return EFL_OK;
}

int efl_instance_destroy(efl_instance* in) {
  if (!in) return EFL_E_INVALID;
  efl_ctx* c = in->ctx;
  hipSetDevice(c->device);
  hipStreamSynchronize(c->stream);
  if (in->d_forces) hipFree(in->d_forces);
  if (in->d_obs_own) hipFree(in->d_obs_own);
  if (in->d_bodies_own) hipFree(in->d_bodies_own);
  if (in->mem) hipFree(in->mem);
  if (in->h_stat) hipHostFree(in->h_stat);
  for (hipEvent_t e : in->stat_ev)
    if (e) hipEventDestroy(e);
  nn_free_activations(in->nn_act);
  c->insts.erase(std::remove(c->insts.begin(), c->insts.end(), in), c->insts.end());
  drop_graphs(c);  // their launches name this instance
  c->gversion++;
  delete in;
  return EFL_OK;
}

int efl_set_projection(efl_instance* in, const efl_projection* p) {
  if (!in || !p) return EFL_E_INVALID;
  if (!valid_projection(*p)) return fail(in->ctx, EFL_E_INVALID, "invalid projection");
  in->s.projection = *p;
  in->ctx->gversion++;
  return EFL_OK;
}

int efl_set_reinit(efl_instance* in, const efl_reinit* r) {
  if (!in || !r) return EFL_E_INVALID;
  if (!valid_reinit(*r)) return fail(in->ctx, EFL_E_INVALID, "invalid reinitialisation method");
  in->s.reinit = *r;
  in->ctx->gversion++;
  return EFL_OK;
}

int efl_set_uniforms(efl_instance* in, float rho, const float g[2], const float T[16]) {
  if (!in) return EFL_E_INVALID;
  in->h.rho = rho;
  in->s.rho = rho;
  if (g) { in->h.gx = g[0]; in->h.gy = g[1]; in->s.gravity[0] = g[0]; in->s.gravity[1] = g[1]; }
  if (T) for (int k = 0; k < 16; ++k) in->h.T[k] = T[k];
  in->dirty = true;
  in->ctx->gversion++;  // the body integrator takes gravity and rho as arguments
  return EFL_OK;
}

int efl_set_sources(efl_instance* in, const efl_source* src, uint32_t n, int oneshot) {
  if (!in || (n && !src)) return EFL_E_INVALID;
  efl_ctx* c = in->ctx;
  if (n > EFL_MAX_FLUID_SOURCES)  // fluid_source_uniform.rs:92-97 warns and drops the rest
    return fail(c, EFL_E_RANGE, "%u sources > EFL_MAX_FLUID_SOURCES", n);
  HIPCHK(c, hipSetDevice(c->device));
  efl_source* dst = oneshot ? in->h.src_init : in->h.src;
  if (n) HIPCHK(c, hipMemcpyAsync(dst, src, sizeof(efl_source) * n, hipMemcpyHostToDevice, c->stream));
  if (oneshot) in->h.nsrc_init = n;
  else in->h.nsrc = n;
  in->dirty = true;
  return EFL_OK;
}

int efl_set_forces(efl_instance* in, const efl_force* f, uint32_t n) {
  if (!in || (n && !f)) return EFL_E_INVALID;
  efl_ctx* c = in->ctx;
  HIPCHK(c, hipSetDevice(c->device));
  if (n > in->forces_cap) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (in->d_forces) hipFree(in->d_forces);
    in->d_forces = nullptr;
    uint32_t cap = std::max<uint32_t>(64, n);
    if (hipMalloc(&in->d_forces, sizeof(efl_force) * cap) != hipSuccess) {
      in->forces_cap = 0;
      in->h.nforces = 0;
      in->dirty = true;
      return fail(c, EFL_E_OOM, "forces buffer");
    }
    in->forces_cap = cap;
  }
  if (n) HIPCHK(c, hipMemcpyAsync(in->d_forces, f, sizeof(efl_force) * n, hipMemcpyHostToDevice, c->stream));
  in->h.forces = in->d_forces;
  in->h.nforces = n;
  in->dirty = true;
  return EFL_OK;
}

int efl_init_levelset(efl_instance* in, const float* ls) {
  if (!in) return EFL_E_INVALID;
  if (!ls) in->custom_ls.clear();
  else in->custom_ls.assign(ls, ls + (size_t)in->s.width * in->s.height);
  return EFL_OK;
}

int efl_reset(efl_instance* in) {
  if (!in) return EFL_E_INVALID;
  in->initialized = false;
  return EFL_OK;
}

int efl_initialize(efl_instance* in) {
  if (!in) return EFL_E_INVALID;
  efl_ctx* c = in->ctx;
  HIPCHK(c, hipSetDevice(c->device));
  int rc = upload(in);
  if (rc) return rc;
  return run_init(c, IV{in});
}

int efl_set_diagnostics(efl_instance* in, int enabled) {
  if (!in) return EFL_E_INVALID;
  in->diag = enabled != 0;
  return EFL_OK;
}

int efl_step(efl_ctx* c, efl_instance* const* insts, uint32_t n, float dt) {
  if (!c || (n && !insts)) return EFL_E_INVALID;
  if (n > EFL_MAX_BATCH) return fail(c, EFL_E_RANGE, "%u instances > EFL_MAX_BATCH", n);
  if (!(dt > 0.0f)) return fail(c, EFL_E_INVALID, "dt must be positive");
  HIPCHK(c, hipSetDevice(c->device));
  IV init, step;
  for (uint32_t k = 0; k < n; ++k) {
    efl_instance* i = insts[k];
    if (!i || i->ctx != c) return fail(c, EFL_E_INVALID, "instance %u is null or belongs to another context", k);
    if (std::find(init.begin(), init.end(), i) != init.end() || std::find(step.begin(), step.end(), i) != step.end())
      return fail(c, EFL_E_INVALID, "instance %u listed twice", k);
    if (i->h.dt != dt) {
      i->h.dt = dt;
      i->dirty = true;
    }
    int rc = upload(i);
    if (rc) return rc;
    (i->initialized ? step : init).push_back(i);
  }
  if (!init.empty()) {
    int rc = run_init(c, init);
    if (rc) return rc;
  }
  if (!step.empty()) {
    bool graphed = false;
    int rc = run_step_graph(c, step, dt, &graphed);
    if (rc) return rc;
    if (!graphed) {
      rc = run_step(c, step);
      if (rc) return rc;
      integrate_bodies(c, step, dt);
    }
    for (auto* i : step) {  // forces were cleared after use
      rc = upload(i);
      if (rc) return rc;
    }
  }
  return launch_status(c, "kernel launch");
}

int efl_run_stage(efl_instance* in, int stage, float dt) {
  if (!in) return EFL_E_INVALID;
  efl_ctx* c = in->ctx;
  HIPCHK(c, hipSetDevice(c->device));
  if (in->h.dt != dt) {
    in->h.dt = dt;
    in->dirty = true;
  }
  int rc = upload(in);
  if (rc) return rc;
  hipStream_t st = c->stream;
  IV v{in};
  Batch center = make_batch(v, EX_CENTER), bx = make_batch(v, EX_XMAJOR), by = make_batch(v, EX_YMAJOR);
  switch (stage) {
    case EFL_STAGE_MG_BOTTOM:
      if (bottom_level(in) < 0) return fail(c, EFL_E_STATE, "no bottom level at this size");
      mg_setup(st, v, true);
      v_cycle(st, v, bottom_level(in), 0, true);
      break;
    case EFL_STAGE_NN_PRECONDITION:
      if (!in->nn || !is_learned(in)) return fail(c, EFL_E_STATE, "NN_PRECONDITION needs efl_set_learned and a learned projection");
      nn_precondition(st, in, true);
      break;
    case EFL_STAGE_INIT_GRID_CENTER: efl::launch_init_center(st, center); break;
    case EFL_STAGE_INIT_GRID_EDGE: efl::launch_init_edge(st, bx); break;
    case EFL_STAGE_SOURCE_ONESHOT: efl::launch_fluid_source(st, center, 1); break;
    case EFL_STAGE_UPDATE_SOLID: efl::launch_update_solid(st, center); break;
    case EFL_STAGE_AREA_FRACTION: efl::launch_area_fraction(st, center); break;
    case EFL_STAGE_ADVECT_VELOCITY: efl::launch_advect_u(st, bx); efl::launch_advect_v(st, by); break;
    case EFL_STAGE_APPLY_FORCES: efl::launch_forces_u(st, bx); efl::launch_forces_v(st, by); break;
    case EFL_STAGE_DIVERGENCE: efl::launch_divergence(st, center); break;
    case EFL_STAGE_PROJECTION: rc = stage_projection(c, v); break;
    case EFL_STAGE_SOLVE_VELOCITY: efl::launch_solve_u(st, bx); efl::launch_solve_v(st, by); break;
    case EFL_STAGE_EXTRAPOLATE_VELOCITY: stage_extrapolate(st, v); break;
    case EFL_STAGE_ADVECT_LEVELSET: efl::launch_advect_levelset(st, center); break;
    case EFL_STAGE_SOURCE: efl::launch_fluid_source(st, center, 0); break;
    case EFL_STAGE_REINITIALIZE: stage_reinit(st, v); break;
    case EFL_STAGE_LEVELSET_GRADIENT: efl::launch_ls_gradient(st, center); break;
    case EFL_STAGE_FLUID_TO_SOLID: stage_fluid_to_solid(st, v); break;
    case EFL_STAGE_GS_SWEEP: {  // one red + black sweep, weight 1.9, gauss_seidel.rs:142-147
      mg_setup(st, v, false);
      efl::launch_mg_down(st, mg_batch(v, EX_CENTER, 0, false, 0), 0, 1, 1, 1.9f, 0, 0, mg_th_of(in, 0, 0));
      mg_flip(v, 0);
      mg_settle(st, v, 0);
      break;
    }
    case EFL_STAGE_MG_SMOOTH: {  // two red-black sweeps at weight 1 on level 0 in one tiled launch
      mg_setup(st, v, false);
      efl::launch_mg_down(st, mg_batch(v, EX_LEVEL, 0, false, 0), 0, 2, 1, 1.0f, 0, 0, mg_th_of(in, 0, 0));
      mg_flip(v, 0);
      mg_settle(st, v, 0);
      break;
    }
    case EFL_STAGE_GS_SWEEP_SPLIT: {  // one sweep as two kernels (one per colour), as the reference dispatches it
      efl::launch_gs_color(st, center, 1, 1.9f, 0);
      efl::launch_gs_color(st, center, 0, 1.9f, 0);
      break;
    }
    case EFL_STAGE_MG_VCYCLE:
      mg_setup(st, v, true);
      v_cycle(st, v, 0, 0, false);
      break;
    case EFL_STAGE_MG_RESIDUAL: efl::launch_residual(st, make_batch(v, EX_LEVEL, 0), 0); break;
    case EFL_STAGE_CG_SMOOTH: return fail(c, EFL_E_INVALID, "CG_SMOOTH is a timing stage (efl_time_stage)");
    default: return fail(c, EFL_E_INVALID, "unknown stage %d", stage);
  }
  if (rc) return rc;
  return launch_status(c, "kernel launch");
}

namespace {
// efl_time_stage(CG_SMOOTH): k_mg_down<2,5,true> on level 0 as an MGPCG
// iteration launches it.  The solve starts from x = 0 (so the initial test
// does not stop it), its first preconditioner leaves the r . z and z . A z
// partials; the launch then runs as a later iteration (the count set to 1,
// so beta is formed) with an odd iteration's buffers: p_prev = 0 in
// pcg_p[0], the previous r . z (slot 1) set to 1e300 and alpha to 1, so
// beta ~ 0 and p . A p ~ z . A z > 0; p = z + beta p_prev, q = A p, r read
// from pcg_r and written to pcg_r2, x += alpha p, z known to be zero,
// residual and restriction of b, reps times between two events (the
// scalars it reads do not change).  The pressure and the buffer-pair flags
// are put back afterwards.
int time_cg_smooth(efl_instance* in, float dt, uint32_t reps, float* ms) {
  efl_ctx* c = in->ctx;
  if (!is_cg(in) || !fused_axpy(in))
    return fail(c, EFL_E_STATE, "CG_SMOOTH needs an MGPCG instance whose level-0 launch carries the CG update");
  if (in->h.dt != dt) {
    in->h.dt = dt;
    in->dirty = true;
  }
  int rc = upload(in);
  if (rc) return rc;
  hipStream_t st = c->stream;
  IV v{in};
  const size_t bytes = (size_t)in->s.width * in->s.height * sizeof(float);
  float* keep = nullptr;
  HIPCHK(c, hipMalloc(&keep, bytes));
  uint8_t xsel[efl::kMaxLevels + 1], xzero[efl::kMaxLevels + 1];
  std::memcpy(xsel, in->xsel, sizeof(xsel));
  std::memcpy(xzero, in->xzero, sizeof(xzero));
  auto done = [&](int code) {
    (void)hipMemcpyAsync(in->h.p0, keep, bytes, hipMemcpyDeviceToDevice, st);
    (void)hipStreamSynchronize(st);
    (void)hipFree(keep);
    std::memcpy(in->xsel, xsel, sizeof(xsel));
    std::memcpy(in->xzero, xzero, sizeof(xzero));
    in->rpar = in->fax = 0;
    return code;
  };
  if (hipMemcpyAsync(keep, in->h.p0, bytes, hipMemcpyDeviceToDevice, st) != hipSuccess ||
      hipMemsetAsync(in->h.p0, 0, bytes, st) != hipSuccess)
    return done(fail(c, EFL_E_HIP, "CG_SMOOTH: saving the pressure"));
  cg_begin(st, v);
  const double one = 1.0, big = 1e300;
  if (hipMemsetAsync(in->h.pcg_p[0], 0, bytes, st) != hipSuccess ||
      hipMemcpyAsync(in->h.scal + efl::SC_ITER, &one, sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(in->h.scal + efl::SC_RZ0 + 1, &big, sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(in->h.scal + efl::SC_AL0 + 1, &one, sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return done(fail(c, EFL_E_HIP, "CG_SMOOTH: setting up the iteration"));
  in->xsel[0] = 0;  // the preconditioner's z starts at zero: not read
  in->xzero[0] = 1;
  in->rpar = 0;  // r read from pcg_r, written to pcg_r2 (an odd iteration's buffers)
  in->fax = 1;
  const Batch b = mg_batch(v, EX_LEVEL, 0, false, 1);  // SEL_AXPY: the update rides on this launch
  efl::launch_mg_down(st, b, 1, 2, 1, 1.0f, 1, 1);       // warm
  if (hipEventRecord(c->ev[0], st) != hipSuccess) return done(fail(c, EFL_E_HIP, "event"));
  for (uint32_t k = 0; k < reps; ++k) efl::launch_mg_down(st, b, 1, 2, 1, 1.0f, 1, 1);
  if (hipEventRecord(c->ev[1], st) != hipSuccess || hipEventSynchronize(c->ev[1]) != hipSuccess)
    return done(fail(c, EFL_E_HIP, "event"));
  float t = 0.0f;
  if (hipEventElapsedTime(&t, c->ev[0], c->ev[1]) != hipSuccess) return done(fail(c, EFL_E_HIP, "event time"));
  *ms = t / (float)reps;
  rc = launch_status(c, "CG_SMOOTH");
  return done(rc);
}
}  // namespace

int efl_time_stage(efl_instance* in, int stage, float dt, uint32_t reps, float* ms) {
  if (!in || !ms || reps == 0) return EFL_E_INVALID;
  efl_ctx* c = in->ctx;
  HIPCHK(c, hipSetDevice(c->device));
  if (stage == EFL_STAGE_CG_SMOOTH) return time_cg_smooth(in, dt, reps, ms);
  int rc = efl_run_stage(in, stage, dt);  // warm (and validate the stage id)
  if (rc) return rc;
  if (stage == EFL_STAGE_MG_BOTTOM) {  // the bottom launch alone, reps times back to back
    hipStream_t st = c->stream;
    IV v{in};
    HIPCHK(c, hipEventRecord(c->ev[0], st));
    for (uint32_t k = 0; k < reps; ++k) v_cycle(st, v, bottom_level(in), 0, true);
    HIPCHK(c, hipEventRecord(c->ev[1], st));
    HIPCHK(c, hipEventSynchronize(c->ev[1]));
    float t = 0.0f;
    HIPCHK(c, hipEventElapsedTime(&t, c->ev[0], c->ev[1]));
    *ms = t / (float)reps;
    return EFL_OK;
  }
  if (stage == EFL_STAGE_GS_SWEEP || stage == EFL_STAGE_MG_SMOOTH) {
    // the tiled smoother alone: reps launches back to back (ping-ponging the
    // x pair), without the stencil setup and the final copy-back
    hipStream_t st = c->stream;
    IV v{in};
    const bool gs = stage == EFL_STAGE_GS_SWEEP;
    Batch b = mg_batch(v, gs ? EX_CENTER : EX_LEVEL, 0, false, 0);
    HIPCHK(c, hipEventRecord(c->ev[0], st));
    for (uint32_t k = 0; k < reps; ++k) {
      b.sel[0] = in->xsel[0];
      efl::launch_mg_down(st, b, 0, gs ? 1 : 2, 1, gs ? 1.9f : 1.0f, 0, 0, mg_th_of(in, 0, 0));
      mg_flip(v, 0);
    }
    HIPCHK(c, hipEventRecord(c->ev[1], st));
    mg_settle(st, v, 0);
    HIPCHK(c, hipEventSynchronize(c->ev[1]));
    float t = 0.0f;
    HIPCHK(c, hipEventElapsedTime(&t, c->ev[0], c->ev[1]));
    *ms = t / (float)reps;
    HIPCHK(c, hipStreamSynchronize(st));
    return EFL_OK;
  }
  HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
  for (uint32_t k = 0; k < reps; ++k) {
    rc = efl_run_stage(in, stage, dt);
    if (rc) return rc;
  }
  HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
  HIPCHK(c, hipEventSynchronize(c->ev[1]));
  float t = 0.0f;
  HIPCHK(c, hipEventElapsedTime(&t, c->ev[0], c->ev[1]));
  *ms = t / (float)reps;
  return EFL_OK;
}

int efl_field_info(const efl_instance* in, int field, size_t* count, uint32_t* w, uint32_t* h, uint32_t* ch) {
  if (!in) return EFL_E_INVALID;
  FieldDesc d = field_desc(const_cast<efl_instance*>(in), field);
  if (!d.p) return EFL_E_INVALID;
  if (count) *count = d.n;
  if (w) *w = d.w;
  if (h) *h = d.h;
  if (ch) *ch = d.c;
  return EFL_OK;
}

int efl_get_field(efl_instance* in, int field, void* host, size_t count) {
  if (!in || !host) return EFL_E_INVALID;
  efl_ctx* c = in->ctx;
  FieldDesc d = field_desc(in, field);
  if (!d.p) return fail(c, EFL_E_INVALID, "unknown field %d", field);
  if (count != d.n) return fail(c, EFL_E_INVALID, "field %d holds %zu elements, not %zu", field, d.n, count);
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMemcpyAsync(host, d.p, d.n * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return EFL_OK;
}

int efl_set_field(efl_instance* in, int field, const void* host, size_t count) {
  if (!in || !host) return EFL_E_INVALID;
  efl_ctx* c = in->ctx;
  FieldDesc d = field_desc(in, field);
  if (!d.p) return fail(c, EFL_E_INVALID, "unknown field %d", field);
  if (count != d.n) return fail(c, EFL_E_INVALID, "field %d holds %zu elements, not %zu", field, d.n, count);
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMemcpyAsync(d.p, host, d.n * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return EFL_OK;
}

int efl_field_device_ptr(efl_instance* in, int field, void** dptr) {
  if (!in || !dptr) return EFL_E_INVALID;
  FieldDesc d = field_desc(in, field);
  if (!d.p) return EFL_E_INVALID;
  *dptr = d.p;
  return EFL_OK;
}

int efl_get_solid_forces(efl_instance* in, efl_solid_force* out, uint32_t n) {
  if (!in || !out) return EFL_E_INVALID;
  if (n > EFL_MAX_SOLIDS) return fail(in->ctx, EFL_E_RANGE, "n > EFL_MAX_SOLIDS");
  efl_ctx* c = in->ctx;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMemcpyAsync(out, in->h.fts, sizeof(efl_solid_force) * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return EFL_OK;
}

int efl_get_fluid_diagnostics(efl_instance* in, efl_fluid_diagnostics* out) {
  if (!in || !out) return EFL_E_INVALID;
  efl_ctx* c = in->ctx;
  HIPCHK(c, hipSetDevice(c->device));
  int rc = upload(in);
  if (rc) return rc;
  efl::FluidDiagArgs da{in->h.ls0, in->h.u0, in->h.v0, in->h.W, in->h.H, in->h.fdp, in->h.fdr};
  efl::launch_fluid_diagnostics(c->stream, da);
  HIPCHK(c, hipGetLastError());
  float r[4];
  HIPCHK(c, hipMemcpyAsync(r, in->h.fdr, sizeof(r), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  out->volume = r[0];
  out->min_velocity = r[1];
  out->max_velocity = r[2];
  out->reserved = 0.0f;
  return EFL_OK;
}

int efl_fluid_diagnostics_noalloc(efl_ctx* c, const float* ls, const float* u, const float* v, uint32_t width,
                                  uint32_t height, float* out) {
  if (!c || !ls || !u || !v || !out) return EFL_E_INVALID;
  if (width == 0 || height == 0 || width > 16384 || height > 16384)
    return fail(c, EFL_E_INVALID, "grid size %ux%u outside [1, 16384]", width, height);
  const size_t blocks = (size_t)(width / 16) * (height / 16);
  if (blocks > (size_t)EFL_NOALLOC_MAX_BLOCKS)
    return fail(c, EFL_E_RANGE, "%ux%u grid: %zu 16x16 blocks > EFL_NOALLOC_MAX_BLOCKS (%d)", width, height, blocks,
                EFL_NOALLOC_MAX_BLOCKS);
  HIPCHK(c, hipSetDevice(c->device));
  efl::FluidDiagArgs da{ls, u, v, (int)width, (int)height, c->scratch_ptr, out};
  efl::launch_fluid_diagnostics(c->stream, da);
  return launch_status(c, "fluid diagnostics");
}

int efl_measure(efl_instance* in, float dt, float* rel, float* mdiv) {
  if (!in) return EFL_E_INVALID;
  efl_ctx* c = in->ctx;
  HIPCHK(c, hipSetDevice(c->device));
  if (in->h.dt != dt) {
    in->h.dt = dt;
    in->dirty = true;
  }
  int rc = upload(in);
  if (rc) return rc;
  IV v{in};
  efl::launch_residual_norm(c->stream, cg_batch(v));
  rc = read_scalars(c, v);
  if (rc) return rc;
  if (rel) *rel = (float)rel_of(c->h_scal);
  bool saved = in->diag;
  in->diag = true;
  rc = measure_divergence(c, v);
  in->diag = saved;
  if (rc) return rc;
  if (mdiv) *mdiv = in->stats.max_norm_divergence;
  return EFL_OK;
}

int efl_get_stats(efl_instance* in, efl_stats* out) {
  if (!in || !out) return EFL_E_INVALID;
  int rc = resolve_stats(in, true);
  if (rc) return rc;
  *out = in->stats;
  return EFL_OK;
}

// ---- autotuner -----------------------------------------------------------------

int efl_autotune(efl_instance* in, float dt, int apply, efl_tuning* res) {
  if (!in) return EFL_E_INVALID;
  efl_ctx* c = in->ctx;
  const efl_projection orig = in->s.projection;
  const int orig_bottom = in->bottom_cells, orig_th = in->mg_th;
  const bool learned = orig.kind == EFL_PROJECTION_LEARNED;
  const bool tunable = orig.kind == EFL_PROJECTION_MGPCG || learned ||
                       (orig.kind == EFL_PROJECTION_MULTIGRID && orig.rel_tol > 0.0f);
  if (!tunable) return fail(c, EFL_E_STATE, "autotune needs a projection with a residual target");
  if (!in->initialized) return fail(c, EFL_E_STATE, "autotune needs an initialised instance");
  if (learned && !in->nn) return fail(c, EFL_E_STATE, "learned projection without a network");
  if (!(dt > 0.0f)) return fail(c, EFL_E_INVALID, "dt must be positive");
  HIPCHK(c, hipSetDevice(c->device));
  if (in->h.dt != dt) {
    in->h.dt = dt;
    in->dirty = true;
  }
  int rc = upload(in);
  if (rc) return rc;
  hipStream_t st = c->stream;
  IV v{in};
  rc = resolve_stats(in, true);
  if (rc) return rc;
  struct KeepMargin {  // the trials read each solve's outcome: synchronous solves meanwhile
    efl_ctx* c;
    uint32_t m;
    ~KeepMargin() { c->async_margin = m; }
  } keep_margin{c, c->async_margin};
  c->async_margin = 0;
  const size_t n = (size_t)in->s.width * in->s.height;
  float* saved = nullptr;
  if (hipMalloc(&saved, n * 4) != hipSuccess) return fail(c, EFL_E_OOM, "autotune");
  struct Free {
    float* p;
    ~Free() { (void)hipFree(p); }
  } guard{saved};
  HIPCHK(c, hipMemcpyAsync(saved, in->h.p0, n * 4, hipMemcpyDeviceToDevice, st));
  phase_pre(st, v);  // a fresh right-hand side from the current state
  uint32_t trials = 0;
  // one solve from the saved initial guess, best of two; +inf if it misses rel_tol
  auto solve_ms = [&](float* ms) -> int {
    float best = 1e30f;
    for (int rep = 0; rep < 2; ++rep) {
      HIPCHK(c, hipMemcpyAsync(in->h.p0, saved, n * 4, hipMemcpyDeviceToDevice, st));
      HIPCHK(c, hipEventRecord(c->ev[0], st));
      int r = stage_projection(c, v);
      if (r) return r;
      HIPCHK(c, hipEventRecord(c->ev[1], st));
      HIPCHK(c, hipEventSynchronize(c->ev[1]));
      float t = 0.0f;
      HIPCHK(c, hipEventElapsedTime(&t, c->ev[0], c->ev[1]));
      if (!in->stats.converged) t = 1e30f;
      best = std::min(best, t);
      ++trials;
    }
    *ms = best;
    return EFL_OK;
  };
  auto set = [&](uint32_t pre, uint32_t post, uint32_t coarse, int bottom, int th) {
    in->s.projection.mg_pre_iterations = pre;
    in->s.projection.mg_post_iterations = post;
    in->s.projection.mg_coarsest_iterations = coarse;
    in->bottom_cells = bottom;
    in->mg_th = th;
    c->gversion++;
  };
  float def_ms = 0.0f;
  rc = solve_ms(&def_ms);
  if (rc) return rc;
  uint32_t best[3] = {orig.mg_pre_iterations, orig.mg_post_iterations, orig.mg_coarsest_iterations};
  int best_bottom = in->bottom_cells, best_th = in->mg_th;
  float best_ms = def_ms;
  static const uint32_t kSweeps[][3] = {{2, 2, 3}, {1, 1, 2}, {1, 1, 3}, {2, 2, 2}, {2, 2, 4},
                                        {3, 3, 3}, {1, 2, 3}, {2, 1, 3}, {3, 3, 4}};
  const int big = efl::bottom_cells_enable(c->device, 4096);
  const int bottoms[2] = {1024, big};
  for (int bi = 0; bi < (big > 1024 ? 2 : 1); ++bi)
    for (const auto& sw : kSweeps) {
      set(sw[0], sw[1], sw[2], bottoms[bi], best_th);
      float ms = 0.0f;
      rc = solve_ms(&ms);
      if (rc) return rc;
      if (ms < 0.97f * best_ms) {  // a clear gain only: single-solve timings carry a few % of noise
        best_ms = ms;
        best[0] = sw[0];
        best[1] = sw[1];
        best[2] = sw[2];
        best_bottom = bottoms[bi];
      }
    }
  // the smoother tile height, with the chosen sweeps
  for (int th : {0, efl::mg_tile_h(), 8}) {
    if (th == best_th) continue;
    set(best[0], best[1], best[2], best_bottom, th);
    float ms = 0.0f;
    rc = solve_ms(&ms);
    if (rc) return rc;
    if (ms < 0.97f * best_ms) {
      best_ms = ms;
      best_th = th;
    }
  }
  set(best[0], best[1], best[2], best_bottom, best_th);
  // the network's convolution strips, one level at a time
  int strips_before[efl::nn::kLevels];
  uint32_t strip_bits = 0;
  if (learned) {
    for (int l = 0; l < efl::nn::kLevels; ++l) strips_before[l] = in->nn_act.strip[l];
    for (int l = 0; l < efl::nn::kLevels; ++l) {
      float t[2] = {0.0f, 0.0f};
      const int choice[2] = {64, 16};
      for (int k = 0; k < 2; ++k) {
        in->nn_act.strip[l] = choice[k];
        nn_precondition(st, in, true);  // warm
        HIPCHK(c, hipEventRecord(c->ev[0], st));
        for (int r = 0; r < 5; ++r) nn_precondition(st, in, true);
        HIPCHK(c, hipEventRecord(c->ev[1], st));
        HIPCHK(c, hipEventSynchronize(c->ev[1]));
        HIPCHK(c, hipEventElapsedTime(&t[k], c->ev[0], c->ev[1]));
        trials += 5;
      }
      in->nn_act.strip[l] = t[0] <= t[1] ? 64 : 16;
      if (t[0] <= t[1]) strip_bits |= 1u << l;
    }
  }
  HIPCHK(c, hipMemcpyAsync(in->h.p0, saved, n * 4, hipMemcpyDeviceToDevice, st));
  if (!apply) {
    set(orig.mg_pre_iterations, orig.mg_post_iterations, orig.mg_coarsest_iterations, orig_bottom, orig_th);
    if (learned)
      for (int l = 0; l < efl::nn::kLevels; ++l) in->nn_act.strip[l] = strips_before[l];
  }
  HIPCHK(c, hipStreamSynchronize(st));
  if (res) {
    res->pre = best[0];
    res->post = best[1];
    res->coarsest = best[2];
    res->bottom_cells = (uint32_t)best_bottom;
    res->nn_strips = strip_bits;
    res->tile_h = (uint32_t)best_th;
    res->trials = trials;
    res->best_ms = best_ms;
    res->default_ms = def_ms;
  }
  return EFL_OK;
}

int efl_apply_tuning(efl_instance* in, const efl_tuning* t) {
  if (!in || !t) return EFL_E_INVALID;
  efl_ctx* c = in->ctx;
  for (uint32_t k : {t->pre, t->post, t->coarsest})
    if (k < 1 || k > 8) return fail(c, EFL_E_INVALID, "sweep counts must lie in 1..8");
  if (t->tile_h != 0 && t->tile_h != 16 && t->tile_h != 8)
    return fail(c, EFL_E_INVALID, "tile_h must be 0, 16 or 8");
  if (t->bottom_cells != 1024 && t->bottom_cells != 4096)
    return fail(c, EFL_E_INVALID, "bottom_cells must be 1024 or 4096");
  HIPCHK(c, hipSetDevice(c->device));
  if (t->bottom_cells == 4096 && efl::bottom_cells_enable(c->device, 4096) != 4096)
    return fail(c, EFL_E_INVALID, "the device does not grant the 4096-cell bottom solve");
  HIPCHK(c, hipStreamSynchronize(c->stream));  // queued steps keep the settings they were queued with
  in->s.projection.mg_pre_iterations = t->pre;
  in->s.projection.mg_post_iterations = t->post;
  in->s.projection.mg_coarsest_iterations = t->coarsest;
  in->bottom_cells = (int)t->bottom_cells;
  in->mg_th = (int)t->tile_h;
  if (in->nn)
    for (int l = 0; l < efl::nn::kLevels; ++l) in->nn_act.strip[l] = (t->nn_strips >> l) & 1u ? 64 : 16;
  c->gversion++;
  return EFL_OK;
}

// ---- learned preconditioner -------------------------------------------------

size_t efl_nn_param_count(void) { return efl::nn::param_count(); }
double efl_nn_flops_per_pixel(void) { return efl::nn::flops_per_pixel(); }

int efl_nn_create(efl_ctx* c, efl_nn** out) {
  if (!c || !out) return EFL_E_INVALID;
  *out = nullptr;
  HIPCHK(c, hipSetDevice(c->device));
  efl_nn* nn = new efl_nn();
  nn->ctx = c;
  const size_t np = efl::nn::param_count();
  efl::nn::Trainer& t = nn->tr;
  bool ok = hipMalloc(&t.params, np * 4) == hipSuccess && hipMalloc(&t.grads, np * 4) == hipSuccess &&
            hipMalloc(&t.m, np * 4) == hipSuccess && hipMalloc(&t.v, np * 4) == hipSuccess &&
            hipMalloc(&t.loss, sizeof(double)) == hipSuccess;
  for (int k = 0; ok && k < efl::nn::kLayers; ++k) {
    const efl::nn::LayerDef& L = efl::nn::kNet[k];
    const size_t nw = 9 * (size_t)L.cout * (L.ca + L.cb);
    ok = hipMalloc(&nn->pk.wf[k], nw * 2) == hipSuccess && hipMalloc(&nn->pk.wb[k], nw * 2) == hipSuccess;
  }
  if (!ok) {
    efl_nn_destroy(nn);
    return fail(c, EFL_E_OOM, "network parameters");
  }
  HIPCHK(c, hipMemsetAsync(t.params, 0, np * 4, c->stream));
  HIPCHK(c, hipMemsetAsync(t.grads, 0, np * 4, c->stream));
  HIPCHK(c, hipMemsetAsync(t.m, 0, np * 4, c->stream));
  HIPCHK(c, hipMemsetAsync(t.v, 0, np * 4, c->stream));
  efl::nn::pack(c->stream, t.params, nn->pk);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  *out = nn;
  return EFL_OK;
}

int efl_nn_destroy(efl_nn* nn) {
  if (!nn) return EFL_E_INVALID;
  efl_ctx* c = nn->ctx;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (auto* i : c->insts)
    if (i->nn == nn) {
      i->nn = nullptr;
      nn_free_activations(i->nn_act);
    }
  efl::nn::Trainer& t = nn->tr;
  for (void* p : {(void*)t.params, (void*)t.grads, (void*)t.m, (void*)t.v, (void*)t.loss, (void*)t.data,
                  (void*)t.partial, (void*)t.idx, (void*)t.target, (void*)t.mask, (void*)t.rsums, (void*)nn->d_tmp,
                  (void*)nn->d_hist})
    if (p) (void)hipFree(p);
  nn_free_activations(t.act);
  for (int k = 0; k < efl::nn::kLayers; ++k) {
    if (nn->pk.wf[k]) (void)hipFree(nn->pk.wf[k]);
    if (nn->pk.wb[k]) (void)hipFree(nn->pk.wb[k]);
  }
  delete nn;
  return EFL_OK;
}

int efl_nn_set_params(efl_nn* nn, const float* params, size_t n) {
  if (!nn || !params) return EFL_E_INVALID;
  efl_ctx* c = nn->ctx;
  if (n != efl::nn::param_count()) return fail(c, EFL_E_INVALID, "%zu parameters, expected %zu", n, efl::nn::param_count());
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMemcpyAsync(nn->tr.params, params, n * 4, hipMemcpyHostToDevice, c->stream));
  efl::nn::pack(c->stream, nn->tr.params, nn->pk);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return EFL_OK;
}

int efl_nn_get_params(efl_nn* nn, float* params, size_t n) {
  if (!nn || !params) return EFL_E_INVALID;
  efl_ctx* c = nn->ctx;
  if (n != efl::nn::param_count()) return fail(c, EFL_E_INVALID, "%zu parameters, expected %zu", n, efl::nn::param_count());
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMemcpyAsync(params, nn->tr.params, n * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return EFL_OK;
}

int efl_nn_init_params(efl_nn* nn, uint64_t seed) {
  if (!nn) return EFL_E_INVALID;
  std::vector<float> h(efl::nn::param_count());
  efl::nn::init_params(h.data(), seed);
  nn->tr.adam_t = 0;
  efl_ctx* c = nn->ctx;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMemsetAsync(nn->tr.m, 0, h.size() * 4, c->stream));
  HIPCHK(c, hipMemsetAsync(nn->tr.v, 0, h.size() * 4, c->stream));
  return efl_nn_set_params(nn, h.data(), h.size());
}

int efl_nn_dataset_reserve(efl_nn* nn, uint32_t w, uint32_t h, uint32_t capacity) {
  if (!nn) return EFL_E_INVALID;
  efl_ctx* c = nn->ctx;
  if (w == 0 || h == 0 || w % 8 || h % 8 || capacity == 0)
    return fail(c, EFL_E_INVALID, "dataset samples must be multiples of 8 on a side");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  efl::nn::Trainer& t = nn->tr;
  if (t.data) (void)hipFree(t.data);
  t.data = nullptr;
  t.count = t.capacity = 0;
  t.act_n = 0;  // batch buffers follow the sample size
  nn_free_activations(t.act);
  t.w = (int)w;
  t.h = (int)h;
  if (hipMalloc(&t.data, (size_t)capacity * efl::nn::kSamplePlanes * w * h * 4) != hipSuccess)
    return fail(c, EFL_E_OOM, "dataset");
  t.capacity = capacity;
  if (nn->tmp_floats < 2ull * w * h) {
    if (nn->d_tmp) (void)hipFree(nn->d_tmp);
    if (hipMalloc(&nn->d_tmp, 2ull * w * h * 4) != hipSuccess) return fail(c, EFL_E_OOM, "dataset staging");
    nn->tmp_floats = 2ull * w * h;
  }
  return EFL_OK;
}

int efl_nn_add_sample(efl_nn* nn, efl_instance* in, const float* rhs, const float* target) {
  if (!nn || !in || !rhs || !target) return EFL_E_INVALID;
  efl_ctx* c = nn->ctx;
  efl::nn::Trainer& t = nn->tr;
  if (in->ctx != c) return fail(c, EFL_E_INVALID, "instance of another context");
  if ((int)in->s.width != t.w || (int)in->s.height != t.h)
    return fail(c, EFL_E_INVALID, "sample %ux%u, dataset %dx%d", in->s.width, in->s.height, t.w, t.h);
  if (t.count >= t.capacity) return fail(c, EFL_E_RANGE, "dataset full (%u samples)", t.capacity);
  const size_t n = (size_t)t.w * t.h;
  double ss = 0.0, tt = 0.0;
  for (size_t k = 0; k < n; ++k) {
    if (!std::isfinite(rhs[k]) || !std::isfinite(target[k]))
      return fail(c, EFL_E_INVALID, "non-finite sample value at cell %zu", k);
    ss += (double)rhs[k] * rhs[k];
    tt += (double)target[k] * target[k];
  }
  const float sigma = ss > 0.0 ? (float)std::sqrt(ss / (double)n) : 1.0f;
  const float trms = tt > 0.0 ? (float)std::sqrt(tt / (double)n) : 1.0f;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMemcpyAsync(nn->d_tmp, rhs, n * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(nn->d_tmp + n, target, n * 4, hipMemcpyHostToDevice, c->stream));
  efl::nn::make_sample(c->stream, t, t.count, in->d, t.w, t.h, nn->d_tmp, nn->d_tmp + n, sigma, trms);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  t.count += 1;
  return EFL_OK;
}

int efl_nn_sample_count(const efl_nn* nn, uint32_t* count) {
  if (!nn || !count) return EFL_E_INVALID;
  *count = nn->tr.count;
  return EFL_OK;
}

int efl_nn_get_sample(const efl_nn* nn, uint32_t index, float* planes) {
  if (!nn || !planes) return EFL_E_INVALID;
  efl_ctx* c = nn->ctx;
  const efl::nn::Trainer& t = nn->tr;
  if (index >= t.count) return fail(c, EFL_E_RANGE, "sample %u of %u", index, t.count);
  const size_t n = 5ull * t.w * t.h;  // the feature, mask and target planes; the stencil stays on the device
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMemcpyAsync(planes, t.data + (size_t)index * efl::nn::kSamplePlanes * t.w * t.h, n * 4,
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return EFL_OK;
}

int efl_nn_set_loss(efl_nn* nn, int loss) {
  if (!nn) return EFL_E_INVALID;
  if (loss != EFL_NN_LOSS_MSE && loss != EFL_NN_LOSS_RESIDUAL)
    return fail(nn->ctx, EFL_E_INVALID, "unknown loss %d", loss);
  nn->tr.loss_mode = loss;
  return EFL_OK;
}

int efl_nn_train(efl_nn* nn, uint32_t iterations, uint32_t batch, float lr, uint64_t seed, float* loss_history) {
  if (!nn) return EFL_E_INVALID;
  efl_ctx* c = nn->ctx;
  efl::nn::Trainer& t = nn->tr;
  if (t.count == 0 || batch == 0) return fail(c, EFL_E_STATE, "no training samples");
  HIPCHK(c, hipSetDevice(c->device));
  if (efl::nn::trainer_reserve_batch(t, (int)batch)) return fail(c, EFL_E_OOM, "training batch");
  if (nn->hist_cap < iterations) {
    if (nn->d_hist) (void)hipFree(nn->d_hist);
    if (hipMalloc(&nn->d_hist, (size_t)iterations * 8) != hipSuccess) return fail(c, EFL_E_OOM, "loss history");
    nn->hist_cap = iterations;
  }
  uint64_t s = seed;
  std::vector<int> idx((size_t)iterations * batch);  // every batch's samples, SplitMix64
  for (int& x : idx) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    x = (int)(z % t.count);
  }
  for (uint32_t it = 0; it < iterations; ++it) {
    efl::nn::eval_batch(c->stream, t, nn->pk, idx.data() + (size_t)it * batch, (int)batch, true);
    HIPCHK(c, hipMemcpyAsync(nn->d_hist + it, t.loss, 8, hipMemcpyDeviceToDevice, c->stream));
    efl::nn::adam_step(c->stream, t, lr);
    efl::nn::pack(c->stream, t.params, nn->pk);
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (loss_history) {
    std::vector<double> hist(iterations);
    HIPCHK(c, hipMemcpy(hist.data(), nn->d_hist, (size_t)iterations * 8, hipMemcpyDeviceToHost));
    for (uint32_t it = 0; it < iterations; ++it) loss_history[it] = (float)hist[it];
  }
  return launch_status(c, "training");
}

int efl_nn_eval(efl_nn* nn, uint32_t first, uint32_t count, float* loss, float* grads, float* out) {
  if (!nn || count == 0) return EFL_E_INVALID;
  efl_ctx* c = nn->ctx;
  efl::nn::Trainer& t = nn->tr;
  if ((uint64_t)first + count > t.count) return fail(c, EFL_E_RANGE, "samples [%u, %u) of %u", first, first + count, t.count);
  HIPCHK(c, hipSetDevice(c->device));
  if (efl::nn::trainer_reserve_batch(t, (int)count)) return fail(c, EFL_E_OOM, "evaluation batch");
  std::vector<int> idx(count);
  for (uint32_t k = 0; k < count; ++k) idx[k] = (int)(first + k);
  efl::nn::eval_batch(c->stream, t, nn->pk, idx.data(), (int)count, grads != nullptr);
  double l = 0.0;
  HIPCHK(c, hipMemcpyAsync(&l, t.loss, 8, hipMemcpyDeviceToHost, c->stream));
  if (grads)
    HIPCHK(c, hipMemcpyAsync(grads, t.grads, efl::nn::param_count() * 4, hipMemcpyDeviceToHost, c->stream));
  if (out) {
    const size_t npx = (size_t)t.w * t.h * count;
    float* d_out = nullptr;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (hipMalloc(&d_out, npx * 4) != hipSuccess) return fail(c, EFL_E_OOM, "outputs");
    efl::nn::read_output(c->stream, t.act, (int)count, d_out);
    hipError_t e = hipMemcpyAsync(out, d_out, npx * 4, hipMemcpyDeviceToHost, c->stream);
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(d_out);
    if (e != hipSuccess) return fail(c, EFL_E_HIP, "outputs: %s", hipGetErrorString(e));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  int rc = launch_status(c, "evaluation");
  if (rc) return rc;
  if (loss) *loss = (float)l;
  return EFL_OK;
}

int efl_set_learned(efl_instance* in, efl_nn* nn) {
  if (!in) return EFL_E_INVALID;
  efl_ctx* c = in->ctx;
  if (nn && nn->ctx != c) return fail(c, EFL_E_INVALID, "network of another context");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  nn_free_activations(in->nn_act);
  in->nn = nn;
  in->nn_iters = in->nn_fallback = 0;
  if (nn) {
    efl::nn::Activations& a = in->nn_act;
    a.s = efl::nn::Shapes{1, (int)((in->s.width + 7) / 8 * 8), (int)((in->s.height + 7) / 8 * 8)};
    for (int k = 0; k < efl::nn::kTensors; ++k)
      if (hipMalloc(&a.t[k], a.s.elems(k) * 2) != hipSuccess) {
        nn_free_activations(a);
        in->nn = nullptr;
        return fail(c, EFL_E_OOM, "network activations");
      }
  }
  c->gversion++;
  return EFL_OK;
}

int efl_nn_last_solve(efl_instance* in, uint32_t out[2], float* learned_rel) {
  if (!in || !out) return EFL_E_INVALID;
  out[0] = in->nn_iters;
  out[1] = in->nn_fallback;
  if (learned_rel) *learned_rel = in->nn_rel;
  return EFL_OK;
}

}  // extern "C"
