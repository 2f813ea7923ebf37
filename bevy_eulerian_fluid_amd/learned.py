"""Training and evaluation driver of the learned preconditioner.

The network (include/efl.h, efl_nn_*) is trained on the device by the
library's own forward/backward kernels, on pairs that the plain solver
produces while a scene runs.  At a capture step the pipeline is run stage by
stage up to the pressure solve, and from the system A x = b of that step
(MGPCG to ||r|| / ||b|| <= 1e-7 gives x*) it records

- (b, x*): the right-hand side and its solution (zero initial guess);
- (b - A p_prev, x* - p_prev): the residual of the previous step's pressure,
  which is where a step's solve starts, and the correction it needs;
- (b - A x_k, x* - x_k) for x_k = k MGPCG iterations from zero (k = 1, 2, 4):
  residuals of partly converged iterates, like the ones CG feeds the
  preconditioner later in a solve.

Scenes alternate between the coupled-tank generator (lower 40-70 % fluid,
0-4 bodies of each shape) and the batch generator (random gravity direction,
30-70 % fluid box, 0-8 bodies), at a fixed size.

`python -m bevy_eulerian_fluid_amd.learned --out DIR` writes DIR/weights.f32
(efl_nn_param_count() float32 values) and DIR/weights.json (the training
record).  The weights shipped with the package are data/learned_preconditioner.f32.
"""
import argparse
import json
import os
import time

import numpy as np

from . import native, scenes

_HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(_HERE, "data", "learned_preconditioner.f32")
SEED = 0xE0F1D004  # the learned configuration's seed base
DT = 1.0 / 60.0
PRE = ["UPDATE_SOLID", "AREA_FRACTION", "ADVECT_VELOCITY", "APPLY_FORCES", "DIVERGENCE"]
POST = ["SOLVE_VELOCITY", "EXTRAPOLATE_VELOCITY", "ADVECT_LEVELSET", "SOURCE", "REINITIALIZE",
        "LEVELSET_GRADIENT", "FLUID_TO_SOLID"]


def load_weights(path=WEIGHTS):
    """The shipped parameters (float32), or None if the file is absent."""
    if not os.path.exists(path):
        return None
    w = np.fromfile(path, dtype=np.float32)
    if w.size != native.Network.param_count():
        raise ValueError(f"{path}: {w.size} parameters, the network has {native.Network.param_count()}")
    return w


def training_scene(k, size):
    rng = scenes.SplitMix64(SEED + k, 0)
    solve = scenes.projection("mgpcg", max_cycles=200, rel_tol=1e-7)
    if k % 2 == 0:
        fill = rng.uniform(0.4, 0.7)
        n_each = int(rng.uniform(0, 5))
        s = scenes.settings(size, size, proj=solve)
        src = [scenes.source((0.0, -0.5 * size + 0.5 * fill * size), "aabb", half_size=(0.5 * size, 0.5 * fill * size))]
        obs, dens = scenes.random_bodies(rng, size, size, {c: n_each for c in ("circle", "rectangle", "capsule", "triangle")},
                                         rho=99.7)
        return dict(settings=s, init_sources=src, sources=[], obstacles=obs, densities=dens)
    return scenes.batch_instance(rng, proj=solve, size=size)


def _add(net, inst, rhs, target, skipped):
    if np.isfinite(rhs).all() and np.isfinite(target).all() and np.any(rhs) and np.any(target):
        net.add_sample(inst, rhs, target)
    else:
        skipped.append((int(np.isfinite(rhs).all()), int(np.isfinite(target).all())))


def capture(net, inst, skipped, truncated=(1, 2, 4)):
    """Run one step stage by stage and add its samples to net's dataset."""
    for st in PRE:
        inst.run_stage(st, DT)
    p_prev = inst.get("P0").copy()
    inst.run_stage("MG_RESIDUAL", DT)
    r0 = inst.get("MG_R0").copy()
    b = inst.get("DIV").copy()
    inst.run_stage("PROJECTION", DT)
    xs = inst.get("P0").copy()
    _add(net, inst, b, xs, skipped)
    _add(net, inst, r0, xs - p_prev, skipped)
    solve = inst.settings.projection
    for k in truncated:
        inst.set("P0", np.zeros_like(xs))
        inst.set_projection(scenes.projection("mgpcg", max_cycles=k, rel_tol=0.0))
        inst.run_stage("PROJECTION", DT)
        xk = inst.get("P0").copy()
        inst.run_stage("MG_RESIDUAL", DT)
        _add(net, inst, inst.get("MG_R0"), xs - xk, skipped)
    inst.set_projection(solve)
    inst.set("P0", xs)
    for st in POST:
        inst.run_stage(st, DT)


def dataset_summary(net):
    """Per plane: max |value| over the dataset, and whether all are finite."""
    mx = np.zeros(5)
    fin = True
    for i in range(net.count()):
        s = net.sample(i)
        fin = fin and bool(np.isfinite(s).all())
        mx = np.maximum(mx, np.nan_to_num(np.abs(s), nan=np.inf).reshape(5, -1).max(axis=1))
    return dict(finite=fin, max_abs=[float(x) for x in mx])


def generate(ctx, net, size, n_scenes, captures=(4, 12, 30), truncated=(1, 2, 4), log=print):
    """Dataset of n_scenes x len(captures) x (2 + len(truncated)) samples of
    size x size."""
    net.reserve(size, size, n_scenes * len(captures) * (2 + len(truncated)))
    t0 = time.time()
    skipped = []
    for k in range(n_scenes):
        sc = training_scene(k, size)
        inst = ctx.instance(sc["settings"])
        inst.settings = sc["settings"]
        inst.set_sources(sc["init_sources"], True)
        inst.set_obstacles(sc["obstacles"], scenes.bodies_for(sc["obstacles"], sc["densities"]))
        inst.step(DT)  # initialisation
        step = 0
        for c in captures:
            while step < c - 1:
                inst.step(DT)
                step += 1
            capture(net, inst, skipped, truncated)
            step += 1
        inst.close()
        if (k + 1) % max(8, n_scenes // 50) == 0:
            log(f"[learned] scenes {k + 1}/{n_scenes}, samples {net.count()}, {time.time() - t0:.1f} s")
    if skipped:
        log(f"[learned] skipped {len(skipped)} samples (rhs finite, target finite): {skipped[:8]}")
    return net.count()


def train(net, iterations, batch, lr, seed, log=print):
    """Adam with the learning rate dropped 10x for the last third."""
    a = int(iterations * 2 / 3)
    t0 = time.time()
    h0 = net.train(10, batch, lr, seed + 2)
    log(f"[learned] first 10 steps: {' '.join(f'{x:.4g}' for x in h0)}")
    h1 = net.train(a - 10, batch, lr, seed)
    h1 = np.concatenate([h0, h1])
    log(f"[learned] {a} steps at lr {lr}: loss {h1[:20].mean():.4g} -> {h1[-20:].mean():.4g}, {time.time() - t0:.1f} s")
    h2 = net.train(iterations - a, batch, lr / 10, seed + 1)
    log(f"[learned] {iterations - a} steps at lr {lr / 10}: loss -> {h2[-20:].mean():.4g}, {time.time() - t0:.1f} s")
    return np.concatenate([h1, h2])


def evaluate(ctx, weights, size=1024, steps=10, warmup=10, rel_tol=1e-5, seed=scenes.SEED_BASE + 4):
    """The learned projection against plain MGPCG on the same states: a
    coupled-tank scene (lower 55 % fluid, 16 bodies, static) is stepped with
    MGPCG; before each measured step the learned instance gets the MGPCG
    instance's state, then both solve the same system to rel_tol.  Returns
    per-step iterations and projection times (hipEvents, eager)."""
    sc = scenes.coupled_tank(size, seed=seed, proj=scenes.projection("mgpcg", max_cycles=200, rel_tol=rel_tol))
    a = ctx.instance(sc["settings"])
    b = ctx.instance(sc["settings"])
    n_obs = len(sc["obstacles"])
    a.set_obstacles(sc["obstacles"], scenes.bodies_for(sc["obstacles"], sc["densities"]))  # two-way coupled
    b.set_obstacles(sc["obstacles"])  # follows a's bodies
    for i in (a, b):
        i.set_sources(sc["init_sources"], True)
        i.step(DT)  # initialisation
    net = native.Network(ctx)
    if weights is None:
        net.init_params(1)
    else:
        net.set_params(weights)
    b.set_learned(net)
    b.set_projection(scenes.projection("learned", max_cycles=200, rel_tol=rel_tol))
    for _ in range(warmup):
        a.step(DT)
    ctx.set_timing(True)
    rows = []
    for _ in range(steps):
        for f in ("U0", "V0", "P0", "LS_AIR0", "LS_AIR1"):
            b.set(f, a.get(f))
        b.set_obstacles(a.get_obstacles(n_obs))
        a.step(DT)
        b.step(DT)
        sa, sb = a.stats(), b.stats()
        nn_it, fb, nn_rel = b.learned_iterations(with_rel=True)
        rows.append(dict(mgpcg_iters=sa.cycles, mgpcg_ms=sa.projection_ms, learned_iters=nn_it, fallback_iters=fb,
                         network_part_rel=nn_rel,
                         learned_ms=sb.projection_ms, learned_converged=int(sb.converged),
                         learned_rel=sb.rel_residual))
    ctx.set_timing(False)
    net.close()
    a.close()
    b.close()
    mean = lambda k: float(np.mean([r[k] for r in rows]))
    return dict(size=size, steps=steps, rel_tol=rel_tol, mgpcg_iters=mean("mgpcg_iters"), mgpcg_ms=mean("mgpcg_ms"),
                learned_iters=mean("learned_iters"), fallback_iters=mean("fallback_iters"),
                learned_ms=mean("learned_ms"), fallbacks=int(sum(r["fallback_iters"] > 0 for r in rows)),
                converged=int(sum(r["learned_converged"] for r in rows)), rows=rows)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--evaluate", default=None, help="weights file: evaluate instead of training")
    ap.add_argument("--eval-size", type=int, default=1024)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--scenes", type=int, default=32)
    ap.add_argument("--iterations", type=int, default=1500)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--loss", choices=("mse", "residual"), default="mse")
    ap.add_argument("--captures", default="4,12,30", help="steps of each scene at which samples are taken")
    ap.add_argument("--truncated", default="1,2,4", help="MGPCG iteration counts of the partly converged samples")
    ap.add_argument("--init", default=None, help="weights file to start from (default: He-normal from --seed)")
    a = ap.parse_args()
    ints = lambda t: tuple(int(x) for x in t.split(",") if x)
    os.makedirs(a.out, exist_ok=True)
    ctx = native.Context(0)
    if a.evaluate:
        w = None if a.evaluate == "untrained" else np.fromfile(a.evaluate, dtype=np.float32)
        res = evaluate(ctx, w, size=a.eval_size)
        json.dump(res, open(os.path.join(a.out, f"eval_{a.eval_size}.json"), "w"), indent=1)
        print(json.dumps({k: v for k, v in res.items() if k != "rows"}), flush=True)
        ctx.close()
        return
    net = native.Network(ctx)
    net.init_params(a.seed)
    if a.init:
        net.set_params(np.fromfile(a.init, dtype=np.float32))
    net.set_loss(a.loss)
    t0 = time.time()
    n = generate(ctx, net, a.size, a.scenes, ints(a.captures), ints(a.truncated), log=lambda m: print(m, flush=True))
    t_gen = time.time() - t0
    print(f"[learned] {n} samples of {a.size}x{a.size} in {t_gen:.1f} s; {dataset_summary(net)}", flush=True)
    t0 = time.time()
    hist = train(net, a.iterations, a.batch, a.lr, a.seed, log=lambda m: print(m, flush=True))
    t_train = time.time() - t0
    w = net.get_params()
    w.tofile(os.path.join(a.out, "weights.f32"))
    rec = dict(params=int(w.size), size=a.size, scenes=a.scenes, samples=n, loss=a.loss, captures=a.captures,
               truncated=a.truncated, init=os.path.basename(a.init) if a.init else None, iterations=a.iterations,
               batch=a.batch, lr=a.lr, seed=a.seed, seed_base=SEED, generate_s=round(t_gen, 1),
               train_s=round(t_train, 1), loss_first=float(hist[:20].mean()), loss_last=float(hist[-20:].mean()),
               loss_every_50=[float(x) for x in hist[::50]])
    json.dump(rec, open(os.path.join(a.out, "weights.json"), "w"), indent=1)
    print(json.dumps({k: v for k, v in rec.items() if k != "loss_every_50"}), flush=True)
    net.close()
    ctx.close()


if __name__ == "__main__":
    main()
