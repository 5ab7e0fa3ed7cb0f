"""Learned warm start of the atlas refinement (project-specified; the reference
has no counterpart).

Per texel, the 8 highest-confidence projected samples ``[R, G, B, confidence,
cos]`` (pdr_bake_features) go through a 40 -> 64 -> 64 -> 3 GELU network on
matrix cores (pdr_mlp_infer); its output, added to the samples'
confidence-weighted mean colour, starts the Adam refinement in place of the
weighted-average bake.  The network is trained on the GPU by the library's own
gradient kernels (pdr_mlp_adam_step) on (features, refined atlas) pairs that
the plain method - bake followed by the full refinement - produces on seeded
scenes.  The plain method stays the fallback and the yardstick: a warm-started
refinement is scored by the same photometric PSNR.

Run ``python -m photara_drender_amd.warmstart --train`` to regenerate the
shipped weights (``warmstart_mlp.npy``).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from dataclasses import dataclass

import numpy as np

from . import _native

WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "warmstart_mlp.npy")
IN, HID, OUT = 40, 64, 3
PARAMS = _native.MLP_PARAMS

# Training scenes: the 2048^2 / 64-view refinement workload scaled to a 512^2
# atlas (same texels per chart, mesh vertices per texel, views per texel and
# pixels per texel in each view: 64 views of 192 x 256).
TRAIN_SCENE = dict(nx=250, ny=125, charts=4, atlas_size=512, views=64, size=(192, 256), noise=0.02)
TRAIN_SEED0 = 1000


def init_params(seed=0) -> np.ndarray:
    """He-initialised hidden layers and a zero output layer (the network starts as the base colour)."""
    rng = np.random.default_rng(seed)
    w1 = rng.normal(0.0, math.sqrt(2.0 / IN), (IN, HID))
    w2 = rng.normal(0.0, math.sqrt(2.0 / HID), (HID, HID))
    parts = [w1, np.zeros(HID), w2, np.zeros(HID), np.zeros((HID, OUT)), np.zeros(OUT)]
    return np.concatenate([p.ravel() for p in parts]).astype(np.float32)


def _stream():
    import torch

    return torch.cuda.current_stream().cuda_stream


class WarmStartMLP:
    """Float32 master parameters (and Adam moments) of the warm-start network on one GPU."""

    def __init__(self, context, params=None, device="cuda:0"):
        import torch

        self.context = context
        self.device = torch.device(device)
        p = init_params(0) if params is None else np.asarray(params, np.float32).reshape(-1)
        if p.size != PARAMS:
            raise ValueError(f"warm-start parameters must have {PARAMS} values")
        self.params = torch.tensor(p, device=self.device)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.step = 0
        self._loss = torch.zeros(2, dtype=torch.float32, device=self.device)

    @classmethod
    def load(cls, context, path: str = WEIGHTS, device="cuda:0") -> "WarmStartMLP":
        return cls(context, np.load(path, allow_pickle=False), device)

    def save(self, path: str) -> None:
        np.save(path, self.params.detach().cpu().numpy().astype(np.float32), allow_pickle=False)

    def _features(self, features):
        import torch

        x = features.reshape(-1, IN)
        if x.dtype != torch.float32 or not x.is_contiguous() or x.data_ptr() % 16:
            x = x.to(torch.float32).contiguous().clone()
        return x

    def predict(self, features, clamp=(0.0, 1.0)):
        """Warm-start colour [..., 3] of features [..., 40] (0 where a texel has no sample)."""
        import torch

        x = self._features(features)
        out = torch.empty((x.shape[0], OUT), dtype=torch.float32, device=x.device)
        _native.check(_native.lib().pdr_mlp_infer(self.context.handle, x.data_ptr(), x.shape[0],
                                                  self.params.data_ptr(), out.data_ptr(), _stream()))
        if clamp is not None:
            out.clamp_(clamp[0], clamp[1])
        return out.reshape(features.shape[:-1] + (OUT,))

    def gradient(self, features, target, valid=None):
        """(gradient [PARAMS], loss, valid count) of the mean squared error over valid texels."""
        import torch

        x = self._features(features)
        t = target.reshape(-1, OUT).to(torch.float32).contiguous()
        v = None if valid is None else valid.reshape(-1).to(torch.float32).contiguous()
        g = torch.empty_like(self.params)
        lc = torch.empty(2, dtype=torch.float32, device=x.device)
        _native.check(_native.lib().pdr_mlp_gradient(self.context.handle, x.data_ptr(), t.data_ptr(),
                                                     None if v is None else v.data_ptr(), x.shape[0],
                                                     self.params.data_ptr(), g.data_ptr(), lc.data_ptr(), _stream()))
        return g, float(lc[0]), float(lc[1])

    def train_step(self, features, target, valid=None, lr=1e-3, betas=(0.9, 0.999), eps=1e-8):
        """One Adam step on a batch; returns the device tensor (loss, valid count) of the batch."""
        import torch

        x = self._features(features)
        t = target.reshape(-1, OUT).to(torch.float32).contiguous()
        v = None if valid is None else valid.reshape(-1).to(torch.float32).contiguous()
        self.step += 1
        _native.check(_native.lib().pdr_mlp_adam_step(
            self.context.handle, x.data_ptr(), t.data_ptr(), None if v is None else v.data_ptr(), x.shape[0],
            self.params.data_ptr(), self.m.data_ptr(), self.v.data_ptr(), self.step, float(lr), float(betas[0]),
            float(betas[1]), float(eps), self._loss.data_ptr(), _stream()))
        return self._loss


def problem_features(context, problem, **kw):
    """Warm-start features [S, S, 40] of a refinement problem's views (same inputs as its bake)."""
    from .bake import bake_features

    return bake_features(context, problem.positions, problem.normals, problem.uv, problem.indices, problem.images,
                         problem.world_to_clip, problem.camera_positions, problem.masks,
                         resolution=(problem.atlas_size, problem.atlas_size), **kw)


def warm_start_atlas(context, problem, mlp: WarmStartMLP, padding=8):
    """Returns (atlas [S, S, 3], features, seconds): features, MLP inference and
    the same guard-band padding as the plain method's bake, timed to completion."""
    import torch

    from .bake import pad_atlas

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    feats = problem_features(context, problem)
    atlas = mlp.predict(feats).contiguous()
    if padding:
        pad_atlas(context, atlas, (feats[..., 3::5].sum(-1) > 0).to(torch.float32), padding)
    torch.cuda.synchronize()
    return atlas, feats, time.perf_counter() - t0


def plain_refine_kwargs(problem, steps=1000):
    """The plain method's refinement: graph-captured Adam, lr 1e-2 cosine to 2.5e-4, batch 4,
    chart TV 1e-3, joint seam 1e-2, then the 30-step seam polish."""
    return dict(steps=steps, batch_size=4, learning_rate=1e-2, minimum_learning_rate=2.5e-4, seam_polish_steps=30,
                tv_weight=1e-3, chart_ids=problem.chart_texels, joint_seam_weight=1e-2, use_graph=True)


def refine_problem_atlas(context, problem, init, checkpoint_every=0, **overrides):
    from .refine import refine_atlas

    kw = plain_refine_kwargs(problem)
    kw.update(overrides)
    return refine_atlas(context, init, problem.positions, problem.uv, problem.indices, problem.images,
                        problem.world_to_clip, problem.masks, None, problem.seam_pairs,
                        checkpoint_every=checkpoint_every, **kw)


@dataclass
class TrainingSet:
    features: object  # [N, 40] device
    target: object    # [N, 3] device
    scenes: int
    seconds: float


def training_pairs(context, seed, log=None):
    """(features [n, 40], refined colour [n, 3]) over the texels with samples of one
    seeded training scene; the target is the plain method's refined atlas."""
    import torch

    from . import workloads

    prob = workloads.refine_problem(context, seed=seed, **TRAIN_SCENE)
    init, _, _ = workloads.bake_initial_atlas(context, prob)
    refined = refine_problem_atlas(context, prob, init).color
    feats = problem_features(context, prob)
    has = feats[..., 3::5].sum(-1) > 0
    if log:
        log(f"scene {seed}: bake {workloads.atlas_psnr(context, prob, init):.2f} dB -> refined "
            f"{workloads.atlas_psnr(context, prob, refined):.2f} dB, {int(has.sum())} texels")
    out = feats[has].contiguous(), refined[has].contiguous()
    del prob, init, refined, feats
    torch.cuda.empty_cache()
    return out


def build_training_set(context, scenes=200, seed0=TRAIN_SEED0, log=print) -> TrainingSet:
    import torch

    t0 = time.perf_counter()
    xs, ys = [], []
    for k in range(scenes):
        x, y = training_pairs(context, seed0 + k, log)
        xs.append(x)
        ys.append(y)
    x = torch.cat(xs)
    y = torch.cat(ys)
    return TrainingSet(x, y, scenes, time.perf_counter() - t0)


def train(context, data: TrainingSet, steps=4000, batch=65536, lr=2e-3, min_lr=2e-5, seed=0, log=print,
          mlp: WarmStartMLP | None = None) -> WarmStartMLP:
    """Adam with a cosine learning rate over shuffled mini-batches, all on the GPU."""
    import torch

    mlp = mlp or WarmStartMLP(context, init_params(seed))
    n = data.features.shape[0]
    gen = torch.Generator(device=data.features.device)
    gen.manual_seed(seed)
    order, pos = None, n
    t0 = time.perf_counter()
    running = torch.zeros(2, dtype=torch.float64, device=data.features.device)
    if log:
        _, base_loss, _ = mlp.gradient(data.features[:1 << 20], data.target[:1 << 20])
        log(f"initial loss {base_loss:.3e} (rmse {math.sqrt(base_loss):.4f})")
    window = 0
    for step in range(steps):
        if pos + batch > n:
            order = torch.randperm(n, generator=gen, device=data.features.device)
            pos = 0
        sel = order[pos:pos + batch]
        pos += batch
        rate = min_lr + 0.5 * (lr - min_lr) * (1.0 + math.cos(math.pi * step / max(steps - 1, 1)))
        lc = mlp.train_step(data.features[sel], data.target[sel], None, rate)
        running += lc.to(torch.float64)
        window += 1
        if log and ((step + 1) % 500 == 0 or step + 1 == steps):
            r = float(running[0]) / window
            log(f"step {step + 1}: loss {r:.3e} (rmse {math.sqrt(max(r, 0)):.4f}), lr {rate:.2e}, "
                f"{time.perf_counter() - t0:.1f} s")
            running.zero_()
            window = 0
    torch.cuda.synchronize()
    return mlp


def _steps_to(psnr, target, every):
    return next((every * (i + 1) for i, p in enumerate(psnr) if p >= target), None)


def evaluate(context, problem, mlp: WarmStartMLP, warm=None, checkpoint_every=25, log=None):
    """Plain bake + full refinement against the MLP warm start on one problem.

    Runs are timed without checkpoints (pdr_refine is synchronous), then
    repeated with atlas checkpoints every ``checkpoint_every`` steps for the PSNR
    curve (masked, covered pixels of every view against the noise-free renders).
    The target is the plain method's final PSNR - 0.1 dB.  Time to target is
    measured: the wall time of a run that makes the initial atlas, precomputes
    and stops at the first checkpoint reaching the target (the same schedule,
    stop_after = that step, no seam polish), whose atlas is that checkpoint;
    beside it the estimate initial atlas + precompute + steps x time per step.
    ``warm`` overrides the warm start's refinement options (default: the plain
    schedule started at a learning rate of 2e-3); the ablation runs the plain
    method's initial atlas with those same options, which isolates the
    network's part."""
    import torch

    from . import workloads

    warm = dict(learning_rate=2e-3) if warm is None else dict(warm)

    def run(init, init_s, make_init, target=None, **kw):
        timed = refine_problem_atlas(context, problem, init, **kw)
        ck = refine_problem_atlas(context, problem, init, checkpoint_every=checkpoint_every, **kw)
        steps = kw.get("steps", 1000)
        psnr = [workloads.atlas_psnr(context, problem, ck.checkpoints[i]) for i in range(ck.checkpoints.shape[0])]
        per_step = timed.optimization_seconds / (steps + 30)
        out = dict(init_s=init_s, init_psnr_db=workloads.atlas_psnr(context, problem, init),
                   final_psnr_db=workloads.atlas_psnr(context, problem, timed.color),
                   precompute_s=timed.precompute_seconds, optimize_s=timed.optimization_seconds,
                   ms_per_step=per_step * 1e3, psnr_checkpoints_db=[round(p, 3) for p in psnr])
        if target is None:
            target = psnr[-1] - 0.1
        n = _steps_to(psnr, target, checkpoint_every)
        out["steps_to_target"] = n
        out["seconds_to_target_estimate"] = None if n is None else init_s + timed.precompute_seconds + n * per_step
        out["seconds_to_target"] = out["psnr_at_stop_db"] = None
        if n is not None:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            start = make_init()[0]
            stopped = refine_problem_atlas(context, problem, start, stop_after=n, **dict(kw, seam_polish_steps=0))
            torch.cuda.synchronize()
            out["seconds_to_target"] = time.perf_counter() - t0
            out["psnr_at_stop_db"] = workloads.atlas_psnr(context, problem, stopped.color)
        return out, target

    bake = lambda: workloads.bake_initial_atlas(context, problem)  # noqa: E731
    warm_init = lambda: warm_start_atlas(context, problem, mlp)  # noqa: E731
    bake()  # warm-up
    init, _, bake_s = bake()
    plain, target = run(init, bake_s, bake)
    warm_init()  # warm-up
    winit, _, ws_s = warm_init()
    ws, _ = run(winit, ws_s, warm_init, target, **warm)
    ablation, _ = run(init, bake_s, bake, target, **warm)
    ws["options"] = ablation["options"] = warm
    speed = lambda r: None if r["seconds_to_target"] is None else plain["seconds_to_target"] / r["seconds_to_target"]
    out = {"target_psnr_db": target, "plain": plain, "warm_start": ws, "ablation_plain_init_warm_options": ablation,
           "speedup_time_to_target": speed(ws), "ablation_speedup": speed(ablation)}
    if log:
        log(json.dumps(out))
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--train", action="store_true", help="generate the training scenes and train")
    ap.add_argument("--scenes", type=int, default=200)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--batch", type=int, default=65536)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--evaluate", action="store_true",
                    help="compare plain bake + refinement with the warm start on the 2048^2 / 64-view scene")
    ap.add_argument("--warm-lr", type=float, nargs="*", default=[2e-3],
                    help="starting learning rates of the warm-started refinement to evaluate")
    ap.add_argument("--warm-steps", type=int, nargs="*", default=[1000],
                    help="schedule lengths (Adam steps) of the warm-started refinement to evaluate")
    args = ap.parse_args(argv)
    if not (args.train or args.evaluate):
        ap.print_help()
        return 0

    def log(msg):
        print(msg, flush=True)

    ctx = _native.Context(0, flags=0)
    if not args.train:
        from . import workloads

        mlp = WarmStartMLP.load(ctx, args.out)
        prob = workloads.refine_problem(ctx)
        for lr in args.warm_lr:
            for steps in args.warm_steps:
                evaluate(ctx, prob, mlp, warm=dict(learning_rate=lr, steps=steps), log=log)
        return 0
    data = build_training_set(ctx, args.scenes, log=log)
    log(f"training set: {data.features.shape[0]} texels from {data.scenes} scenes in {data.seconds:.1f} s")
    t0 = time.perf_counter()
    mlp = train(ctx, data, steps=args.steps, batch=args.batch, lr=args.lr, log=log)
    train_s = time.perf_counter() - t0
    mlp.save(args.out)
    meta = {"scenes": data.scenes, "seed0": TRAIN_SEED0, "scene": {k: list(v) if isinstance(v, tuple) else v
                                                                 for k, v in TRAIN_SCENE.items()},
            "texels": int(data.features.shape[0]), "steps": args.steps, "batch": args.batch, "lr": args.lr,
            "data_seconds": data.seconds, "train_seconds": train_s}
    with open(os.path.splitext(args.out)[0] + ".json", "w") as f:
        json.dump(meta, f, indent=1)
    log(json.dumps(meta))
    return 0


if __name__ == "__main__":
    sys.exit(main())
