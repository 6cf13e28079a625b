"""Train the learned warm start's network on the GPU.

Training pairs come from randomised dam breaks (seeded block sizes at box
scales 1, 2 and 4 of the 3.2 x 1.6 x 1.6 box, jitter uniform in
[0, 0.05 h]).  A pair is (features of a substep's first density pass,
(Lambda - Lambda_prev) / S), Lambda being the sum of the multipliers a solve
applied and Lambda_prev the previous substep's, S = PBF_WARM_LAMBDA_SCALE:
the network learns how far the multiplier sum moves from one substep to the
next.  Two sources of trajectories:

* "tolerance" (default; sph-cpp_amd/warm_model.npy): each scene advances
  with the solves the warm start serves, iterating to a mean density error of
  ``tol`` warm-started from the previous sums.
* "retry" (sph-cpp_amd/warm_model_retry.npy): as "tolerance" with
  PBF_OPT_WARM_RETRY, so that substeps whose start was retried at half its
  multipliers (and kept) are among the samples.
* "plain50" (sph-cpp_amd/warm_model_plain50.npy): each scene advances with
  the plain solver run to 50 Jacobi iterations per substep, so Lambda is the
  multiplier the plain solve converges to.  The features come from a zero
  warm start, which moves nothing, so the step is the plain one.

The bench runs both networks.  After every ``every``-th sampled substep the
features become samples (a stride over the cell order).  The network is
trained with the library's own MFMA forward/backward and Adam kernels
(pbf_warm_model_train).  Afterwards the output layer is scaled by the
least-squares gain of the output on samples of held-out scenes at the largest
box scale (clamped to [0, 1]), so that only as much of the correction is
applied as carries over to large scenes the network has not seen.

    python sph-cpp_amd/train_warm.py [--pairs 2000000] [--epochs 40] [--source tolerance|retry|plain50]
                                     [--out sph-cpp_amd/warm_model.npy]

Prints one JSON line per scene and a summary; writes the 4801 parameters
(float32, numpy .npy).
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
pbf = importlib.import_module("sph-cpp_amd")
scenes = importlib.import_module("sph-cpp_amd.scenes")

MODEL_PATH = os.path.join(HERE, "warm_model.npy")
PBF_HIDDEN = pbf.PBF_WARM_HIDDEN


def random_scene(rng, cap=60, scale=None):
    scale = int(scale or rng.choice([1, 2, 4]))
    box = (3.2 * scale, 1.6 * scale, 1.6 * scale)
    counts = (int(rng.integers(16, 49)) * scale, int(rng.integers(16, 33)) * scale, int(rng.integers(12, 33)) * scale)
    amp = float(rng.uniform(0.0, 0.05 * scenes.H))
    seed = int(rng.integers(1 << 31))
    params, pos, vel = scenes.dam_break(counts, box, seed=seed, cap=cap, jitter_amplitude=amp)
    params.options |= pbf.PBF_OPT_SOLVE_MEAN_ERROR | pbf.PBF_OPT_KEEP_WARM_FEATURES
    return params, pos, vel, {"scale": scale, "counts": counts, "jitter": round(amp, 6), "seed": seed}


PLAIN_ITERATIONS = 50


def generate(model, pairs, seed=0, steps=200, cap=60, tol=0.01, max_slots=128, per_sample=10_000,
             scenes_min=8, log=print, budget=None, scale=None, source="tolerance"):
    """Add samples until the model holds `pairs`.  Each scene contributes at
    most `budget` samples (default: pairs / scenes_min), each sampled substep
    at most per_sample (a stride over the cell order), and the sampled
    substeps are spread over the whole trajectory."""
    rng = np.random.default_rng(seed)
    k = 0
    t0 = time.time()
    budget = budget or pairs // scenes_min + 1
    while model.sample_count() < pairs:
        start = model.sample_count()
        params, pos, vel, info = random_scene(rng, cap, scale)
        if source == "retry":
            params.options |= pbf.PBF_OPT_WARM_RETRY
        t = pbf.Particles(params)
        t.add(pos, pos, vel)
        log(json.dumps({"scene": k, **info, "particles": len(pos)}))
        stride = max(1, len(pos) // per_sample)
        per_substep = (len(pos) + stride - 1) // stride
        every = max(2, steps * per_substep // budget)
        kept = 0
        zero = np.zeros(len(pos), np.float32)
        for step in range(1, steps + 1):
            if source == "plain50":
                # the plain solver to 50 iterations; the zero warm start only
                # computes the first density pass's features
                t.set_warm_lambdas(zero)
                t.solve(0.0, PLAIN_ITERATIONS, pbf.PBF_WARM_GIVEN if step > 1 else pbf.PBF_WARM_NONE)
            else:
                t.solve(tol, max_slots, pbf.PBF_WARM_PREVIOUS if step > 1 else pbf.PBF_WARM_NONE)
            if step > 1 and step % every == 0 and t.solve_info()["warm_start_kept"]:
                model.add_from(t, stride)
                kept += 1
                if model.sample_count() >= pairs or model.sample_count() - start >= budget:
                    break
        t.close()
        log(json.dumps({"scene": k, "sampled_substeps": kept, "samples": model.sample_count(),
                        "seconds": round(time.time() - t0, 1)}))
        k += 1


VALIDATION_PAIRS = 80_000
VALIDATION_SCENES = 4
VALIDATION_SCALE = 4  # held-out scenes at the largest box scale (the million-particle scenes the start serves)


def output_gain(model, first, count):
    """Least-squares gain g of the network output on samples [first, first +
    count): the loss of g * net is a parabola in g, fitted from g = 0, 1/2, 1
    (the output layer w3, b3 is the last 65 parameters, so scaling it scales
    the output).  Returns g clamped to [0, 1] and the three losses."""
    p = model.params()
    losses = []
    for g in (0.0, 0.5, 1.0):
        q = p.copy()
        q[-(PBF_HIDDEN + 1):] *= np.float32(g)
        model.set_params(q)
        losses.append(model.gradient(first, count)[1])
    model.set_params(p)
    return gain_from_losses(*losses), losses


def gain_from_losses(c, lh, l1):
    """Minimiser of the parabola through (0, c), (1/2, lh), (1, l1), clamped to
    [0, 1]; 0 when the three losses do not curve upwards."""
    a = 2.0 * (l1 - c) - 4.0 * (lh - c)   # L(g) = a g^2 - 2 b g + c
    b = 0.5 * (a - (l1 - c))
    g = b / a if a > 0 else 0.0
    return float(min(max(g, 0.0), 1.0))


def train(pairs=2_000_000, epochs=40, batch=4096, lr=3e-3, seed=0, log=print, scale=None, scenes_min=8,
          source="tolerance"):
    model = pbf.WarmModel(pairs + 2 * VALIDATION_PAIRS)
    t0 = time.time()
    generate(model, pairs, seed, log=log, scale=scale, scenes_min=scenes_min, source=source)
    t1 = time.time()
    model.init(seed)
    n = model.sample_count()
    _, l0 = model.gradient(0, n)
    half = epochs // 2
    m1 = model.train(half, batch, lr, seed + 1)
    m2 = model.train(epochs - half, batch, lr / 3, seed + 2)
    _, l1 = model.gradient(0, n)
    t2 = time.time()
    # held-out scenes (appended after the trained range): the output gain that
    # minimises their loss is folded into the output layer, so a correction
    # that does not carry over to new scenes shrinks towards zero and the warm
    # start towards the previous multiplier sums it corrects
    generate(model, n + VALIDATION_PAIRS, seed + 1000, log=log, budget=VALIDATION_PAIRS // VALIDATION_SCENES + 1,
             scale=VALIDATION_SCALE, source=source)
    nv = model.sample_count() - n
    gain, vl = output_gain(model, n, nv)
    p = model.params()
    p[-(PBF_HIDDEN + 1):] *= np.float32(gain)
    model.set_params(p)
    summary = {"source": source, "samples": n, "epochs": epochs, "batch": batch, "loss_before": l0, "loss_after": l1,
               "mean_batch_loss": [m1, m2], "validation_samples": nv,
               "validation_loss_at_gain_0_half_1": vl, "output_gain": gain,
               "data_seconds": round(t1 - t0, 1), "train_seconds": round(t2 - t1, 1)}
    return model, summary


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--pairs", type=int, default=2_000_000)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--scale", type=int, default=0, choices=(0, 1, 2, 4),
                    help="box scale of the training scenes (0: 1, 2 and 4 at random)")
    ap.add_argument("--scenes", type=int, default=8, help="training scenes (each gives pairs / scenes samples)")
    ap.add_argument("--source", default="tolerance", choices=("tolerance", "retry", "plain50"),
                    help="plain50: plain solves of 50 iterations; tolerance: warm-started tolerance solves; "
                         "retry: the same with PBF_OPT_WARM_RETRY")
    ap.add_argument("--out", default=MODEL_PATH)
    a = ap.parse_args(argv)
    model, summary = train(a.pairs, a.epochs, a.batch, a.lr, a.seed, log=lambda s: print(s, flush=True),
                           scale=a.scale or None, scenes_min=a.scenes, source=a.source)
    np.save(a.out, model.params())
    print(json.dumps({"out": a.out, **summary}), flush=True)
    model.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
