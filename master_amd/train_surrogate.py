"""Trains the warm-start surrogate on steady fields that the GPU DSMC itself
generates, then compares cold-start and surrogate warm-start DSMC on held-out
geometries.

python -m master_amd.train_surrogate [--train N] [--epochs E] [--eval G] [--out DIR]

Writes DIR/surrogate_params.npy (the network, float16, layout of
include/dsmc_unet.h) and DIR/surrogate_report.json (training losses, error
curves of cold and warm starts, steps and wall time to the 2% threshold).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from master_amd import surrogate as S  # noqa: E402


def log(*a):
    print(*a, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=1024, help="training geometries")
    ap.add_argument("--train-seed", type=int, default=100)
    ap.add_argument("--t-eq", type=int, default=600)
    ap.add_argument("--t-avg", type=int, default=3000)
    ap.add_argument("--epochs", type=int, default=50)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--eval", type=int, default=32, help="held-out geometries (seeds 5000..)")
    ap.add_argument("--eval-steps", type=int, default=2000)
    ap.add_argument("--window", type=int, default=200)
    ap.add_argument("--replicas", type=int, default=8)
    ap.add_argument("--out", default=os.path.join(ROOT, "surrogate_out"))
    ap.add_argument("--params", default=None, help="evaluate these parameters (.npy) instead of training")
    ap.add_argument("--synthetic", type=int, default=0,
                    help="train on this many random fields for --epochs, no DSMC and no evaluation "
                         "(kernel timing of the training step)")
    args = ap.parse_args()
    os.makedirs(args.out, exist_ok=True)
    net = S.Net(edge=S.EDGE, max_batch=args.batch, seed=1)
    losses, t_gen, t_train, nfields = [], 0.0, 0.0, 0
    if args.synthetic:
        rng = np.random.default_rng(0)
        cells = (S.EDGE // 4) ** 3
        net.set_data((rng.random((args.synthetic, S.EDGE ** 3)) < 0.4).astype(np.uint8),
                     rng.normal(size=(args.synthetic, 2)).astype(np.float32),
                     rng.normal(size=(args.synthetic, cells, 5)).astype(np.float32),
                     np.ones((args.synthetic, cells), np.float32))
        t1 = time.time()
        losses = net.train(args.epochs, args.batch, args.lr, seed=0)
        n_batches = args.epochs * (args.synthetic // args.batch)
        log(f"[surrogate] {n_batches} training batches of {args.batch} in {time.time() - t1:.2f}s, "
            f"losses {losses}")
        return
    if args.params:
        net.set_params(np.load(args.params).astype(np.float32))
        args.epochs = 0
    else:
        t0 = time.time()
        specs = S.sample_specs(args.train, args.train_seed)
        masks, scal, tgt, w = S.generate(specs, args.t_eq, args.t_avg, log=log)
        t_gen = time.time() - t0
        nfields = len(masks)
        log(f"[surrogate] data: {nfields} fields in {t_gen:.1f}s; target rms {np.sqrt((tgt ** 2).mean(axis=(0, 1)))}")
        net.set_data(masks, scal, tgt, w)
        del masks, scal, tgt, w
    t1 = time.time()
    for ep in range(args.epochs):
        # cosine decay from lr to lr / 10
        lr = args.lr * (0.1 + 0.9 * 0.5 * (1.0 + np.cos(np.pi * ep / max(args.epochs - 1, 1))))
        losses.append(float(net.train(1, args.batch, lr, seed=ep)[0]))
        log(f"[surrogate] epoch {ep + 1}/{args.epochs} loss {losses[-1]:.5f} ({time.time() - t1:.1f}s)")
    if not args.params:
        t_train = time.time() - t1
        np.save(os.path.join(args.out, "surrogate_params.npy"), net.params().astype(np.float16))
    ev_specs = S.sample_specs(args.eval, 5000)
    ev = S.evaluate(net, ev_specs, replicas=args.replicas, steps=args.eval_steps, window=args.window,
                    t_ss=(args.eval_steps, args.eval_steps + 3000), log=log)
    sc = S.steps_to(ev["cold_err"], args.window, 0.02)
    sw = S.steps_to(ev["warm_err"], args.window, 0.02)
    rep = {
        "train_fields": nfields, "params": args.params, "generate_s": t_gen, "train_s": t_train, "epochs": args.epochs,
        "losses": losses, "eval_geometries": args.eval, "replicas": args.replicas, "window": args.window,
        "cold_err": ev["cold_err"].tolist(), "warm_err": ev["warm_err"].tolist(),
        "cold_steps_to_2pct": sc, "warm_steps_to_2pct": sw,
        "cold_ms_per_step": ev["cold_ms"], "warm_ms_per_step": ev["warm_ms"], "inference_s": ev["infer_s"],
        "cold_wall_s_to_2pct": None if sc is None else sc * ev["cold_ms"] / 1000,
        "warm_wall_s_to_2pct": None if sw is None else sw * ev["warm_ms"] / 1000 + ev["infer_s"],
    }
    with open(os.path.join(args.out, "surrogate_report.json"), "w") as f:
        json.dump(rep, f, indent=1)
    log(json.dumps({k: v for k, v in rep.items() if not isinstance(v, list)}))
    log("cold err " + " ".join(f"{x:.3f}" for x in ev["cold_err"]))
    log("warm err " + " ".join(f"{x:.3f}" for x in ev["warm_err"]))


if __name__ == "__main__":
    main()
