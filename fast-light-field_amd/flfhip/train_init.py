"""Train the learned initialiser for one PSF set on the GPU.

    python -m flfhip.train_init --workload rl1455 --pairs 6144 --steps 16000
    python -m flfhip.train_init --workload ragged64 --pairs 20000 --rl-iters 8 --seed 4 \
        --test-seed 5 --heldout 512 --gen-batch 64 --steps 20000

(the second: 600 x 600 px frames of the 31-plane PSF set shared by the
ragged time series, 20k training pairs with 8-iteration targets, evaluated
on 512 held-out frames).

Pairs are made by the method itself (lensletnet.make_pairs): seeded blob
volumes at L x L lenslets are forward-projected through the workload's PSF
set, Poisson noise is drawn at a peak in [500, 5000] counts, and plain RL from
a uniform start gives the target (its lenslet-block means) after --rl-iters
iterations.  With --heldout, the held-out frames (--test-seed) are also
reconstructed both ways (lensletnet.evaluate_heldout): plain RL for 8
iterations, and the guarded learned start followed by RL to the same data
fit; the summary (guard acceptance, iterations to the target, error
against the ground-truth volumes) is printed and saved with the weights.  Training runs the network's own
BF16 MFMA forward/backward kernels with Adam (warm-up, cosine decay, random
dihedral transforms).  The weights (with the training settings and losses)
go to flfhip/weights/ by default.
"""
import argparse
import os
import sys
import time

import numpy as np


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="rl1455", help="PSF set: a workload or a ragged time series")
    ap.add_argument("--pairs", type=int, default=6144)
    ap.add_argument("--val-pairs", type=int, default=64)
    ap.add_argument("--lenslets", type=int, default=40, help="training frames are L x L lenslets")
    ap.add_argument("--rl-iters", type=int, default=16, help="plain RL iterations that make a target")
    ap.add_argument("--steps", type=int, default=16000)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=3e-4)
    ap.add_argument("--width", type=int, default=128)
    ap.add_argument("--blocks", type=int, default=4)
    ap.add_argument("--seed", type=int, default=1000, help="seed of the training pairs")
    ap.add_argument("--test-seed", type=int, default=None, help="seed of the held-out frames (default seed + 7919)")
    ap.add_argument("--heldout", type=int, default=0, help="held-out frames to evaluate RL from the guess on")
    ap.add_argument("--gen-batch", type=int, default=16, help="frames per projector call when making pairs")
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)

    import torch
    import json
    from .lensletnet import LensletNet, default_weights_path, evaluate_heldout, make_pairs, train
    from .projector import Projector_hip
    from .workloads import psf_set

    def log(*a):
        print(*a, flush=True)

    t0 = time.time()
    hm, zs, sides = psf_set(args.workload)
    proj = Projector_hip()
    t = time.time()
    test_seed = args.seed + 7919 if args.test_seed is None else args.test_seed
    imgs, tgts = make_pairs(proj, hm, args.pairs, args.lenslets, seed=args.seed, iters=args.rl_iters,
                            batch=args.gen_batch, log=log)
    vimgs, vtgts = make_pairs(proj, hm, args.val_pairs, args.lenslets, seed=test_seed, iters=args.rl_iters,
                              batch=args.gen_batch)
    torch.cuda.synchronize()
    t_pairs = time.time() - t
    log("pairs: %d + %d held out, %.1fs" % (args.pairs, args.val_pairs, t_pairs))
    net = LensletNet(hm.Nnum, hm.numZ, args.width, args.blocks, seed=0)
    t = time.time()
    hist = train(net, imgs, tgts, args.steps, batch=args.batch, lr=args.lr, seed=args.seed, log=log, log_every=500)
    torch.cuda.synchronize()
    t_train = time.time() - t
    vl = [net.train_step(vimgs[b:b + args.batch], vtgts[b:b + args.batch], 0.0)
          for b in range(0, args.val_pairs, args.batch)]
    log("train %.1fs (%.1f ms/step); loss last-50 %.5f, held-out %.5f"
        % (t_train, 1e3 * t_train / args.steps, float(np.mean(hist[-50:])), float(np.mean(vl))))
    evaluation = {}
    if args.heldout:
        t = time.time()
        _, evaluation = evaluate_heldout(net, proj, hm, args.heldout, args.lenslets, test_seed, iters=8,
                                         batch=args.gen_batch)
        log("held-out evaluation (%.1fs): %s" % (time.time() - t, json.dumps(evaluation)))
    out = args.out or default_weights_path(hm.Nnum, hm.numZ, args.workload)
    os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
    net.save(out, workload=args.workload, pairs=args.pairs, lenslets=args.lenslets, rl_iters=args.rl_iters,
             steps=args.steps, batch=args.batch, lr=args.lr, seed=args.seed, train_loss=float(np.mean(hist[-50:])),
             heldout_loss=float(np.mean(vl)), pair_seconds=t_pairs, train_seconds=t_train, test_seed=test_seed,
             gen_batch=args.gen_batch, evaluation=json.dumps(evaluation))
    log("saved %s (%.1fs total)" % (out, time.time() - t0))
    return 0


if __name__ == "__main__":
    sys.exit(main())
