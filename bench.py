#!/usr/bin/env python
"""Throughput of Richardson-Lucy light-field deconvolution on one MI355X.

Workload (default "rl1455"): Nnum=15, 97x97 lenslets (1455x1455 float32
image), 41 planes z=-20..20, seeded ring PSFs with odd sides growing from 15
to 345 px, 400 seeded Gaussian blobs on a 0.01 background, Poisson counts at
peak 2000, 8 plain RL iterations (Prevedel update) from a uniform start.

One step = one full reconstruction of one volume from its image with all
inputs resident in HBM: the back projection Htf = Ht im, then RL.  Two legs
reach the same data fit ||H X - im|| / ||im||:
  plain    8 Prevedel RL iterations from a uniform start (the reference's
           DeconvRL with a uniform Xguess); its final fit is the target;
  learned  the guarded learned initialiser (network guess, kept only if its
           fit beats one plain iteration), then RL until the fit is <= the
           target.
value = volumes/s of the learned leg (of the plain leg when no trained
weights are packaged, or with --plain-only).

Also reported:
  roofline      the multiply-accumulate kernel (dominant): algorithmic bytes
                (compulsory cached-spectrum reads, 8 B per stored complex bin)
                / measured kernel time (HIP events on its stream) vs 8 TB/s.
  rl_iteration  the same ratio for a whole RL iteration: S_H + S_Ht +
                B*(4*nz+2)*Y*X*4 bytes / iteration time.
  cpu_baseline  the FP64 numpy oracle (single thread) timed on one plane's
                projection at full size and scaled to a volume.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fast-light-field_amd"))

import numpy as np  # noqa: E402

from flfhip.workloads import WORKLOADS, setup  # noqa: E402

PEAK_HBM = 8.0e12  # MI355X HBM3E, bytes/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(w):
    """FP64 oracle (special-FFT restatement, numpy, one thread) on a bounded
    sample of the workload: the forward and the backward projection of the
    largest-PSF plane and the forward projection of the median plane at full
    size (about 10-20 s), scaled by Fourier bins to one plain reconstruction
    (Htf, then forward + backward per iteration, every plane)."""
    sys.path.insert(0, ROOT)
    from oracle import projector as oproj
    from oracle import shapes as oshapes
    hm, Y, X, N = w["hm"], w["Y"], w["X"], w["N"]
    bins = lambda s2: (lambda f: f[0] * (f[1] // 2 + 1))(oshapes.convolution_shape((Y, X), (s2, s2), N, "5smooth")[0])
    zmax = int(np.argmax(w["sides"]))
    zmed = int(np.argsort(w["sides"])[len(w["sides"]) // 2])
    sample = [(zmax, False), (zmax, True), (zmed, False)]
    t_total, b_total = 0.0, 0
    for z, back in sample:
        src = w["vol"][z][:, :Y, :X]            # (1, Y, X): a plane, or an image-sized source for Ht
        t = time.time()
        oproj.project_special(src, hm.Hcc(z, back), pad="5smooth")
        t_total += time.time() - t
        b_total += bins(w["sides"][z])
    per_bin = t_total / b_total
    all_bins = sum(bins(s2) for s2 in w["sides"])
    per_volume = per_bin * all_bins * (2 * w["iters"] + 1)
    return {"value": 1.0 / per_volume, "unit": "volumes/s", "cores": 1, "kind": "port",
            "sample": "FP64 numpy oracle (project_special), one thread: forward + backward projection of plane "
                      "z=%d (PSF %d px) and forward of plane z=%d (PSF %d px) at %dx%d, %.1f s, scaled by Fourier "
                      "bins to %d planes x (2 x %d iterations + 1) projections"
                      % (w["zs"][zmax], w["sides"][zmax], w["zs"][zmed], w["sides"][zmed], Y, X, t_total,
                         len(w["zs"]), w["iters"])}


def measured_traffic(workload):
    """HBM bytes per k_mac launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE runs of this bench (calibrated by a 1 GiB copy), or None."""
    path = os.path.join(ROOT, "profiles", "k_mac_%s_hbm_traffic.txt" % workload)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_per_launch_bytes"):
            return float(line.split()[1])
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="rl1455", choices=sorted(WORKLOADS))
    ap.add_argument("--weights", default=None, help="learned-initialiser weights (default: the packaged file)")
    ap.add_argument("--frames", type=int, default=1,
                    help="frames reconstructed together per step (value = frames x steps / time)")
    ap.add_argument("--plain-only", action="store_true", help="time plain RL only")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="also run one device copy of this many MiB (reference kernel for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (frames or planes would be "
              "sharded across GPUs with one all-reduce of the forward image per iteration; see DESIGN.md).",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from flfhip.lfdeconv import deconv_rl_device, _fit_from_sums
    from flfhip.projector import Projector_hip
    from flfhip import lensletnet, synthetic

    T = args.frames
    w = setup(args.workload, frames=T)
    hm, Y, X, iters = w["hm"], w["Y"], w["X"], w["iters"]
    nz = len(w["zs"])
    proj = Projector_hip()
    torch.cuda.synchronize()
    t = time.time()
    # room for two frames per call: the learned leg projects its guess and the
    # first plain iterate together
    proj.prepare(hm, T, Y, X, frames_capacity=2 * T)
    torch.cuda.synchronize()
    t_prep = time.time() - t
    log("spectra precompute: %.2fs, device bytes %.2f GB" % (t_prep, proj.lib.flf_device_bytes(proj._ctx) / 1e9))
    tuned = None
    if not args.no_autotune:
        from flfhip import autotune
        t = time.time()
        tuned, _ = autotune.tune(proj, hm, T, Y, X)
        log("autotune %.1fs: %s" % (time.time() - t, tuned))

    vol = proj.asnative(w["vol"])
    clean = proj.asnumpy(proj.forward_device(hm, vol))
    im = proj.asnative(np.concatenate([synthetic.poisson_image(clean[t:t + 1], w["peak"], w["seed"] + t)
                                       for t in range(T)]))
    del vol, clean
    Htf = torch.empty((nz, T, Y, X), dtype=torch.float32, device=proj.device)
    Xv = torch.empty_like(Htf)

    # One step = one reconstruction from the image: Htf = Ht im, then RL.
    def plain_step():
        proj.backward_device(hm, im, out=Htf)
        Xv.fill_(1.0)
        deconv_rl_device(proj, hm, Htf, iters, Xv)

    if args.calib_copy_mib:
        n = args.calib_copy_mib * (1 << 20) // 4
        a_ = torch.ones(n, dtype=torch.float32, device=proj.device)
        b_ = torch.empty_like(a_)
        b_.copy_(a_)
        torch.cuda.synchronize()
        del a_, b_

    s = torch.cuda.current_stream()

    def timed(step, n):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(s)
        for _ in range(n):
            step()
        e1.record(s)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / n

    # plain RL: the reference's DeconvRL from a uniform Xguess, fixed iterations
    for _ in range(args.warmup):
        plain_step()
    proj.profile(True)
    plain_ms = timed(plain_step, args.steps)
    proj.profile(False)
    mac_ms, mac_n, mac_bytes = proj.profile_read()
    fit_plain_t = _fit_from_sums(proj.datafit_device(proj.forward_device(hm, Xv), im))
    fit_plain = float(fit_plain_t[0]) if T == 1 else [float(f) for f in fit_plain_t]
    X_plain = Xv.clone()
    log("plain RL: %.2f ms per step of %d frame(s), data fit %s after %d iterations"
        % (plain_ms, T, fit_plain, iters))

    learned = None
    wpath = args.weights or lensletnet.default_weights_path(w["N"], nz, args.workload)
    if not args.plain_only and not os.path.exists(wpath):
        log("no learned-initialiser weights at %s: timing plain RL only" % wpath)
    if not args.plain_only and os.path.exists(wpath):
        net = lensletnet.LensletNet.load(wpath)
        li = lensletnet.LearnedInit(net, proj, hm)
        li.plain_denominator(T, Y, X)          # cached per PSF set and frame shape, like the spectra
        out = {}

        def learned_step():
            proj.backward_device(hm, im, out=Htf)
            Xl, n_it, info = li.deconvolve(im, Htf, 4 * iters, targetFit=fit_plain_t)
            out.update(X=Xl, n=n_it, info=info)

        for _ in range(args.warmup):
            learned_step()
        proj.profile(True)
        learned_ms = timed(learned_step, args.steps)
        proj.profile(False)
        lm_ms, lm_n, lm_bytes = proj.profile_read()
        mac_ms, mac_n, mac_bytes = mac_ms + lm_ms, mac_n + lm_n, mac_bytes + lm_bytes
        info = out["info"]
        fit_l_t = _fit_from_sums(proj.datafit_device(proj.forward_device(hm, out["X"]), im))
        rel = float(torch.linalg.vector_norm((out["X"] - X_plain).double()) / torch.linalg.vector_norm(X_plain.double()))
        one = lambda a, f=float: f(a[0]) if T == 1 else [f(x) for x in a]
        learned = {"ms": learned_ms, "iterations": int(out["n"]), "guess_accepted": one(info["accepted"], bool),
                   "fit_guess": one(info["fit_guess"]), "fit_plain1": one(info["fit_plain1"]),
                   "datafit_reached": one(fit_l_t), "rel_l2_vs_plain": rel, "weights": os.path.relpath(wpath, ROOT)}
        log("learned init: %.2f ms per step, %d iterations, fit %s (guess %s vs one plain iteration %s)"
            % (learned_ms, out["n"], learned["datafit_reached"], learned["fit_guess"], learned["fit_plain1"]))

    ms_step = learned["ms"] if learned else plain_ms
    # Whole-iteration roofline: S_H + S_Ht + B*(4*nz+2)*Y*X*4 bytes.
    s_dir = 0.0
    for z in range(nz):
        fy, fx = proj.fshape(z)
        u = 36 if fy == fx else 64
        s_dir += u * fy * (fx // 2 + 1) * 8
    it_bytes = 2 * s_dir + T * (4 * nz + 2) * Y * X * 4
    it_s = (plain_ms / 1e3) / (iters + 0.5)      # Htf is half an iteration (all frames of a step)
    mac_avg_ms = mac_ms / max(mac_n, 1)
    achieved = (mac_bytes / max(mac_n, 1)) / (mac_avg_ms / 1e3)
    desc = ("Nnum=15, %dx%d px, %d planes (z=-%d..%d), PSF %d..%d px, %d blobs, Poisson peak %g, %d frame%s per "
            "step; target = data fit of %d plain RL iterations from a uniform start"
            % (Y, X, nz, w["zs"][-1], w["zs"][-1], min(w["sides"]), max(w["sides"]), w["nblobs"], w["peak"], T,
               "" if T == 1 else "s", iters))
    res = {
        "metric": "volumes/s at fixed RL data-fit: learned-init vs plain RL, 1 GPU",
        "value": T * 1e3 / ms_step,
        "unit": "volumes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": desc, "init": "learned (guarded)" if learned else "uniform (plain RL)",
                   "target_datafit": fit_plain,
                   "frames": T, "launch_options": tuned,
                   "plain": {"ms": plain_ms, "volumes_per_s": T * 1e3 / plain_ms, "iterations": iters},
                   "learned": learned,
                   "speedup_vs_plain": (plain_ms / learned["ms"]) if learned else None},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM, "traffic": measured_traffic(args.workload),
                     "kernel": "k_mac", "launch_avg_ms": mac_avg_ms, "launches": mac_n,
                     "bytes_per_launch": mac_bytes / max(mac_n, 1)},
        "rl_iteration": {"ms": it_s * 1e3, "algorithmic_bytes": it_bytes,
                         "achieved_GBps": it_bytes / it_s / 1e9, "frac": it_bytes / it_s / PEAK_HBM},
        "spectra_precompute_s": t_prep,
    }
    if not args.no_cpu_baseline:
        t = time.time()
        res["cpu_baseline"] = cpu_baseline(w)
        log("cpu baseline %.1fs" % (time.time() - t))
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
