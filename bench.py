#!/usr/bin/env python
"""Throughput of Richardson-Lucy light-field deconvolution on one MI355X.

Workload (default "rl1455"): Nnum=15, 97x97 lenslets (1455x1455 float32
image), 41 planes z=-20..20, seeded ring PSFs with odd sides growing from 15
to 345 px, 400 seeded Gaussian blobs on a 0.01 background, Poisson counts at
peak 2000, 8 plain RL iterations (Prevedel update) from a uniform start.

One step = one full reconstruction of one volume (8 RL iterations, i.e. 8
forward + 8 fused backward/update projections) with all inputs resident in
HBM.  value = volumes/s.  The 8-iteration data fit is the fixed target, so
plain RL reaches it in exactly 8 iterations.

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

PEAK_HBM = 8.0e12  # MI355X HBM3E, bytes/s

WORKLOADS = {
    # name: (Nnum, lenslets, zmax, s_min, s_max, blobs, peak, iters, seed)
    "rl1455": (15, 97, 20, 15, 345, 400, 2000.0, 8, 2),
    "rl300": (15, 20, 5, 15, 135, 200, 1000.0, 4, 1),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def setup(name):
    from flfhip import synthetic
    N, L, zmax, smin, smax, nblobs, peak, iters, seed = WORKLOADS[name]
    zs = list(range(-zmax, zmax + 1))
    sides = synthetic.psf_sides(zs, smin, smax)
    t = time.time()
    hm = synthetic.make_psf_set(N, zs, sides, seed=seed)
    log("psf set: %d planes, sides %d..%d, %.1fs" % (len(zs), min(sides), max(sides), time.time() - t))
    Y = X = N * L
    t = time.time()
    vol = synthetic.blob_volume(len(zs), Y, X, nblobs, seed=seed)
    log("volume: %s, %.1fs" % (vol.shape, time.time() - t))
    return dict(N=N, Y=Y, X=X, zs=zs, sides=sides, hm=hm, vol=vol, peak=peak, iters=iters, seed=seed)


def cpu_baseline(w, target_s=15.0):
    """FP64 oracle (special-FFT restatement, numpy, one thread) on one plane of
    the workload: forward projection of the largest-PSF plane, scaled to a
    full reconstruction by the Fourier-bin count of every plane."""
    sys.path.insert(0, ROOT)
    from oracle import projector as oproj
    from oracle import shapes as oshapes
    hm, Y, X, N = w["hm"], w["Y"], w["X"], w["N"]
    z = int(np.argmax(w["sides"]))
    src = w["vol"][z][:, :Y, :X]
    t = time.time()
    oproj.project_special(src, hm.Hcc(z, False), pad="5smooth")
    dt = time.time() - t
    f_s = oshapes.convolution_shape((Y, X), (w["sides"][z],) * 2, N, "5smooth")[0]
    bins = lambda f: f[0] * (f[1] // 2 + 1)
    total = 0.0
    for s2 in w["sides"]:
        f = oshapes.convolution_shape((Y, X), (s2, s2), N, "5smooth")[0]
        total += dt * bins(f) / bins(f_s)
    per_volume = total * 2 * w["iters"]          # forward + backward per iteration
    return {"value": 1.0 / per_volume, "unit": "volumes/s", "cores": 1, "kind": "port",
            "sample": "FP64 numpy oracle (project_special), forward projection of plane z=%d (s2=%d) at %dx%d, "
                      "%.1f s, scaled by Fourier bins to %d planes x 2 directions x %d iterations"
                      % (w["zs"][z], w["sides"][z], Y, X, dt, len(w["zs"]), w["iters"])}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="rl1455", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
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

    w = setup(args.workload)
    hm, Y, X, iters = w["hm"], w["Y"], w["X"], w["iters"]
    nz = len(w["zs"])
    proj = Projector_hip()
    torch.cuda.synchronize()
    t = time.time()
    proj.prepare(hm, 1, Y, X)
    torch.cuda.synchronize()
    t_prep = time.time() - t
    log("spectra precompute: %.2fs, device bytes %.2f GB" % (t_prep, proj.lib.flf_device_bytes(proj._ctx) / 1e9))

    vol = proj.asnative(w["vol"])
    clean = proj.forward_device(hm, vol)
    from flfhip import synthetic
    im = proj.asnative(synthetic.poisson_image(proj.asnumpy(clean), w["peak"], w["seed"]))
    del vol, clean
    Htf = proj.backward_device(hm, im)
    Xv = torch.empty_like(Htf)

    def step():
        Xv.fill_(1.0)
        deconv_rl_device(proj, hm, Htf, iters, Xv, im=im)

    if args.calib_copy_mib:
        n = args.calib_copy_mib * (1 << 20) // 4
        a_ = torch.ones(n, dtype=torch.float32, device=proj.device)
        b_ = torch.empty_like(a_)
        b_.copy_(a_)
        torch.cuda.synchronize()
        del a_, b_
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    proj.profile(True)
    s = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.time()
    e0.record(s)
    for _ in range(args.steps):
        step()
    e1.record(s)
    torch.cuda.synchronize()
    wall = time.time() - t0
    proj.profile(False)
    mac_ms, mac_n, mac_bytes = proj.profile_read()
    ms_step = e0.elapsed_time(e1) / args.steps
    log("wall %.3fs for %d steps; event %.2f ms/step" % (wall, args.steps, ms_step))

    # Data fit reached (the fixed target of the metric).
    fit = _fit_from_sums(proj.datafit_device(proj.forward_device(hm, Xv), im))
    # Whole-iteration roofline: S_H + S_Ht + B*(4*nz+2)*Y*X*4 bytes.
    s_dir = 0.0
    for z in range(nz):
        fy, fx = proj.fshape(z)
        u = 36 if fy == fx else 64
        s_dir += u * fy * (fx // 2 + 1) * 8
    it_bytes = 2 * s_dir + 1 * (4 * nz + 2) * Y * X * 4
    it_s = ms_step / 1e3 / iters
    mac_avg_ms = mac_ms / max(mac_n, 1)
    achieved = (mac_bytes / max(mac_n, 1)) / (mac_avg_ms / 1e3)
    res = {
        "metric": "volumes/s at fixed RL data-fit: learned-init vs plain RL, 1 GPU",
        "value": 1e3 / ms_step,
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
        "config": {"workload": "plain RL, Nnum=15, %dx%d px, %d planes (z=-%d..%d), PSF %d..%d px, %d blobs, "
                               "Poisson peak %g, %d iterations from uniform start, 1 frame"
                               % (Y, X, nz, w["zs"][-1], w["zs"][-1], min(w["sides"]), max(w["sides"]),
                                  WORKLOADS[args.workload][5], w["peak"], iters),
                   "init": "uniform (plain RL)", "iterations": iters, "datafit_reached": float(fit[0])},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM, "traffic": None,
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
