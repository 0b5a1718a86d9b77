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
  rl_iteration  the same ratio for a whole RL iteration of one frame:
                S_H + S_Ht + (4*nz+2)*Y*X*4 bytes / iteration time
                (rl_iteration_batch: the same for all B frames of a step,
                where the MAC is FP32-bound instead).
  cpu_baseline  the FP64 numpy oracle (single thread) timed on one plane's
                projection at full size and scaled to a volume.

--workload ragged64 measures the time series of uneven regions instead
(64 frames of 300..900 px square, 31 planes, PSF 15..255 px, 8 Prevedel RL
iterations from Htf per frame, all frames of a canvas bucket in one call per
kernel): value = frames reconstructed per second.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fast-light-field_amd"))

import numpy as np  # noqa: E402

from flfhip.workloads import BATCHES, RAGGED, WORKLOADS, setup  # noqa: E402

PEAK_HBM = 8.0e12  # MI355X HBM3E, bytes/s
PEAK_FP32 = 157.3e12  # MI355X FP32, flop/s


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
                      "z=%g (PSF %d px) and forward of plane z=%g (PSF %d px) at %dx%d, %.1f s, scaled by Fourier "
                      "bins to %d planes x (2 x %d iterations + 1) projections"
                      % (round(w["zs"][zmax], 3), w["sides"][zmax], round(w["zs"][zmed], 3), w["sides"][zmed], Y, X,
                         t_total,
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


def ragged_cpu_baseline(w):
    """FP64 oracle (one thread) on the median-size frame: forward and backward
    projection of the largest-PSF plane and forward of the median plane,
    scaled by Fourier bins (each frame planned at its own size, as the
    reference does) to every frame's Htf + 2 projections per iteration."""
    sys.path.insert(0, ROOT)
    from oracle import projector as oproj
    from oracle import shapes as oshapes
    hm, N, shapes = w["hm"], w["N"], w["shapes"]
    bins = lambda sh, s2: (lambda f: f[0] * (f[1] // 2 + 1))(oshapes.convolution_shape(sh, (s2, s2), N, "5smooth")[0])
    tmed = int(np.argsort([a * b for a, b in shapes])[len(shapes) // 2])
    sh = shapes[tmed]
    zmax = int(np.argmax(w["sides"]))
    zmed = int(np.argsort(w["sides"])[len(w["sides"]) // 2])
    t_total, b_total = 0.0, 0
    for z, back in [(zmax, False), (zmax, True), (zmed, False)]:
        src = w["vols"][tmed][z][None]
        t = time.time()
        oproj.project_special(src, hm.Hcc(z, back), pad="5smooth")
        t_total += time.time() - t
        b_total += bins(sh, w["sides"][z])
    per_bin = t_total / b_total
    all_bins = sum(bins(s, s2) for s in shapes for s2 in w["sides"])
    per_series = per_bin * all_bins * (2 * w["iters"] + 1)
    return {"value": len(shapes) / per_series, "unit": "volumes/s", "cores": 1, "kind": "port",
            "sample": "FP64 numpy oracle (project_special), one thread: forward + backward projection of plane z=%d "
                      "(PSF %d px) and forward of plane z=%d (PSF %d px) of the median frame (%dx%d), %.1f s, "
                      "scaled by Fourier bins to %d frames x %d planes x (2 x %d iterations + 1) projections"
                      % (w["zs"][zmax], w["sides"][zmax], w["zs"][zmed], w["sides"][zmed], sh[0], sh[1], t_total,
                         len(shapes), len(w["zs"]), w["iters"])}


def ragged_main(args):
    import torch
    from flfhip import autotune, synthetic
    from flfhip.lfdeconv import _fit_from_sums
    from flfhip.ragged import RaggedBatch
    from flfhip.workloads import setup_ragged

    w = setup_ragged(args.workload)
    hm, shapes, iters, N = w["hm"], w["shapes"], w["iters"], w["N"]
    nz, T = len(w["zs"]), len(shapes)
    rb = RaggedBatch(hm, shapes, nbuckets=args.buckets)
    log("buckets: %s" % [(c, len(i)) for c, i in rb.buckets])
    torch.cuda.synchronize()
    t = time.time()
    for k, ((yc, xc), idx) in enumerate(rb.buckets):
        rb.projectors[k].prepare(hm, len(idx), yc, xc)
    torch.cuda.synchronize()
    t_prep = time.time() - t
    dev_bytes = sum(p.lib.flf_device_bytes(p._ctx) for p in rb.projectors)
    log("spectra precompute: %.2fs, device bytes %.2f GB" % (t_prep, dev_bytes / 1e9))
    tuned = []
    for k, ((yc, xc), idx) in enumerate(rb.buckets):
        if args.no_autotune:
            tuned.append(None)
            continue
        t = time.time()
        opt, _ = autotune.tune(rb.projectors[k], hm, len(idx), yc, xc, reps=1)
        tuned.append(opt)
        log("autotune bucket %d (%dx%d, %d frames) %.1fs: %s" % (k, yc, xc, len(idx), time.time() - t, opt))
    clean = rb.forward(w["vols"])
    ims = [synthetic.poisson_image(c[None], w["peak"], w["seed"] + t)[0] for t, c in enumerate(clean)]
    del clean
    staged = rb.stage(ims)
    bufs = [(torch.empty((nz, len(idx), yc, xc), dtype=torch.float32, device=rb.projectors[k].device),
             torch.empty((nz, len(idx), yc, xc), dtype=torch.float32, device=rb.projectors[k].device))
            for k, ((yc, xc), idx) in enumerate(rb.buckets)]

    def step():
        rb.deconv_staged(staged, iters, out=bufs, graph=args.graph)

    for _ in range(args.warmup):
        step()
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for p in rb.projectors:
        p.profile(True)
    torch.cuda.synchronize()
    e0.record(s)
    for _ in range(args.steps):
        step()
    e1.record(s)
    torch.cuda.synchronize()
    ms_step = e0.elapsed_time(e1) / args.steps
    for p in rb.projectors:
        p.profile(False)
    prof_steps = args.steps
    if args.graph:
        # graph replays record no per-launch events: take the kernel figures
        # from one eager step
        for p in rb.projectors:
            p.profile_read()
            p.profile(True)
        rb.deconv_staged(staged, iters, out=bufs, graph=False)
        torch.cuda.synchronize()
        for p in rb.projectors:
            p.profile(False)
        prof_steps = 1
    mac_ms = mac_n = mac_bytes = 0
    for p in rb.projectors:
        a, b, c = p.profile_read()
        mac_ms, mac_n, mac_bytes = mac_ms + a, mac_n + b, mac_bytes + c
    # data fit of every frame's reconstruction
    fits = np.zeros(T)
    for k, ((yc, xc), idx) in enumerate(rb.buckets):
        with rb._bucket(k) as (p, _, _, _):
            f = _fit_from_sums(p.datafit_device(p.forward_device(hm, bufs[k][1]), staged[k]))
        fits[idx] = f
    # algorithmic flops of the multiply-accumulate: 8 N^2 per half-plane bin,
    # plane and frame, per projection (Htf + 2 per iteration)
    mac_flops = 0.0
    for k, ((yc, xc), idx) in enumerate(rb.buckets):
        p = rb.projectors[k]
        per = sum((lambda f: f[0] * (f[1] // 2 + 1))(p.fshape(z)) for z in range(nz))
        mac_flops += 8.0 * N * N * per * len(idx) * (2 * iters + 1)
    mac_flops *= prof_steps
    mac_avg_ms = mac_ms / max(mac_n, 1)
    achieved = (mac_bytes / max(mac_n, 1)) / (mac_avg_ms / 1e3)
    flop_rate = mac_flops / (mac_ms / 1e3)
    desc = ("ragged time series: %d frames, Nnum=%d, lenslet grids %d..%d square (%d..%d px), %d planes "
            "(z=-%d..%d), PSF %d..%d px, %d..%d drifting blobs per frame, Poisson peak %g, %d Prevedel RL "
            "iterations from Htf per frame; %d canvas bucket(s) %s"
            % (T, N, min(s[0] for s in shapes) // N, max(s[0] for s in shapes) // N, min(s[0] for s in shapes),
               max(s[0] for s in shapes), nz, w["zs"][-1], w["zs"][-1], min(w["sides"]), max(w["sides"]),
               min(w["nblobs"]), max(w["nblobs"]), w["peak"], iters, len(rb.buckets),
               [(c, len(i)) for c, i in rb.buckets]))
    res = {
        "metric": "volumes/s, ragged time series (%d uneven frames, %d RL iterations), 1 GPU" % (T, iters),
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
        "config": {"workload": desc, "frames": T, "graph": bool(args.graph), "launch_options": tuned,
                   "datafit": {"min": float(fits.min()), "median": float(np.median(fits)),
                               "max": float(fits.max())}},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM, "traffic": None, "kernel": "k_mac",
                     "launch_avg_ms": mac_avg_ms, "launches": mac_n, "bytes_per_launch": mac_bytes / max(mac_n, 1)},
        "valu_fp32": {"kernel": "k_mac", "achieved_TFLOPs": flop_rate / 1e12, "peak_TFLOPs": PEAK_FP32 / 1e12,
                      "frac": flop_rate / PEAK_FP32, "mac_share_of_step": mac_ms / prof_steps / ms_step},
        "spectra_precompute_s": t_prep,
        "device_bytes": dev_bytes,
    }
    if not args.no_cpu_baseline:
        t = time.time()
        res["cpu_baseline"] = ragged_cpu_baseline(w)
        log("cpu baseline %.1fs" % (time.time() - t))
    print(json.dumps(res), flush=True)

def _rel_to_truth(X, vol, scale):
    """Per-frame relative L2 of a reconstruction against the ground-truth
    volume: X is in image counts, vol in volume units; scale[t] converts
    (the image is Poisson(scale[t] * H vol_t))."""
    import torch
    out = []
    for t in range(X.shape[1]):
        a = X[:, t].double() / float(scale[t])
        b = vol[:, t].double()
        out.append(float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b)))
    return out


def _iters_to_target(fits, target):
    """First iteration count whose fit is <= the frame's target, from a list
    of per-iteration fit arrays (entry 0: the starting volume)."""
    F = np.array([np.asarray(f, np.float64) for f in fits])
    out = []
    for t in range(F.shape[1]):
        k = np.nonzero(F[:, t] <= target[t])[0]
        out.append(int(k[0]) if len(k) else -1)
    return out


def oracle_check(proj, hm, im_t, X0_t, iters=1, planes=(0, 20), start="the learned leg's starting volume"):
    """Part of the CPU leg (runs only with it, after every timed region).
    Bounded FP64-oracle comparison of one frame: Prevedel RL restricted to
    two planes (the largest-PSF plane and z = 0) from the given start, on the
    GPU and in the FP64 oracle (numpy, one thread); relative L2 per plane."""
    import torch
    sys.path.insert(0, ROOT)
    from oracle import projector as oproj
    from oracle import rl as orl
    from flfhip.lfdeconv import deconv_rl_device
    planes = list(planes)
    Y, X = im_t.shape[-2:]
    proj.prepare(hm, 1, Y, X)
    im1 = im_t.reshape(1, Y, X).contiguous()
    Htf = proj.backward_device(hm, im1, planes=planes)
    Xv = torch.zeros((hm.numZ, 1, Y, X), dtype=torch.float32, device=proj.device)
    Xv[planes] = X0_t[planes].reshape(len(planes), 1, Y, X)
    deconv_rl_device(proj, hm, Htf, iters, Xv, planes=planes)
    Qf = [hm.Hcc(z, False) for z in planes]
    Qb = [hm.Hcc(z, True) for z in planes]
    t = time.time()
    Xr = orl.deconv_rl(lambda v: oproj.forward_project(v, Qf), lambda i: oproj.backward_project(i, Qb), None, iters,
                       Xguess=X0_t[planes].reshape(len(planes), 1, Y, X).cpu().numpy().astype(np.float64),
                       im=im1.cpu().numpy().astype(np.float64))
    secs = time.time() - t
    Xg = Xv.cpu().numpy()
    errs = [float(np.linalg.norm(Xg[z, 0] - Xr[k, 0]) / np.linalg.norm(Xr[k, 0])) for k, z in enumerate(planes)]
    return {"frame": 0, "planes": [int(hm.PSFShape(z)[0]) for z in planes], "plane_ids": planes,
            "iterations": iters, "start": start, "rel_l2_per_plane": errs,
            "oracle_s": secs,
            "what": "Prevedel RL of frame 0 restricted to planes %s (PSF sides %s px) on the GPU vs the FP64 numpy "
                    "oracle from the same start" % (planes, [int(hm.PSFShape(z)[0]) for z in planes])}


def batch_main(args):
    """Batches of full-size frames (workloads.BATCHES): plain RL against the
    guarded learned start on T frames reconstructed together."""
    import torch
    from flfhip import lensletnet
    from flfhip.lfdeconv import deconv_rl_device, _fit_from_sums
    from flfhip.projector import Projector_hip
    from flfhip.workloads import setup_batch

    w = setup_batch(args.workload)
    hm, Y, X, iters, T = w["hm"], w["Y"], w["X"], w["iters"], w["frames"]
    if args.frames > 1:
        T = args.frames
    nz = len(w["zs"])
    proj = Projector_hip()
    torch.cuda.synchronize()
    t = time.time()
    proj.prepare(hm, T, Y, X, frames_capacity=2 * T)
    torch.cuda.synchronize()
    t_prep = time.time() - t
    log("spectra precompute: %.2fs, device bytes %.2f GB" % (t_prep, proj.lib.flf_device_bytes(proj._ctx) / 1e9))
    tuned = None
    if args.options:
        tuned = {}
        for kv in args.options.split(","):
            k, v = kv.split("=")
            proj.set_option(k, int(v))
            tuned[k] = int(v)
        # a plan option re-plans: keep the room for two frames per call
        proj.prepare(hm, T, Y, X, frames_capacity=2 * T)
        log("launch options (given): %s" % tuned)
    elif not args.no_autotune:
        from flfhip import autotune
        t = time.time()
        tuned, _ = autotune.tune(proj, hm, T, Y, X)
        log("autotune %.1fs: %s" % (time.time() - t, tuned))
    dev = proj.device
    if args.calib_copy_mib:
        # one device copy of known size: the reference kernel for the HBM counters
        n = args.calib_copy_mib * (1 << 20) // 4
        a_ = torch.ones(n, dtype=torch.float32, device=dev)
        b_ = torch.empty_like(a_)
        b_.copy_(a_)
        torch.cuda.synchronize()
        del a_, b_
    t = time.time()
    vol = lensletnet.blob_volumes_gpu(nz, Y, X, T, w["nblobs"], w["seed"], dev)
    proj.prepare(hm, T, Y, X)
    clean = proj.forward_device(hm, vol)
    scale = w["peak"] / clean.reshape(T, -1).amax(dim=1)
    gen = torch.Generator(device=dev)
    gen.manual_seed(w["seed"])
    im = torch.poisson(clean.clamp_min(0) * scale[:, None, None], generator=gen)
    del clean
    torch.cuda.synchronize()
    log("volumes and images: %d frames, %.1fs" % (T, time.time() - t))
    Htf = torch.empty((nz, T, Y, X), dtype=torch.float32, device=dev)
    Xv = torch.empty_like(Htf)
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

    # plain leg: Htf = Ht im, then iters Prevedel iterations from X = 1, all frames together
    def plain_step():
        proj.backward_device(hm, im, out=Htf)
        Xv.fill_(1.0)
        deconv_rl_device(proj, hm, Htf, iters, Xv)

    for _ in range(args.warmup):
        plain_step()
    plain_ms = timed(plain_step, args.steps)
    fit_plain_t = _fit_from_sums(proj.datafit_device(proj.forward_device(hm, Xv), im))
    gt_plain = _rel_to_truth(Xv, vol, scale)
    X_plain = Xv.clone()
    log("plain RL: %.1f ms per step of %d frames (%.1f volumes/s)" % (plain_ms, T, T * 1e3 / plain_ms))

    wpath = args.weights or lensletnet.default_weights_path(w["N"], nz, w["base"])
    net = lensletnet.LensletNet.load(wpath)
    li = lensletnet.LearnedInit(net, proj, hm)
    proj.prepare(hm, T, Y, X)
    li.plain_denominator(Y, X)
    out = {}

    def learned_step():
        fits = []
        Xl, n_it, info = li.deconvolve(im, None, 4 * iters, targetFit=fit_plain_t, Htf_out=Htf, fitOut=fits,
                                       minIter=args.min_iter)
        out.update(X=Xl, n=n_it, info=info, fits=fits)

    for _ in range(args.warmup):
        learned_step()
    proj.profile(True)
    learned_ms = timed(learned_step, args.steps)
    proj.profile(False)
    mac_ms, mac_n, mac_bytes = proj.profile_read()
    info = out["info"]
    fit_l_t = _fit_from_sums(proj.datafit_device(proj.forward_device(hm, out["X"]), im))
    gt_learned = _rel_to_truth(out["X"], vol, scale)
    rel_vs_plain = [float(torch.linalg.vector_norm((out["X"][:, t] - X_plain[:, t]).double())
                          / torch.linalg.vector_norm(X_plain[:, t].double())) for t in range(T)]
    need = _iters_to_target(out["fits"], fit_plain_t)
    log("learned: %.1f ms per step (%.1f volumes/s), %d iterations" % (learned_ms, T * 1e3 / learned_ms, out["n"]))

    # the same legs on one frame (frame 0), for the single-frame figures
    X1s = out["X"][:, :1].clone()
    proj.prepare(hm, 1, Y, X)
    im0 = im[:1].contiguous()
    Htf0 = torch.empty((nz, 1, Y, X), dtype=torch.float32, device=dev)
    X0v = torch.empty_like(Htf0)

    def plain1():
        proj.backward_device(hm, im0, out=Htf0)
        X0v.fill_(1.0)
        deconv_rl_device(proj, hm, Htf0, iters, X0v)

    plain1()
    plain1_ms = timed(plain1, args.steps)

    def learned1():
        li.deconvolve(im0, None, 4 * iters, targetFit=fit_plain_t[:1], Htf_out=Htf0, minIter=args.min_iter)

    learned1()
    learned1_ms = timed(learned1, args.steps)
    log("one frame: plain %.2f ms, learned %.2f ms" % (plain1_ms, learned1_ms))

    # algorithmic bytes and flops of the multiply-accumulate launches
    s_dir = 0.0
    bins = 0
    for z in range(nz):
        fy, fx = proj.fshape(z)
        u = 36 if fy == fx else 64
        s_dir += u * fy * (fx // 2 + 1) * 8
        bins += fy * (fx // 2 + 1)
    mac_avg_ms = mac_ms / max(mac_n, 1)
    # flops of the learned step's MAC launches: Htf (T frames), the guard's
    # pair (2T frames) and one forward + backward (2T frames) per iteration
    nproj_frames = T + 2 * T + 2 * T * out["n"]
    mac_flops = 8.0 * hm.Nnum ** 2 * bins * nproj_frames * args.steps
    flop_rate = mac_flops / (mac_ms / 1e3)
    hbm_rate = (mac_bytes / max(mac_n, 1)) / (mac_avg_ms / 1e3)
    it_bytes = 2 * s_dir + T * (4 * nz + 2) * Y * X * 4
    it_s = (plain_ms / 1e3) / (iters + 0.5)
    # one frame (Htf + iters iterations of frame 0): the HBM-bound case
    it1_bytes = 2 * s_dir + (4 * nz + 2) * Y * X * 4
    it1_s = (plain1_ms / 1e3) / (iters + 0.5)
    one = lambda a, f=float: [f(x) for x in a]
    desc = ("%s: %d frames of Nnum=%d, %dx%d px, %d planes (z=-%d..%d), PSF %d..%d px, %d..%d seeded blobs per "
            "frame (seed %d), Poisson peak %g; target = each frame's data fit after %d plain RL iterations from a "
            "uniform start" % (args.workload, T, w["N"], Y, X, nz, w["zs"][-1], w["zs"][-1], min(w["sides"]),
                               max(w["sides"]), w["nblobs"][0], w["nblobs"][1], w["seed"], w["peak"], iters))
    summ = lambda a: {"min": float(np.min(a)), "median": float(np.median(a)), "max": float(np.max(a))}
    res = {
        "metric": "volumes/s at fixed RL data-fit: learned-init vs plain RL, 1 GPU",
        "value": T * 1e3 / learned_ms,
        "unit": "volumes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": learned_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": desc, "frames": T, "launch_options": tuned, "min_iterations": args.min_iter,
            "plain": {"ms": plain_ms, "volumes_per_s": T * 1e3 / plain_ms, "iterations": iters,
                      "datafit": one(fit_plain_t), "rel_l2_vs_truth": one(gt_plain),
                      "rel_l2_vs_truth_summary": summ(gt_plain)},
            "learned": {"ms": learned_ms, "volumes_per_s": T * 1e3 / learned_ms,
                        "iterations_run": int(out["n"]), "iterations_needed_per_frame": need,
                        "iterations_saved_per_frame": [iters - k if k >= 0 else None for k in need],
                        "guess_accepted": one(info["accepted"], bool),
                        "guard_margin_per_frame": one(np.asarray(info["fit_plain1"]) - np.asarray(info["fit_guess"])),
                        "fit_guess": one(info["fit_guess"]), "fit_plain1": one(info["fit_plain1"]),
                        "datafit_reached": one(fit_l_t), "rel_l2_vs_truth": one(gt_learned),
                        "rel_l2_vs_truth_summary": summ(gt_learned), "rel_l2_vs_plain": one(rel_vs_plain),
                        "weights": os.path.relpath(wpath, ROOT)},
            "speedup_vs_plain": plain_ms / learned_ms,
            "one_frame": {"plain_ms": plain1_ms, "learned_ms": learned1_ms,
                          "plain_volumes_per_s": 1e3 / plain1_ms, "learned_volumes_per_s": 1e3 / learned1_ms,
                          "speedup": plain1_ms / learned1_ms},
        },
        "roofline": {"bound": "fp32", "achieved": flop_rate / 1e12, "peak": PEAK_FP32 / 1e12, "unit": "TFLOP/s",
                     "frac": flop_rate / PEAK_FP32, "traffic": measured_traffic(args.workload), "kernel": "k_mac",
                     "launch_avg_ms": mac_avg_ms, "launches": mac_n,
                     "flops_per_launch": mac_flops / max(mac_n, 1),
                     "note": "several frames per workgroup: the MAC is bound on the CU, not by HBM: VALU 57% "
                             "busy, one 16-byte LDS table read per 8 FMAs, two to three waves per SIMD "
                             "(profiles/k_mac_rl1455x32_sq_counters.txt)",
                     "hbm": {"achieved": hbm_rate / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                             "frac": hbm_rate / PEAK_HBM, "bytes_per_launch": mac_bytes / max(mac_n, 1)}},
        "rl_iteration": {"ms": it1_s * 1e3, "frames": 1, "algorithmic_bytes": it1_bytes,
                         "achieved_GBps": it1_bytes / it1_s / 1e9, "frac": it1_bytes / it1_s / PEAK_HBM,
                         "note": "one frame: every cached spectrum read once per projection, HBM-bound"},
        "rl_iteration_batch": {"ms": it_s * 1e3, "frames": T, "algorithmic_bytes": it_bytes,
                               "achieved_GBps": it_bytes / it_s / 1e9, "frac": it_bytes / it_s / PEAK_HBM,
                               "note": "all frames of a step: FP32-bound (see roofline), not HBM-bound"},
        "spectra_precompute_s": t_prep,
    }
    del X1s
    if not args.no_cpu_baseline:
        # the CPU leg: the FP64 oracle timed on a bounded sample, and (as the
        # checker, outside every timed GPU region) frame 0 of the learned leg
        # against the oracle from the same start
        t = time.time()
        wb = dict(hm=hm, Y=Y, X=X, N=w["N"], sides=w["sides"], zs=w["zs"], iters=iters,
                  vol=vol[:, :1].cpu().numpy())
        res["cpu_baseline"] = cpu_baseline(wb)
        log("cpu baseline %.1fs" % (time.time() - t))
        if not args.no_oracle_check:
            g0, _, _ = li.initial_volume(im0, li.projector.backward_device(hm, im0))
            t = time.time()
            res["cpu_baseline"]["oracle_check"] = oracle_check(proj, hm, im0[0], g0[:, 0])
            log("oracle check %.1fs: %s" % (time.time() - t, res["cpu_baseline"]["oracle_check"]["rel_l2_per_plane"]))
    print(json.dumps(res), flush=True)


# Published wall times of the reference for this scenario (CuPy GPU and
# C++/FFTW CPU paths, 4 iterations, F(H) not cached): benchmarking.ipynb:371-375,
# 435-436 (batch 1, and the linear fits over batch size).
NILS_PUBLISHED = {"gpu_batch1_s": 19.0, "gpu_fixed_s": 14.01, "gpu_per_frame_s": 4.98,
                  "cpu_batch1_s": 628.8, "cpu_fixed_s": 565.69, "cpu_per_frame_s": 60.55,
                  "hardware": "48-SM GPU with 224 GB/s memory; 8-core CPU (suil-bheag)"}


def spectrum_bytes(proj, hm):
    """Compulsory read of the cached half-plane spectra of one projection
    direction: sum over planes of U * fY * (fX/2 + 1) * 8 B, U = h(h+1)/2
    lenslet offsets for a square grid (transposes derived), h^2 otherwise."""
    h = (hm.Nnum + 1) // 2
    tot = 0.0
    for z in range(hm.numZ):
        fy, fx = proj.fshape(z)
        u = h * (h + 1) // 2 if fy == fx else h * h
        tot += u * fy * (fx // 2 + 1) * 8
    return tot


def nils_main(args):
    """The reference's own benchmark (benchmarking.ipynb:14-28, 'new' and
    'new-batch' of lf_performance_analysis.py:240-250): DeconvRL with
    Htf=None and 4 iterations of the example Nnum=19 image, on a PSF set
    synthesised on the GPU at the scenario's optics.  One step = Htf = Ht im,
    then 4 Prevedel iterations from X = Htf, inputs resident in HBM; batch 1
    and a batch of 24 copies of the image (the reference's batch runs tile
    the same image).  value = seconds per image in the batch."""
    import torch
    from flfhip.lfdeconv import DeconvRL, deconv_rl_device, _fit_from_sums
    from flfhip.projector import Projector_hip
    from flfhip.workloads import setup_nils
    w = setup_nils(os.path.join(ROOT, "tests", "golden", "nils_lfimage_1273x1463.tif"))
    hm, Y, X, iters = w["hm"], w["Y"], w["X"], w["iters"]
    nz = hm.numZ
    B = args.frames if args.frames > 1 else 24
    proj = Projector_hip()
    dev = proj.device
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

    # first call as a user makes it: numpy image in, numpy volume out, the
    # spectra computed inside (the reference's batch-1 figure includes F(H))
    torch.cuda.synchronize()
    t = time.time()
    first = DeconvRL(hm, None, iters, None, im=w["im"][None], projector=proj, logPrint=False)
    torch.cuda.synchronize()
    first_s = time.time() - t
    log("nils: first DeconvRL call (spectra + Htf + %d iterations, host arrays) %.2fs" % (iters, first_s))
    res_frames = {}
    for T in (1, B):
        t = time.time()
        proj.prepare(hm, T, Y, X)
        torch.cuda.synchronize()
        prep_s = time.time() - t
        im = torch.from_numpy(w["im"]).to(dev)[None].expand(T, Y, X).contiguous()
        Htf = torch.empty((nz, T, Y, X), dtype=torch.float32, device=dev)
        Xv = torch.empty_like(Htf)

        def step():
            proj.backward_device(hm, im, out=Htf)
            Xv.copy_(Htf)
            deconv_rl_device(proj, hm, Htf, iters, Xv)

        for _ in range(args.warmup):
            step()
        ms = timed(step, args.steps)
        fit = _fit_from_sums(proj.datafit_device(proj.forward_device(hm, Xv), im))
        res_frames[T] = {"ms_per_step": ms, "s_per_image": ms / 1e3 / T, "prepare_s": prep_s,
                         "fit_after": float(fit[0])}
        if T == 1:
            X1 = Xv.clone()
            d = float(np.linalg.norm(X1[:, 0].cpu().numpy() - first[:, 0]) / np.linalg.norm(first[:, 0]))
            res_frames[T]["rel_l2_vs_first_call"] = d
        else:
            d = float(torch.linalg.vector_norm((Xv[:, -1] - X1[:, 0]).double()) /
                      torch.linalg.vector_norm(X1[:, 0].double()))
            res_frames[T]["rel_l2_last_frame_vs_batch1"] = d
        log("nils: batch %d: %.1f ms per step, %.2f ms per image, prepare %.2fs" % (T, ms, ms / T, prep_s))
        del Htf, Xv, im
    s_dir = spectrum_bytes(proj, hm)
    per = res_frames[B]["s_per_image"]
    it_bytes = 2 * s_dir + B * (4 * nz + 2) * Y * X * 4
    it_s = res_frames[B]["ms_per_step"] / 1e3 / (iters + 0.5)
    res = {
        "metric": "s per image: 4-iteration Richardson-Lucy of the reference's benchmark image (DeconvRL, Htf=None)",
        "value": per, "unit": "s/image", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": res_frames[B]["ms_per_step"], "higher_is_better": False, "scaling": "weak",
        "vs_baseline": per / NILS_PUBLISHED["gpu_per_frame_s"], "dtype": "f32",
        "data": "real camera image (the reference's example data), PSF synthesised on the GPU at the scenario optics",
        "config": {"workload": "nils: %d copies of the 1273x1463 Nnum=19 image, %d planes z=-60..60 um, PSF %d..%d "
                               "px, %d Prevedel iterations from Htf" % (B, nz, min(w["sides"]), max(w["sides"]),
                                                                         iters),
                   "frames": B, "grids": sorted(set(tuple(proj.fshape(z)) for z in range(nz)))},
        "batch1": res_frames[1], "batch": res_frames[B],
        "first_call_s": first_s, "psf_synthesis_s": w["psf_s"],
        "rl_iteration": {"ms": it_s * 1e3, "frames": B, "algorithmic_bytes": it_bytes,
                         "achieved_GBps": it_bytes / it_s / 1e9, "frac": it_bytes / it_s / PEAK_HBM},
        "reference_published": NILS_PUBLISHED,
        "speedup_vs_published": {"gpu_per_frame": NILS_PUBLISHED["gpu_per_frame_s"] / per,
                                 "gpu_batch1_first_call": NILS_PUBLISHED["gpu_batch1_s"] / first_s,
                                 "cpu_per_frame": NILS_PUBLISHED["cpu_per_frame_s"] / per},
    }
    if not args.no_cpu_baseline:
        # the CPU leg: the FP64 oracle timed on a bounded sample and scaled
        # to one image (Htf + 2 projections per iteration), and as the
        # checker, one iteration on two planes from X = Htf against it
        t = time.time()
        wb = dict(hm=hm, Y=Y, X=X, N=w["N"], sides=w["sides"], zs=w["zs"], iters=iters,
                  vol=[w["im"][None]] * nz)
        cb = cpu_baseline(wb)
        cb["value"], cb["unit"] = 1.0 / cb["value"], "s/image"
        res["cpu_baseline"] = cb
        log("cpu baseline %.1fs: %.0f s/image" % (time.time() - t, cb["value"]))
        if not args.no_oracle_check:
            t = time.time()
            im1 = torch.from_numpy(w["im"]).to(dev)[None].contiguous()
            proj.prepare(hm, 1, Y, X)
            Htf1 = proj.backward_device(hm, im1)
            cb["oracle_check"] = oracle_check(proj, hm, im1[0], Htf1[:, 0], iters=1, planes=(0, nz // 2),
                                              start="X = Htf (DeconvRL's default Xguess)")
            log("nils oracle check %.1fs: %s" % (time.time() - t, cb["oracle_check"]["rel_l2_per_plane"]))
    print(json.dumps(res), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="rl1455x32", choices=sorted(WORKLOADS) + sorted(RAGGED) + sorted(BATCHES) + ["nils"])
    ap.add_argument("--buckets", type=int, default=2, help="ragged workloads: canvas sizes (at most)")
    ap.add_argument("--graph", action="store_true", help="ragged workloads: replay iterations as HIP graphs")
    ap.add_argument("--weights", default=None, help="learned-initialiser weights (default: the packaged file)")
    ap.add_argument("--frames", type=int, default=1,
                    help="frames reconstructed together per step (value = frames x steps / time)")
    ap.add_argument("--plain-only", action="store_true", help="time plain RL only")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-oracle-check", action="store_true", help="batch workloads: skip the FP64 oracle check")
    ap.add_argument("--min-iter", type=int, default=0,
                    help="batch workloads: RL iterations the learned leg runs before its target may stop it")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--options", default="",
                    help="batch workloads: launch options name=value,... instead of the autotuner (for profiler runs)")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="also run one device copy of this many MiB (reference kernel for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this implementation has no multi-GPU path (frames or planes would be "
              "sharded across GPUs with one all-reduce of the forward image per iteration).",
              file=sys.stderr)
        sys.exit(2)
    if args.workload in RAGGED:
        return ragged_main(args)
    if args.workload in BATCHES:
        return batch_main(args)
    if args.workload == "nils":
        return nils_main(args)

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
        li.plain_denominator(Y, X)             # cached per PSF set and frame shape, like the spectra
        out = {}

        def learned_step():
            # Htf = Ht(im) into Htf, beside the network's guess
            Xl, n_it, info = li.deconvolve(im, None, 4 * iters, targetFit=fit_plain_t, Htf_out=Htf)
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
