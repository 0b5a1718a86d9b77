"""Synthetic workloads of the benchmark and of the learned initialiser's
training: seeded ring PSF sets (synthetic.make_psf_set) and blob volumes.

    name     Nnum  lenslets  planes      PSF sides   blobs  Poisson peak  RL iterations
    rl1455   15    97x97     z=-20..20   15..345     400    2000          8
    rl300    15    20x20     z=-5..5     15..135     200    1000          4

Time series of uneven regions (RAGGED): T frames whose lenslet grids are
drawn uniformly from L_min..L_max (square), one shared PSF set, blobs
drifting 0.5 px per frame (synthetic.drifting_blob_frames).

    name       frames  Nnum  lenslets  planes      PSF sides  blobs/frame  peak  RL iterations
    ragged64   64      15    20..60    z=-15..15   15..255    50..300      1000  8

The reference's own benchmark scenario (NILS; benchmarking.ipynb:14-28,
the 'nils-image' and 'nils-matrix' options of lf_performance_analysis.py:141,
161): one 1273x1463 light-field camera image of Nnum=19 lenslets, the PSF of
M 22.222, NA 0.5, ML pitch 125 um, f_ML 3125 um, n 1.33, 520 nm at 25 planes
z=-60..60 um (5 um apart), 4 RL iterations from Htf (DeconvRL with Htf=None).

Batches of full-size frames (BATCHES): T frames of one size reconstructed
together, each a seeded blob volume with its own blob count, on the PSF set
of a base workload (so that a network trained for it applies).

    name        base    frames  blobs/frame  peak  RL iterations  seed
    rl1455x32   rl1455  32      100..800     2000  8              6
"""
import sys
import time

from . import synthetic

WORKLOADS = {
    # name: (Nnum, lenslets, zmax, s_min, s_max, blobs, peak, iters, seed)
    "rl1455": (15, 97, 20, 15, 345, 400, 2000.0, 8, 2),
    "rl300": (15, 20, 5, 15, 135, 200, 1000.0, 4, 1),
}


BATCHES = {
    # name: (base workload, frames, blobs_min, blobs_max, peak, iters, seed)
    "rl1455x32": ("rl1455", 32, 100, 800, 2000.0, 8, 6),
}


# (M, NA, MLPitch, Nnum, OSR, n, fml, lambda, zmin, zmax, zspacing) of
# GeneratePSF, image shape, RL iterations
NILS = {"optics": (22.222, 0.5, 125e-6, 19, 3, 1.33, 3125e-6, 520e-9, -60e-6, 60e-6, 5e-6),
        "shape": (1273, 1463), "iters": 4}


RAGGED = {
    # name: (frames, Nnum, L_min, L_max, zmax, s_min, s_max, blobs_min, blobs_max, peak, iters, seed)
    "ragged64": (64, 15, 20, 60, 15, 15, 255, 50, 300, 1000.0, 8, 3),
}


def _log(*a):
    print(*a, file=sys.stderr, flush=True)


def psf_set(name):
    """(HMatrix, zs, sides) of a workload (WORKLOADS, or the shared PSF set of
    a RAGGED time series)."""
    if name in RAGGED:
        T, N, lmin, lmax, zmax, smin, smax, bmin, bmax, peak, iters, seed = RAGGED[name]
    else:
        N, L, zmax, smin, smax, nblobs, peak, iters, seed = WORKLOADS[name]
    zs = list(range(-zmax, zmax + 1))
    sides = synthetic.psf_sides(zs, smin, smax)
    t = time.time()
    hm = synthetic.make_psf_set(N, zs, sides, seed=seed)
    _log("psf set: %d planes, sides %d..%d, %.1fs" % (len(zs), min(sides), max(sides), time.time() - t))
    return hm, zs, sides


def setup(name, volume=True, frames=1):
    """Workload dictionary; with frames > 1 the volume holds that many seeded
    frames (frame 0 is the single-frame volume)."""
    N, L, zmax, smin, smax, nblobs, peak, iters, seed = WORKLOADS[name]
    hm, zs, sides = psf_set(name)
    Y = X = N * L
    vol = None
    if volume:
        t = time.time()
        vol = synthetic.blob_volume(len(zs), Y, X, nblobs, seed=seed, T=frames)
        _log("volume: %s, %.1fs" % (vol.shape, time.time() - t))
    return dict(name=name, N=N, Y=Y, X=X, zs=zs, sides=sides, hm=hm, vol=vol, peak=peak, iters=iters, seed=seed,
                nblobs=nblobs)


def setup_ragged(name, volume=True):
    """Ragged time-series workload: PSF set, per-frame shapes and (with
    volume) the frames' blob volumes.  Frame t's lenslet grid and blob count
    come from a generator seeded with seed + t."""
    T, N, lmin, lmax, zmax, smin, smax, bmin, bmax, peak, iters, seed = RAGGED[name]
    hm, zs, sides = psf_set(name)
    shapes, nblobs = [], []
    for t in range(T):
        rng = synthetic.np.random.default_rng(seed + t)
        L = int(rng.integers(lmin, lmax + 1))
        shapes.append((N * L, N * L))
        nblobs.append(int(rng.integers(bmin, bmax + 1)))
    vols = None
    if volume:
        t0 = time.time()
        vols = synthetic.drifting_blob_frames(len(zs), shapes, nblobs, seed=seed, field=(N * lmax, N * lmax))
        _log("volumes: %d frames, %.1fs" % (T, time.time() - t0))
    return dict(name=name, N=N, zs=zs, sides=sides, hm=hm, shapes=shapes, nblobs=nblobs, vols=vols, peak=peak,
                iters=iters, seed=seed)


def setup_batch(name):
    """Batch workload: the base workload's PSF set and geometry, and the
    per-frame parameters; the volumes are made on the GPU by
    lensletnet.blob_volumes_gpu(nz, Y, X, frames, (blobs_min, blobs_max),
    seed) (frame t's blob count is the t-th draw of that seeded generator)."""
    base, T, bmin, bmax, peak, iters, seed = BATCHES[name]
    N, L = WORKLOADS[base][0], WORKLOADS[base][1]
    hm, zs, sides = psf_set(base)
    Y = X = N * L
    return dict(name=name, base=base, N=N, Y=Y, X=X, zs=zs, sides=sides, hm=hm, frames=T, nblobs=(bmin, bmax),
                peak=peak, iters=iters, seed=seed)


def setup_nils(image_path):
    """The reference's benchmark scenario: the camera image (float32, as the
    reference reads it) and its PSF set synthesised on the GPU
    (psfgen.GeneratePSF at the scenario's optics; seconds, where the
    reference's CPU generation takes tens of minutes)."""
    import numpy as np
    import torch
    from . import lfimage, psfgen
    im = lfimage.LoadLightFieldTiff(image_path)
    if im.shape != NILS["shape"]:
        raise ValueError("%s: expected a %s image, found %s" % (image_path, NILS["shape"], im.shape))
    torch.cuda.synchronize()
    t = time.time()
    hm, info = psfgen.GeneratePSF(*NILS["optics"])
    torch.cuda.synchronize()
    t_psf = time.time() - t
    sides = [int(hm.PSFShape(z)[0]) for z in range(hm.numZ)]
    _log("nils: PSF set of %d planes, sides %d..%d, synthesised in %.1fs" % (hm.numZ, min(sides), max(sides), t_psf))
    return dict(name="nils", N=NILS["optics"][3], Y=im.shape[0], X=im.shape[1], im=im.astype(np.float32), hm=hm,
                zs=[z * 1e6 for z in info["zs"]], sides=sides, iters=NILS["iters"], psf_s=t_psf, psf_info=info)
