"""Deconvolve a time series of rectified light-field images on the MI355X
path: the reference's batch driver (deconvolve_time_series.py:9-85) with
the same options, output files and RL semantics.

    python -m flfhip.deconvolve_time_series -p PSF.mat -d OUT -t img1.tif img2.tif ...

Per batch of up to --batch-size timepoints: Htf = Ht(images), then
--iters Prevedel RL iterations starting from Xguess = Htf
(deconvolve_time_series.py:27-28); each result is written to
OUT/<name>_deconvolved.tif as one float32 page per plane.  The PSF file is
the reference's memmap cache next to PSF.mat (psfmatrix.LoadMatrix).

Beyond the reference: the reference refuses a batch whose images differ in
shape (deconvolve_time_series.py:20-23); --ragged deconvolves such a batch
in canvas buckets (flfhip.ragged).  --learned WEIGHTS starts RL from the
guarded learned initialiser instead of Htf (flfhip.lensletnet.LearnedInit:
the guess is kept only where its data fit beats one plain iteration), and
--target-fit stops each frame's RL once its data fit reaches that value.
Deconvolution always runs on the GPU: --mode cpu is accepted for
command-line compatibility and prints a note.
"""
import argparse
import os
import sys
import time

import numpy as np

from . import lfimage


def _log(*a):
    print(*a, flush=True)


def _check_shapes(images, timepoints, ragged):
    shapes = [im.shape for im in images]
    if len(set(shapes)) > 1 and not ragged:
        # deconvolve_time_series.py:21-23
        print("ERROR: not all images have the same dimensions - cannot deconvolve these all together "
              "(use --ragged to deconvolve them in canvas buckets)")
        sys.exit(1)
    return shapes


def ProcessBatch(projector, hMatrix, timepoints, destDir, maxIter, ragged=False, nbuckets=2, learned=None,
                 targetFit=None):
    """deconvolve_time_series.py:9-34 for one batch; returns the list of
    written file paths."""
    os.makedirs(destDir, exist_ok=True)
    images = [lfimage.LoadLightFieldTiff(t) for t in timepoints]
    shapes = _check_shapes(images, timepoints, ragged)
    if len(set(shapes)) > 1:
        from .ragged import RaggedBatch
        rb = RaggedBatch(hMatrix, shapes, nbuckets=nbuckets, projectors=[projector])
        vols = rb.deconv(images, maxIter, targetFit=targetFit)
    else:
        vols = _deconv_uniform(projector, hMatrix, np.array(images), maxIter, learned, targetFit)
    out = []
    for i, t in enumerate(timepoints):
        dest = "%s/%s_deconvolved.tif" % (destDir, os.path.splitext(os.path.basename(t))[0])
        lfimage.SaveVolumeTiff(dest, vols[i])
        out.append(dest)
    return out


def _deconv_uniform(projector, hMatrix, images, maxIter, learned, targetFit):
    from . import lfdeconv
    if learned is None:
        # deconvolve_time_series.py:27-28
        Htf = lfdeconv.BackwardProjectACC(hMatrix, images, progress=None, logPrint=False, projector=projector)
        Xguess = Htf.copy() if isinstance(Htf, np.ndarray) else Htf.clone()
        res = lfdeconv.DeconvRL(hMatrix, Htf, maxIter, Xguess, logPrint=False, projector=projector, im=images,
                                targetFit=targetFit)
        return [res[:, i] for i in range(images.shape[0])]
    from .lensletnet import LearnedInit
    T, Y, X = images.shape
    projector.prepare(hMatrix, T, Y, X, frames_capacity=2 * T)
    im = projector.asnative(images)
    init = LearnedInit(learned, projector, hMatrix)
    Xv, n, info = init.deconvolve(im, None, maxIter, targetFit=targetFit)     # Htf computed beside the guess
    _log("learned start: guess kept for %d of %d frame(s); %d RL iteration(s)"
         % (int(np.sum(info["accepted"])), T, n))
    Xn = projector.asnumpy(Xv)
    return [Xn[:, i] for i in range(T)]


def main(argv, maxIter=8, numParallel=32):
    """deconvolve_time_series.py:37-82."""
    parser = argparse.ArgumentParser(
        description="Deconvolves a time series of (pre-rectified) light field microscopy images.",
        formatter_class=argparse.ArgumentDefaultsHelpFormatter)
    parser.add_argument("-m", "--mode", dest="platform", metavar="PLATFORM", choices=["cpu", "gpu", "gpubig"],
                        default="gpu", help="accepted for compatibility; deconvolution runs on the GPU")
    parser.add_argument("-b", "--batch-size", dest="numParallel", metavar="BATCH-SIZE", type=int,
                        default=numParallel, help="number of simultaneous deconvolutions")
    parser.add_argument("-i", "--iters", dest="numIter", metavar="RL-ITERS", type=int, default=maxIter,
                        help="number of iterations of Richardson-Lucy to run")
    parser.add_argument("-v", "--volumes-on-gpu", dest="storeVolumesOnGPU", action="store_true", default=False,
                        help="keep volumes in GPU memory between calls")
    parser.add_argument("-c", "--cacheFH", dest="cacheFH", action="store_true", default=False,
                        help="accepted for compatibility; PSF spectra are always cached in HBM")
    parser.add_argument("-d", "--dest", dest="destDir", metavar="DEST-DIR", default=".",
                        help="directory in which deconvolved output will be saved")
    parser.add_argument("-p", "--psf", dest="psfFile", metavar="PSF-FILE", required=True,
                        help="path to the PSF file (its memmap cache directory must exist)")
    parser.add_argument("-t", "--timepoints", dest="timepoints", metavar="IMG", required=True, nargs="+",
                        help="files to deconvolve")
    parser.add_argument("--ragged", action="store_true",
                        help="deconvolve batches whose images differ in shape (canvas buckets)")
    parser.add_argument("--buckets", type=int, default=2, help="canvas sizes per ragged batch (at most)")
    parser.add_argument("--learned", metavar="WEIGHTS", default=None,
                        help="start RL from the guarded learned initialiser with these weights (.npz)")
    parser.add_argument("--target-fit", dest="targetFit", type=float, default=None,
                        help="stop a frame's RL once ||H X - im|| / ||im|| <= this value")
    args = parser.parse_args(argv)

    # deconvolve_time_series.py:65-70
    if args.numParallel < 1:
        print("\033[0;31mERROR: Invalid batch size: %d\033[0m" % args.numParallel)
        sys.exit(1)
    if args.numIter < 1:
        print("\033[0;31mERROR: Invalid number of RL iterations: %d\033[0m" % args.numIter)
        sys.exit(1)
    if args.platform == "cpu":
        print("\033[0;33mNOTE: deconvolution runs on the GPU (there is no CPU path)\033[0m")
    if args.learned and args.ragged:
        print("\033[0;31mERROR: --learned applies to batches of one image shape\033[0m")
        sys.exit(1)

    from . import psfmatrix
    from .projector import Projector_hip
    projector = Projector_hip()
    projector.storeVolumesOnGPU = args.storeVolumesOnGPU
    hMatrix = psfmatrix.LoadMatrix(args.psfFile)
    net = None
    if args.learned:
        from .lensletnet import LensletNet
        net = LensletNet.load(args.learned)

    n = len(args.timepoints)
    if n > args.numParallel:
        print("Deconvolving timepoints in batches of %d" % args.numParallel)
    else:
        print("Deconvolving all %d timepoints simultaneously" % n)
    written = []
    t0 = time.time()
    for t in range(0, n, args.numParallel):
        tEnd = min(t + args.numParallel, n)
        written += ProcessBatch(projector, hMatrix, args.timepoints[t:tEnd], args.destDir, args.numIter,
                                ragged=args.ragged, nbuckets=args.buckets, learned=net, targetFit=args.targetFit)
        _log("deconvolved %d/%d timepoints (%.1f s)" % (tEnd, n, time.time() - t0))
    return written


if __name__ == "__main__":
    main(sys.argv[1:])
