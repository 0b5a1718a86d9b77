"""Batches of frames with uneven lenslet grids (regions of interest of
different sizes) in one call per kernel.

The reference deconvolves a time series only when every frame has one shape
(deconvolve_time_series.py:21-23) and plans one convolution shape per image
size (special_fftconvolve.py:142-174).  Here frames are grouped into a few
canvas sizes ("buckets"); the frames of a bucket are zero-padded into a
(T, Yc, Xc) batch that shares one plan and one PSF spectrum cache, and the
library is told each frame's own size (flf_set_frame_dims), so that
projections read only that region and write zeros outside it.  Zero padding
does not change a linear convolution's cropped output, so every frame's
result is its projection (and its RL reconstruction) at its own size.

Each bucket has its own Projector_hip (its own Fourier grids and spectra).
Fewer buckets mean fewer spectrum sets to hold and read; more buckets mean
less padding.  plan_buckets chooses the split.
"""
import contextlib

import numpy as np
import torch

from .lfdeconv import RLGraph, deconv_rl_device
from .projector import Projector_hip


def plan_buckets(shapes, nbuckets=1, spectrum_weight=4.0):
    """Split frames into at most nbuckets canvases.

    shapes: [(Y_t, X_t)].  Frames are ordered by canvas side and split into
    contiguous runs by dynamic programming over the cost
    sum_b (frames_b + spectrum_weight) * Yc_b * Xc_b: padded pixels
    transformed per bucket, plus a per-bucket charge for its spectrum set
    (read once per projection, shared by the bucket's frames).  Returns
    [((Yc, Xc), [frame indices])] with frames in input order within a bucket.
    """
    n = len(shapes)
    if n == 0:
        return []
    nb = max(1, min(int(nbuckets), n))
    order = sorted(range(n), key=lambda i: (max(shapes[i]), shapes[i][0], shapes[i][1], i))

    def canvas(i, j):      # frames order[i:j]
        return (max(shapes[k][0] for k in order[i:j]), max(shapes[k][1] for k in order[i:j]))

    def cost(i, j):
        yc, xc = canvas(i, j)
        return ((j - i) + spectrum_weight) * float(yc) * float(xc)

    INF = float("inf")
    best = [[INF] * (n + 1) for _ in range(nb + 1)]
    cut = [[0] * (n + 1) for _ in range(nb + 1)]
    best[0][0] = 0.0
    for b in range(1, nb + 1):
        for j in range(1, n + 1):
            for i in range(b - 1, j):
                if best[b - 1][i] == INF:
                    continue
                c = best[b - 1][i] + cost(i, j)
                if c < best[b][j]:
                    best[b][j], cut[b][j] = c, i
    b = min(range(1, nb + 1), key=lambda k: (best[k][n], k))
    runs, j = [], n
    while b > 0:
        i = cut[b][j]
        runs.append((i, j))
        j, b = i, b - 1
    out = []
    for i, j in reversed(runs):
        out.append((canvas(i, j), sorted(order[i:j])))
    return out


class RaggedBatch(object):
    """Frames of uneven size, deconvolved together.

    hMatrix: the shared PSF set.  shapes: [(Y_t, X_t)] per frame.  nbuckets:
    at most this many canvas sizes (plan_buckets).  projectors: optional list
    of Projector_hip to reuse, one per bucket (created on `device` otherwise).
    """

    def __init__(self, hMatrix, shapes, nbuckets=1, projectors=None, device=0, spectrum_weight=4.0):
        self.hMatrix = hMatrix
        self.shapes = [(int(y), int(x)) for y, x in shapes]
        self.buckets = plan_buckets(self.shapes, nbuckets, spectrum_weight)
        projectors = list(projectors or [])
        while len(projectors) < len(self.buckets):
            projectors.append(Projector_hip(device))
        self.projectors = projectors[:len(self.buckets)]
        self._graphs = {}

    @contextlib.contextmanager
    def _bucket(self, k):
        """Plan bucket k and set its frames' sizes (on the current stream);
        the projector returns to full-size frames afterwards."""
        (yc, xc), idx = self.buckets[k]
        p = self.projectors[k]
        p.prepare(self.hMatrix, len(idx), yc, xc)
        p.set_frame_dims([self.shapes[i] for i in idx])
        try:
            yield p, idx, yc, xc
        finally:
            p.set_frame_dims(None)

    def _pad_images(self, images, idx, yc, xc, device):
        out = torch.zeros((len(idx), yc, xc), dtype=torch.float32, device=device)
        for j, i in enumerate(idx):
            y, x = self.shapes[i]
            im = torch.as_tensor(np.asarray(images[i], np.float32).reshape(y, x))
            out[j, :y, :x] = im.to(device)
        return out

    def _pad_volumes(self, vols, idx, yc, xc, device):
        nz = self.hMatrix.numZ
        out = torch.zeros((nz, len(idx), yc, xc), dtype=torch.float32, device=device)
        for j, i in enumerate(idx):
            y, x = self.shapes[i]
            v = torch.as_tensor(np.asarray(vols[i], np.float32).reshape(nz, y, x))
            out[:, j, :y, :x] = v.to(device)
        return out

    def _check_shapes(self, arrays, lead):
        if len(arrays) != len(self.shapes):
            raise ValueError("%d arrays for %d frames" % (len(arrays), len(self.shapes)))
        for i, a in enumerate(arrays):
            want = lead + self.shapes[i]
            if tuple(np.shape(a)) != want:
                raise ValueError("frame %d has shape %s, expected %s" % (i, tuple(np.shape(a)), want))

    def forward(self, vols):
        """vols: [(numZ, Y_t, X_t)] -> [(Y_t, X_t)] images (numpy)."""
        self._check_shapes(vols, (self.hMatrix.numZ,))
        out = [None] * len(self.shapes)
        for k in range(len(self.buckets)):
            with self._bucket(k) as (p, idx, yc, xc):
                img = p.forward_device(self.hMatrix, self._pad_volumes(vols, idx, yc, xc, p.device))
                img = img.cpu().numpy()
            for j, i in enumerate(idx):
                y, x = self.shapes[i]
                out[i] = img[j, :y, :x].copy()
        return out

    def backward(self, images):
        """images: [(Y_t, X_t)] -> [(numZ, Y_t, X_t)] volumes (numpy)."""
        self._check_shapes(images, ())
        out = [None] * len(self.shapes)
        for k in range(len(self.buckets)):
            with self._bucket(k) as (p, idx, yc, xc):
                vol = p.backward_device(self.hMatrix, self._pad_images(images, idx, yc, xc, p.device))
                vol = vol.cpu().numpy()
            for j, i in enumerate(idx):
                y, x = self.shapes[i]
                out[i] = vol[:, j, :y, :x].copy()
        return out

    def stage(self, images):
        """Per bucket, the zero-padded (T_b, Yc, Xc) image batch on its
        projector's device."""
        self._check_shapes(images, ())
        out = []
        for k, ((yc, xc), idx) in enumerate(self.buckets):
            out.append(self._pad_images(images, idx, yc, xc, self.projectors[k].device))
        return out

    def unstage(self, batches):
        """Per-bucket (numZ, T_b, Yc, Xc) volumes -> [(numZ, Y_t, X_t)] numpy."""
        out = [None] * len(self.shapes)
        for k, ((yc, xc), idx) in enumerate(self.buckets):
            Xn = batches[k].cpu().numpy()
            for j, i in enumerate(idx):
                y, x = self.shapes[i]
                out[i] = Xn[:, j, :y, :x].copy()
        return out

    def deconv_staged(self, staged, maxIter, Xguess=None, prevedel=True, targetFit=None, fitOut=None,
                      graph=False, out=None):
        """RL on staged image batches (stage()); returns per-bucket volumes
        (numZ, T_b, Yc, Xc), zero outside each frame.  Xguess: per-bucket
        volumes or None (start from Htf).  out: per-bucket (Htf, X) buffers
        to reuse.  See deconv for the other arguments."""
        if graph and (not prevedel or targetFit is not None or fitOut is not None):
            raise ValueError("graph replay runs plain Prevedel iterations only")
        res = []
        for k in range(len(self.buckets)):
            with self._bucket(k) as (p, idx, yc, xc):
                im = staged[k]
                Htf = p.backward_device(self.hMatrix, im, out=None if out is None else out[k][0])
                if out is None:
                    X = Htf.clone() if Xguess is None else Xguess[k].clone()
                else:
                    X = out[k][1]
                    X.copy_(Htf if Xguess is None else Xguess[k])
                if graph:
                    # one captured iteration per bucket and buffer pair; the
                    # bucket's frame sizes live in its projector (set above)
                    key = (k, Htf.data_ptr(), X.data_ptr())
                    g = self._graphs.get(key) if out is not None else None
                    if g is None:
                        g = RLGraph(p, self.hMatrix, Htf, X)
                        if out is not None:
                            self._graphs[key] = g
                    g.run(maxIter)
                else:
                    tgt = None if targetFit is None else np.broadcast_to(
                        np.asarray(targetFit, np.float64), (len(self.shapes),))[idx]
                    fits = [] if fitOut is not None else None
                    deconv_rl_device(p, self.hMatrix, Htf, maxIter, X, im=im, prevedel=prevedel, targetFit=tgt,
                                     fitOut=fits)
                    if fitOut is not None:
                        fitOut.append((list(idx), fits))
                res.append(X)
        return res

    def deconv(self, images, maxIter, Xguess=None, prevedel=True, targetFit=None, fitOut=None, graph=False):
        """Richardson-Lucy for every frame (lfdeconv.py:72-202 semantics per
        frame: Htf = H^T im, start from Htf or Xguess[t]).  Returns
        [(numZ, Y_t, X_t)] volumes.  fitOut (a list) receives, per bucket,
        (frame indices, [per-iteration fit arrays]).  graph=True replays each
        Prevedel iteration from a captured HIP graph (no fit reporting)."""
        staged = self.stage(images)
        Xg = None
        if Xguess is not None:
            self._check_shapes(Xguess, (self.hMatrix.numZ,))
            Xg = [self._pad_volumes(Xguess, idx, yc, xc, self.projectors[k].device)
                  for k, ((yc, xc), idx) in enumerate(self.buckets)]
        return self.unstage(self.deconv_staged(staged, maxIter, Xguess=Xg, prevedel=prevedel, targetFit=targetFit,
                                               fitOut=fitOut, graph=graph))
