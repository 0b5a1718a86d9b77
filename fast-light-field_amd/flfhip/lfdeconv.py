"""Deconvolution driver with the reference's call surface (lfdeconv.py:27-202).

``ForwardProjectACC``, ``BackwardProjectACC`` and ``DeconvRL`` take the same
arguments as the reference functions.  With a ``Projector_hip`` (the default)
DeconvRL keeps the volume, Htf and the images in HBM and runs each iteration
as one forward projection plus one backward projection fused with the RL
update; with any other projector object it follows the reference's generic
loop through that projector's methods.

Beyond the reference, ``DeconvRL`` can stop at a target forward-model data
fit (``targetFit``) and report the fit after every iteration (``fitOut``):
rho(X) = ||F X - im||_2 / ||im||_2 per frame, FP64 sums.
"""
import time

import numpy as np
import torch

from .projector import Projector_hip

_default_projector = None


def default_projector():
    global _default_projector
    if _default_projector is None:
        _default_projector = Projector_hip()
    return _default_projector


def _noprogress(work, desc=None, leave=True, **kw):
    return work


def BackwardProjectACC(hMatrix, projection, planes=None, progress=None, logPrint=True, numjobs=1, projector=None,
                       rlUpdateFunc=None):
    """lfdeconv.py:27-49."""
    if len(projection.shape) != 3:
        raise ValueError("projection must be (T, Y, X)")
    projector = projector or default_projector()
    planes = range(hMatrix.numZ) if planes is None else planes
    t1 = time.time()
    out = projector.BackwardProjectACC(hMatrix, projection, planes, progress or _noprogress, logPrint, numjobs,
                                       rlUpdateFunc)
    if logPrint:
        print("Backward project: elapsed wallclock time %f" % (time.time() - t1))
    return out


def ForwardProjectACC(hMatrix, realspace, planes=None, progress=None, logPrint=True, numjobs=1, projector=None):
    """lfdeconv.py:51-70."""
    if len(realspace.shape) != 4:
        raise ValueError("realspace must be (numZ, T, Y, X)")
    projector = projector or default_projector()
    planes = range(hMatrix.numZ) if planes is None else planes
    t1 = time.time()
    out = projector.ForwardProjectACC(hMatrix, realspace, planes, progress or _noprogress, logPrint, numjobs)
    if logPrint:
        print("Forward project: elapsed wallclock time %f" % (time.time() - t1))
    return out


def _fit_from_sums(sums):
    return np.sqrt(sums[:, 0]) / np.sqrt(np.maximum(sums[:, 1], 1e-300))


def deconv_rl_device(projector, hMatrix, Htf, maxIter, X, im=None, prevedel=True, trust=None, targetFit=None,
                     fitOut=None, planes=None, HX0=None, fit0=None, minIter=0):
    """Device-resident RL iterations.  Htf, X (numZ, T, Y, X) and im (T, Y, X)
    are float32 tensors on the projector's device; X is updated in place.
    Returns the number of iterations run.  With targetFit (array-like per
    frame, or scalar) it stops once every frame's fit of the current X is
    <= target (the fit of X is measured by the forward projection that
    starts the next iteration, so this costs no extra projection).  HX0, if
    given, is the forward projection of the starting X (saves one); fit0,
    with HX0, is its per-frame data fit (saves one reduction).  The standard
    update (prevedel=False) takes its error image im / HX from the pass that
    writes HX (flf_forward_ratio) and ignores HX0.  minIter: iterations run
    before the target may stop the loop."""
    T, Y, Xw = X.shape[1:]
    projector._need(X, "X", (hMatrix.numZ, T, Y, Xw))
    if prevedel:
        projector._need(Htf, "Htf", (hMatrix.numZ, T, Y, Xw))
    if im is not None:
        projector._need(im, "im", (T, Y, Xw))
    if not prevedel:
        if im is None:
            raise ValueError("standard RL needs the image im")
        HX0 = fit0 = None
    it = 0
    target = None if targetFit is None else np.broadcast_to(np.asarray(targetFit, np.float64), (X.shape[1],))
    need_fit = im is not None and (fitOut is not None or target is not None)
    while True:
        if it >= maxIter and not need_fit:
            break
        known = None
        err = None
        if HX0 is not None:
            HX, HX0 = HX0, None
            known, fit0 = fit0, None
        elif prevedel:
            HX = projector.forward_device(hMatrix, X, planes)
        else:
            HX, err = projector.forward_ratio_device(hMatrix, X, im, planes, store_image=need_fit)
        if need_fit:
            fit = np.asarray(known, np.float64) if known is not None else _fit_from_sums(projector.datafit_device(HX, im))
            if fitOut is not None:
                fitOut.append(fit)
            if target is not None and it >= minIter and np.all(fit <= target):
                break
        if it >= maxIter:
            break
        if prevedel:
            projector.rl_update_device(hMatrix, HX, X, Htf=Htf, mode=1, planes=planes)
        else:
            projector.rl_update_device(hMatrix, err, X, trust=trust, mode=2, planes=planes)
        it += 1
    return it


class RLGraph(object):
    """One Prevedel RL iteration (forward projection + fused back projection
    and update of X in place) captured as a HIP graph on the current device,
    replayed once per iteration: the launches of an iteration go to the GPU
    as one graph instead of one call each.  X and Htf are the static buffers
    the graph reads and writes (copy new data into them, do not rebind).
    The capture happens on first use, after one eager warm-up iteration on a
    scratch copy, so every lazy allocation and table upload is done before
    capture; the captured work contains no memset or host-memory copy.  The
    plane selection is part of the captured kernel arguments; the plan's
    device tables (and per-frame sizes) are not, so the graph is captured
    again whenever the projector has been planned again or its frame sizes
    or frame count changed since the capture (Projector_hip.state_key)."""

    def __init__(self, projector, hMatrix, Htf, X, planes=None):
        self.projector, self.hMatrix, self.Htf, self.X, self.planes = projector, hMatrix, Htf, X, planes
        self.graph = None
        self._state = None

    def _capture(self):
        p, hm = self.projector, self.hMatrix
        Xw = self.X.clone()
        HX = p.forward_device(hm, Xw, self.planes)
        p.rl_update_device(hm, HX, Xw, Htf=self.Htf, mode=1, planes=self.planes)
        torch.cuda.synchronize()
        del Xw, HX
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            HX = p.forward_device(hm, self.X, self.planes)
            p.rl_update_device(hm, HX, self.X, Htf=self.Htf, mode=1, planes=self.planes)
        self._hx = HX
        self._state = p.state_key()

    def run(self, iterations):
        T, Y, Xw = self.X.shape[1:]
        p = self.projector
        # the plan for this shape must be the active one (a no-op when it is)
        p._ensure_plan(self.hMatrix, T, Y, Xw)
        if self.graph is None or self._state != p.state_key():
            self.graph = None
            self._capture()
        for _ in range(iterations):
            self.graph.replay()


def DeconvRL(hMatrix, Htf, maxIter, Xguess, logPrint=True, numjobs=1, projector=None, im=None, progress=None,
             prevedel=True, trustMask=None, targetFit=None, fitOut=None):
    """lfdeconv.py:72-202 (Prevedel update by default; standard RL with an
    optional trust mask when prevedel=False).  Returns a numpy volume."""
    projector = projector or default_projector()
    t1 = time.time()
    if not isinstance(projector, Projector_hip):
        return _deconv_generic(hMatrix, Htf, maxIter, Xguess, logPrint, numjobs, projector, im, prevedel, trustMask)
    if Htf is None:
        assert im is not None
    # shapes are checked before anything moves to the device
    if im is not None:
        if len(np.shape(im)) != 3:
            raise ValueError("im must be (T, Y, X)")
        T, Y, X = np.shape(im)
    else:
        if Htf is None or len(np.shape(Htf)) != 4:
            raise ValueError("Htf must be (numZ, T, Y, X) when im is not given")
        T, Y, X = np.shape(Htf)[1:]
    want = (hMatrix.numZ, T, Y, X)
    for name, arr in (("Htf", Htf), ("Xguess", Xguess)):
        if arr is not None and tuple(np.shape(arr)) != want:
            raise ValueError("%s has shape %s, expected %s (numZ,) + image shape" % (name, tuple(np.shape(arr)), want))
    if trustMask is not None and tuple(np.shape(trustMask)) != (Y, X):
        raise ValueError("trustMask has shape %s, expected %s" % (tuple(np.shape(trustMask)), (Y, X)))
    imT = projector.asnative(im) if im is not None else None
    projector.prepare(hMatrix, T, Y, X)
    if Htf is None:
        if prevedel or Xguess is None:
            HtfT = projector.backward_device(hMatrix, imT)
        else:
            HtfT = None
    else:
        HtfT = projector.asnative(Htf)
    if Xguess is None:
        XT = HtfT.clone()
    else:
        XT = projector.asnative(Xguess).clone()
    trust = None
    if trustMask is not None:
        if prevedel:
            raise AssertionError("trustMask applies to the standard update only (lfdeconv.py:113)")
        tm = np.asarray(trustMask, np.float32)
        wh = np.where(tm == 0)
        for _im in np.asarray(im):
            assert np.sum(_im[wh]) == 0
        trust = projector.backward_device(hMatrix, projector.asnative(tm[np.newaxis]))
    if (not prevedel) and imT is None:
        raise ValueError("standard RL needs the image im")
    n = deconv_rl_device(projector, hMatrix, HtfT, maxIter, XT, im=imT, prevedel=prevedel, trust=trust,
                         targetFit=targetFit, fitOut=fitOut)
    out = projector.asnumpy(XT)
    if logPrint:
        print("Deconvolution elapsed wallclock time %f (%d iterations)" % (time.time() - t1, n))
    return out


def _deconv_generic(hMatrix, Htf, maxIter, Xguess, logPrint, numjobs, projector, im, prevedel, trustMask):
    """The reference loop through an arbitrary projector object
    (lfdeconv.py:86-196), CPU arrays."""
    if Htf is None:
        assert im is not None
        if prevedel or Xguess is None:
            Htf = projector.BackwardProjectACC(hMatrix, im, range(hMatrix.numZ), _noprogress, False, numjobs)
    Xguess = np.array(Htf if Xguess is None else Xguess, dtype=np.float32, copy=True)
    trustBack = None
    if trustMask is not None:
        trustBack = projector.BackwardProjectACC(hMatrix, np.asarray(trustMask, np.float32)[np.newaxis],
                                                 range(hMatrix.numZ), _noprogress, False, numjobs)
    for _ in range(maxIter):
        HX = projector.ForwardProjectACC(hMatrix, Xguess, range(hMatrix.numZ), _noprogress, False, numjobs)
        with np.errstate(divide="ignore", invalid="ignore"):
            if prevedel:
                back = projector.BackwardProjectACC(hMatrix, HX, range(hMatrix.numZ), _noprogress, False, numjobs)
                Xguess *= Htf / back
            else:
                back = projector.BackwardProjectACC(hMatrix, im / HX, range(hMatrix.numZ), _noprogress, False,
                                                    numjobs)
                Xguess *= back
                if trustBack is not None:
                    Xguess /= trustBack
        Xguess[np.isnan(Xguess)] = 0
    return Xguess
