"""Richardson-Lucy deconvolution in FP64 (oracle; test use only).

Restates lfdeconv.DeconvRL (lfdeconv.py:72-202) with the projector passed in
as two callables, and adds the forward-model data-fit this library uses as its
convergence measure (the reference has none; it always runs maxIter).
"""
import numpy as np


def rl_update_prevedel(X, Htf, back):
    """RLUpdateFuncPrevedel (lfdeconv.py:112-126): X *= Htf / back, NaN -> 0."""
    with np.errstate(divide="ignore", invalid="ignore"):
        X = X * (Htf / back)
    X[np.isnan(X)] = 0
    return X


def rl_update_standard(X, back, trust_back=None):
    """RLUpdateFuncStandard (lfdeconv.py:151-164): X *= back [/ trustBack], NaN -> 0."""
    with np.errstate(divide="ignore", invalid="ignore"):
        X = X * back
        if trust_back is not None:
            X = X / trust_back
    X[np.isnan(X)] = 0
    return X


def data_fit(forward_image, im):
    """rho = ||F X - im||_2 / ||im||_2 per frame (FP64 sums)."""
    d = np.asarray(forward_image, np.float64) - np.asarray(im, np.float64)
    num = np.sqrt(np.sum(d.reshape(d.shape[0], -1) ** 2, axis=1))
    den = np.sqrt(np.sum(np.asarray(im, np.float64).reshape(d.shape[0], -1) ** 2, axis=1))
    return num / den


def deconv_rl(forward, backward, Htf, max_iter, Xguess=None, im=None,
              prevedel=True, trust_mask=None, record_fit=False):
    """lfdeconv.py:72-202.  forward: vol -> image; backward: image -> vol."""
    if Htf is None:
        assert im is not None
        if prevedel or Xguess is None:
            Htf = backward(im)
    X = np.array(Htf if Xguess is None else Xguess, dtype=np.float64, copy=True)
    trust_back = None
    if trust_mask is not None:
        trust_back = backward(np.asarray(trust_mask, np.float64)[None])
    fits = []
    for _ in range(max_iter):
        HX = forward(X)
        if record_fit and im is not None:
            fits.append(data_fit(HX, im))
        if prevedel:
            X = rl_update_prevedel(X, Htf, backward(HX))
        else:
            with np.errstate(divide="ignore", invalid="ignore"):
                err = np.asarray(im, np.float64) / HX
            X = rl_update_standard(X, backward(err), trust_back)
    if record_fit:
        return X, fits
    return X
