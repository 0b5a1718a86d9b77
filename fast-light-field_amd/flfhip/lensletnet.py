"""Learned initialiser for Richardson-Lucy: a lenslet-aware conv net.

The network (libflf_hip.so, flf_net_*) maps the rectified image to a volume
guess at lenslet resolution (one value per plane and lenslet), which is then
upsampled to full resolution and used as Xguess.  Training pairs are made by
the method itself: seeded blob volumes are forward-projected through the same
PSF set with Poisson noise and reconstructed by plain RL; the network learns
the lenslet-block mean of that reconstruction.  The reference has no learned
component; plain RL from a uniform start stays the fallback (the guard in
``LearnedInit.initial_volume``).
"""
import ctypes
import os

import numpy as np
import torch

from . import _lib


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class LensletNet(object):
    def __init__(self, nnum, nz, width=128, nblocks=4, device=0, seed=0):
        self.lib = _lib.load()
        self.nnum, self.nz, self.width, self.nblocks = nnum, nz, width, nblocks
        self.device = torch.device("cuda", device)
        h = ctypes.c_void_p()
        rc = self.lib.flf_net_create(device, nnum, nz, width, nblocks, ctypes.byref(h))
        if rc != 0:
            raise _lib.FLFError("flf_net_create failed (%d)" % rc)
        self._h = h
        self.nparam = int(self.lib.flf_net_num_params(h))
        self.set_params(self.init_params(seed))

    def __del__(self):
        lib = getattr(self, "lib", None)
        h = getattr(self, "_h", None)
        if lib is not None and h:
            lib.flf_net_destroy(h)
            self._h = None

    def _check(self, rc, what):
        if rc != 0:
            msg = self.lib.flf_net_last_error(self._h)
            raise _lib.FLFError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))

    # ---- parameters: flat float32, per layer W[Co][9][Ci] then b[Co] ----
    def layer_shapes(self):
        shapes = [(self.nnum * self.nnum, self.width)]
        shapes += [(self.width, self.width)] * (2 * self.nblocks)
        shapes += [(self.width, self.nz)]
        return shapes

    def init_params(self, seed):
        """He-normal weights; the second conv of each residual block and the
        head start small so the initial map is close to identity/constant,
        and the head bias makes the initial output 1 (softplus(b) = 1), the
        mean voxel value in the network's normalised units."""
        rng = np.random.default_rng(seed)
        out = []
        shapes = self.layer_shapes()
        for li, (ci, co) in enumerate(shapes):
            std = np.sqrt(2.0 / (9 * ci))
            if li > 0 and li < len(shapes) - 1 and li % 2 == 0:
                std *= 0.1
            if li == len(shapes) - 1:
                std *= 0.1
            out.append(rng.normal(0.0, std, size=co * 9 * ci).astype(np.float32))
            b0 = np.log(np.e - 1.0) if li == len(shapes) - 1 else 0.0
            out.append(np.full(co, b0, np.float32))
        p = np.concatenate(out)
        assert p.size == self.nparam, (p.size, self.nparam)
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.nparam:
            raise ValueError("expected %d parameters, got %d" % (self.nparam, p.size))
        self._check(self.lib.flf_net_set_params(self._h, p.ctypes.data_as(ctypes.c_void_p)), "flf_net_set_params")

    def get_params(self):
        p = np.empty(self.nparam, np.float32)
        self._check(self.lib.flf_net_get_params(self._h, p.ctypes.data_as(ctypes.c_void_p)), "flf_net_get_params")
        return p

    def get_grads(self):
        p = np.empty(self.nparam, np.float32)
        self._check(self.lib.flf_net_get_grads(self._h, p.ctypes.data_as(ctypes.c_void_p)), "flf_net_get_grads")
        return p

    def save(self, path, **meta):
        np.savez(path, params=self.get_params(), nnum=self.nnum, nz=self.nz, width=self.width,
                 nblocks=self.nblocks, **meta)

    @classmethod
    def load(cls, path, device=0):
        d = np.load(path, allow_pickle=False)
        net = cls(int(d["nnum"]), int(d["nz"]), int(d["width"]), int(d["nblocks"]), device=device)
        net.set_params(d["params"])
        return net

    # ---- compute ----
    def forward(self, img):
        """img: (T, Ly*N, Lx*N) float32 cuda -> (T, nz, Ly, Lx) guess (volume units)."""
        T, Y, X = img.shape
        N = self.nnum
        if Y % N or X % N:
            raise ValueError("image size must be a multiple of Nnum")
        Ly, Lx = Y // N, X // N
        img = img.contiguous()
        out = torch.empty((T, self.nz, Ly, Lx), dtype=torch.float32, device=self.device)
        self._check(self.lib.flf_net_forward(self._h, _ptr(img), T, Ly, Lx, _ptr(out), _stream()), "flf_net_forward")
        return out

    def train_step(self, img, target, lr):
        T, Y, X = img.shape
        N = self.nnum
        loss = ctypes.c_double()
        self._check(self.lib.flf_net_train_step(self._h, _ptr(img.contiguous()), _ptr(target.contiguous()), T,
                                                Y // N, X // N, float(lr), ctypes.byref(loss), _stream()),
                    "flf_net_train_step")
        return loss.value

    def volume_guess(self, img):
        """Full-resolution Xguess (nz, T, Y, X) from the image."""
        g = self.forward(img)
        T, nz, Ly, Lx = g.shape
        N = self.nnum
        vol = torch.empty((nz, T, Ly * N, Lx * N), dtype=torch.float32, device=self.device)
        rc = self.lib.flf_net_upsample(_ptr(g), T, nz, Ly, Lx, N, _ptr(vol), _stream())
        if rc != 0:
            raise _lib.FLFError("flf_net_upsample failed (%d)" % rc)
        return vol


def lenslet_means(vol, nnum):
    """(nz, T, Y, X) -> (T, nz, Ly, Lx): mean over each lenslet block."""
    nz, T, Y, X = vol.shape
    v = vol.reshape(nz, T, Y // nnum, nnum, X // nnum, nnum).mean(dim=(3, 5))
    return v.permute(1, 0, 2, 3).contiguous()


def blob_volumes_gpu(nz, Y, X, T, nblobs_range, seed, device, background=0.01, sigma=(1.0, 3.0), amp=(0.5, 2.0)):
    """Seeded Gaussian-blob volumes (nz, T, Y, X) built on the GPU as sums of
    separable Gaussians (same statistics as synthetic.blob_volume)."""
    rng = np.random.default_rng(seed)
    vol = torch.full((nz, T, Y, X), background, dtype=torch.float32, device=device)
    zz = torch.arange(nz, device=device, dtype=torch.float32)
    yy = torch.arange(Y, device=device, dtype=torch.float32)
    xx = torch.arange(X, device=device, dtype=torch.float32)
    for t in range(T):
        nb = int(rng.integers(nblobs_range[0], nblobs_range[1] + 1))
        zc = torch.tensor(rng.uniform(0, nz, nb), dtype=torch.float32, device=device)
        yc = torch.tensor(rng.uniform(0, Y, nb), dtype=torch.float32, device=device)
        xc = torch.tensor(rng.uniform(0, X, nb), dtype=torch.float32, device=device)
        sg = torch.tensor(rng.uniform(sigma[0], sigma[1], nb), dtype=torch.float32, device=device)
        am = torch.tensor(rng.uniform(amp[0], amp[1], nb), dtype=torch.float32, device=device)
        gz = am[:, None] * torch.exp(-(zz[None] - zc[:, None]) ** 2 / (2 * sg[:, None] ** 2))   # (nb, nz)
        gy = torch.exp(-(yy[None] - yc[:, None]) ** 2 / (2 * sg[:, None] ** 2))                # (nb, Y)
        gx = torch.exp(-(xx[None] - xc[:, None]) ** 2 / (2 * sg[:, None] ** 2))                # (nb, X)
        for z in range(nz):
            # sum_i gz[i,z] gy[i,:]^T gx[i,:]  as one fp32 matmul
            vol[z, t] += (gy * gz[:, z:z + 1]).t().contiguous() @ gx
    return vol


def _frames(proj, hMatrix, T, Y, X, vseed, peaks, gen, nblobs):
    """Seeded blob volumes (nz, T, Y, X), their images Poisson(scale * H vol)
    at the given peak counts (T,), and the per-frame scale."""
    dev = proj.device
    vol = blob_volumes_gpu(hMatrix.numZ, Y, X, T, nblobs, vseed, dev)
    proj.prepare(hMatrix, T, Y, X)
    clean = proj.forward_device(hMatrix, vol)
    pk = torch.tensor(peaks, dtype=torch.float32, device=dev)
    scale = pk / clean.reshape(T, -1).amax(dim=1)
    im = torch.poisson(clean.clamp_min(0) * scale[:, None, None], generator=gen)
    return vol, im, scale


def make_pairs(proj, hMatrix, n, L, seed, peak=(500.0, 5000.0), iters=8, batch=16, nblobs=(50, 300), log=None):
    """n training pairs at L x L lenslets: (images (n, Y, X), targets
    (n, nz, L, L)) on the GPU, targets = lenslet means of plain RL after
    `iters` iterations from a uniform start."""
    from .lfdeconv import deconv_rl_device
    if n % batch:
        raise ValueError("the pair count must be a multiple of the batch (one projector plan per batch size)")
    N = hMatrix.Nnum
    Y = X = L * N
    nz = hMatrix.numZ
    dev = proj.device
    imgs = torch.empty((n, Y, X), dtype=torch.float32, device=dev)
    tgts = torch.empty((n, nz, L, L), dtype=torch.float32, device=dev)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    rng = np.random.default_rng(seed)
    for b0 in range(0, n, batch):
        T = min(batch, n - b0)
        vol, im, _ = _frames(proj, hMatrix, T, Y, X, seed * 100003 + b0, rng.uniform(peak[0], peak[1], T), gen,
                             nblobs)
        Htf = proj.backward_device(hMatrix, im)
        Xv = torch.ones_like(Htf)
        deconv_rl_device(proj, hMatrix, Htf, iters, Xv, im=im)
        imgs[b0:b0 + T] = im
        tgts[b0:b0 + T] = lenslet_means(Xv, N)
        del vol, im, Htf, Xv
        if log and (b0 // batch) % 16 == 0:
            log("pairs %d / %d" % (b0 + T, n))
    return imgs, tgts


def evaluate_heldout(net, proj, hMatrix, n, L, seed, iters=8, batch=16, peak=(500.0, 5000.0), nblobs=(50, 300),
                     max_iter=32):
    """Held-out evaluation of a trained network on the n frames that
    make_pairs(..., seed, batch=batch) would draw (same volumes, peaks and
    noise): per frame, plain Prevedel RL from a uniform start for `iters`
    iterations (its fit is the target), then the guarded learned start and
    RL until the fit reaches that target.  Returns per-frame arrays and a
    summary: guard acceptance, learned iterations to the target (the plain
    leg needs `iters`; a rejected guess counts its plain iterate as one),
    and the relative L2 of both results against the ground-truth volume."""
    from .lfdeconv import deconv_rl_device, _fit_from_sums
    if n % batch:
        raise ValueError("the frame count must be a multiple of the batch")
    Y = X = L * hMatrix.Nnum
    dev = proj.device
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    rng = np.random.default_rng(seed)
    li = LearnedInit(net, proj, hMatrix)
    acc, n_it, gt_p, gt_l, fit_t, fit_l = [], [], [], [], [], []

    def rel_truth(Xv, vol, scale):
        d = (Xv.double() / scale.double()[None, :, None, None] - vol.double()).reshape(vol.shape[0], vol.shape[1], -1)
        return (torch.linalg.vector_norm(d, dim=(0, 2)) /
                torch.linalg.vector_norm(vol.double().reshape(vol.shape[0], vol.shape[1], -1), dim=(0, 2)))

    for b0 in range(0, n, batch):
        T = min(batch, n - b0)
        vol, im, scale = _frames(proj, hMatrix, T, Y, X, seed * 100003 + b0, rng.uniform(peak[0], peak[1], T), gen,
                                 nblobs)
        Htf = proj.backward_device(hMatrix, im)
        Xp = torch.ones_like(Htf)
        deconv_rl_device(proj, hMatrix, Htf, iters, Xp, im=im)
        target = _fit_from_sums(proj.datafit_device(proj.forward_device(hMatrix, Xp), im))
        fits = []
        Xl, _, info = li.deconvolve(im, Htf, max_iter, targetFit=target, fitOut=fits)
        F = np.array([np.asarray(f, np.float64) for f in fits])          # (iterates, T)
        for t in range(T):
            k = np.nonzero(F[:, t] <= target[t])[0]
            n_it.append((int(k[0]) if len(k) else -1) + (0 if info["accepted"][t] else 1))
        acc.extend(bool(a) for a in info["accepted"])
        fit_t.extend(float(v) for v in target)
        fit_l.extend(float(v) for v in F[-1])
        gt_p.extend(rel_truth(Xp, vol, scale).tolist())
        gt_l.extend(rel_truth(Xl, vol, scale).tolist())
        del vol, im, Htf, Xp, Xl
    n_it = np.array(n_it)
    gt_p, gt_l = np.array(gt_p), np.array(gt_l)
    summ = lambda a: {"min": float(np.min(a)), "median": float(np.median(a)), "mean": float(np.mean(a)),
                      "max": float(np.max(a))}
    summary = {
        "frames": n, "lenslets": L, "seed": seed, "plain_iterations": iters,
        "accepted_fraction": float(np.mean(acc)),
        "learned_iterations": summ(n_it[n_it >= 0]) if np.any(n_it >= 0) else None,
        "learned_iterations_histogram": {str(int(k)): int(v) for k, v in zip(*np.unique(n_it, return_counts=True))},
        "not_reached": int(np.sum(n_it < 0)),
        "iterations_saved_mean": float(iters - np.mean(n_it[n_it >= 0])) if np.any(n_it >= 0) else None,
        "gt_rel_l2_plain": summ(gt_p), "gt_rel_l2_learned": summ(gt_l),
        "learned_better_gt_fraction": float(np.mean(gt_l < gt_p)),
    }
    return {"accepted": acc, "iterations": n_it.tolist(), "target_fit": fit_t, "learned_fit": fit_l,
            "gt_plain": gt_p.tolist(), "gt_learned": gt_l.tolist()}, summary


def dihedral(x, k):
    """One of the 8 flips/transposes of the last two axes.  The PSF set is
    symmetric under them jointly with the lenslet offsets (row/column mirror
    and transpose, py_light_field.cpp:1514-1562), so transforming an image
    and its RL volume alike gives another valid pair."""
    if k & 1:
        x = torch.flip(x, dims=(-2,))
    if k & 2:
        x = torch.flip(x, dims=(-1,))
    if k & 4:
        x = x.transpose(-1, -2)
    return x.contiguous()


def train(net, imgs, tgts, steps, batch=32, lr=1e-3, seed=0, log_every=50, log=print, warmup=200, augment=True):
    """Adam on random mini-batches (linear warm-up, then cosine decay; each
    batch under a random dihedral transform when augment and the frames are
    square); returns the loss history."""
    rng = np.random.default_rng(seed)
    n = imgs.shape[0]
    square = imgs.shape[-1] == imgs.shape[-2]
    hist = []
    for s in range(steps):
        idx = torch.tensor(rng.choice(n, size=min(batch, n), replace=False), device=imgs.device)
        lr_s = lr * min(1.0, (s + 1) / max(warmup, 1)) * 0.5 * (1 + np.cos(np.pi * s / max(steps, 1)))
        bi, bt = imgs[idx], tgts[idx]
        if augment:
            k = int(rng.integers(0, 8 if square else 4))
            bi, bt = dihedral(bi, k), dihedral(bt, k)
        loss = net.train_step(bi, bt, lr_s)
        hist.append(loss)
        if log and (s % log_every == 0 or s == steps - 1):
            log("step %d loss %.5f" % (s, loss))
    return hist


WEIGHTS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights")


def default_weights_path(nnum, nz, tag):
    return os.path.join(WEIGHTS_DIR, "lensletnet_%s_n%d_z%d.npz" % (tag, nnum, nz))


def _fit(projector, HX, im):
    from .lfdeconv import _fit_from_sums
    return _fit_from_sums(projector.datafit_device(HX, im))


class LearnedInit(object):
    """Guarded learned starting volume for Prevedel RL on one projector/HMatrix.

    The guess is kept for a frame only if its forward-model data fit is no
    worse than that of one plain RL iteration from the uniform start; else
    that plain iterate is used.  The plain iterate X1 = Htf / Ht(H 1) needs no
    projection once the image-independent Ht(H 1) is cached (one forward and
    one backward projection per HMatrix and frame shape)."""

    def __init__(self, net, projector, hMatrix):
        if net.nnum != hMatrix.Nnum or net.nz != hMatrix.numZ:
            raise ValueError("network (Nnum %d, %d planes) does not match the HMatrix (Nnum %d, %d planes)"
                             % (net.nnum, net.nz, hMatrix.Nnum, hMatrix.numZ))
        self.net, self.projector, self.hMatrix = net, projector, hMatrix
        self._den, self._den_key = None, None
        self._side = None

    def plain_denominator(self, Y, X):
        """Ht(H 1) for one frame, (numZ, 1, Y, X): the same for every frame of
        this size, so one slab per plane serves a whole batch."""
        key = (Y, X)
        if self._den_key != key:
            p, hm = self.projector, self.hMatrix
            ones = torch.ones((hm.numZ, 1, Y, X), dtype=torch.float32, device=p.device)
            H1 = p.forward_device(hm, ones)
            del ones
            self._den = p.backward_device(hm, H1)
            self._den_key = key
        return self._den

    def guess_and_htf(self, im, Htf_out=None):
        """The network's guess and Htf = Ht(im), computed side by side: the
        guess (which needs only the image) on a second stream while the back
        projection runs on the current one.  The network's small GEMMs then
        fill the back projection's FFT phases instead of following it.  Both
        results are the same as when computed one after the other."""
        p, hm = self.projector, self.hMatrix
        main = torch.cuda.current_stream(p.device)
        if self._side is None:
            self._side = torch.cuda.Stream(device=p.device)    # default priority
        self._side.wait_stream(main)                            # im and earlier work are ready
        with torch.cuda.stream(self._side):
            g = self.net.volume_guess(im)
        Htf = p.backward_device(hm, im, out=Htf_out)
        main.wait_stream(self._side)
        g.record_stream(main)                                   # g is used and freed on the current stream
        return g, Htf

    def initial_volume(self, im, Htf, guess=None):
        """im (T, Y, X), Htf (nz, T, Y, X) on the device.  Returns (X, HX,
        info): the starting volume, its forward projection and per-frame
        guard results (accepted, fit_guess, fit_plain1).  When the projector
        was planned for 2T frames (prepare(..., frames_capacity=2T)), the
        guess and the plain iterate are projected in one call, so every
        cached spectrum value read serves both.  guess: the network's guess
        for im if already computed (guess_and_htf)."""
        p, hm = self.projector, self.hMatrix
        T, Y, Xw = im.shape
        den = self.plain_denominator(Y, Xw)
        g = self.net.volume_guess(im) if guess is None else guess
        X1 = torch.empty_like(Htf)
        p.rl_ratio_device(X1, Htf, den)                        # X1 = Htf / Ht(H 1): one iteration from X = 1
        if getattr(p, "_tcap", 0) >= 2 * T:
            H2 = p.forward_pair_device(hm, g, X1)              # (2T, Y, X): both volumes, one call
            fits = _fit(p, H2, im)                              # frame t and T + t against im[t]
            fit_g, fit_1 = fits[:T], fits[T:]
            Hg, HX1 = H2[:T], H2[T:]
            del H2
        else:
            Hg = p.forward_device(hm, g)
            fit_g = _fit(p, Hg, im)
            HX1 = p.forward_device(hm, X1)
            fit_1 = _fit(p, HX1, im)
        accept = fit_g <= fit_1
        for t in np.nonzero(~accept)[0]:
            g[:, t] = X1[:, t]
            Hg[t] = HX1[t]
        return g, Hg, {"accepted": accept, "fit_guess": fit_g, "fit_plain1": fit_1}

    def deconvolve(self, im, Htf, maxIter, targetFit=None, fitOut=None, Htf_out=None, minIter=0):
        """Guarded learned start, then Prevedel RL until every frame's fit is
        <= targetFit (or maxIter iterations), after at least minIter
        iterations.  Returns (X, iterations, info); a rejected frame's plain
        iterate counts as one iteration.  Htf None: Htf = Ht(im) is computed
        here (into Htf_out if given) beside the network's guess
        (guess_and_htf) and returned as info["Htf"].  fitOut receives the fit
        of the starting volume and of every later iterate (per frame)."""
        from .lfdeconv import deconv_rl_device
        guess = None
        if Htf is None:
            guess, Htf = self.guess_and_htf(im, Htf_out)
        X, HX, info = self.initial_volume(im, Htf, guess=guess)
        info["Htf"] = Htf
        fit0 = np.where(info["accepted"], info["fit_guess"], info["fit_plain1"])   # fit of the starting X
        n = deconv_rl_device(self.projector, self.hMatrix, Htf, maxIter, X, im=im, targetFit=targetFit,
                             fitOut=fitOut, HX0=HX, fit0=fit0, minIter=minIter)
        info["iterations"] = n + (0 if info["accepted"].all() else 1)
        return X, info["iterations"], info
