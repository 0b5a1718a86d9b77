"""The 32 x 1455^2 batch configuration (learned start against plain RL at full
size, frames reconstructed together) on its first four frames, through the
C ABI: each frame's guard decision and fits equal the frame's single-frame
run, every frame reaches its own 8-iteration plain target, and frame 0's
learned start refined by Prevedel RL matches the FP64 oracle on a two-plane
restriction (relative L2 <= 1e-4 on the volume)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

T = 4


@pytest.fixture(scope="module")
def batch(projector):
    from flfhip import lensletnet, workloads
    from flfhip.lfdeconv import deconv_rl_device, _fit_from_sums
    w = workloads.setup_batch("rl1455x32")
    hm, Y, X = w["hm"], w["Y"], w["X"]
    nz = hm.numZ
    dev = projector.device
    projector.prepare(hm, T, Y, X, frames_capacity=2 * T)
    vol = lensletnet.blob_volumes_gpu(nz, Y, X, T, w["nblobs"], w["seed"], dev)   # frames 0..3 of the batch
    clean = projector.forward_device(hm, vol)
    scale = w["peak"] / clean.reshape(T, -1).amax(dim=1)
    gen = torch.Generator(device=dev)
    gen.manual_seed(w["seed"])
    im = torch.poisson(clean.clamp_min(0) * scale[:, None, None], generator=gen)
    del clean, vol
    Htf = projector.backward_device(hm, im)
    Xp = torch.ones_like(Htf)
    deconv_rl_device(projector, hm, Htf, w["iters"], Xp)
    target = _fit_from_sums(projector.datafit_device(projector.forward_device(hm, Xp), im))
    del Xp
    path = lensletnet.default_weights_path(hm.Nnum, nz, w["base"])
    assert os.path.exists(path), path
    li = lensletnet.LearnedInit(lensletnet.LensletNet.load(path), projector, hm)
    return w, hm, im, target, li


def test_batch_frames_match_single_frame_runs(projector, batch):
    from flfhip.lfdeconv import _fit_from_sums
    w, hm, im, target, li = batch
    fits = []
    Xb, n, info = li.deconvolve(im, None, 32, targetFit=target, fitOut=fits)
    final = _fit_from_sums(projector.datafit_device(projector.forward_device(hm, Xb), im))
    print("batch: iterations %d, accepted %s, fit_guess %s, fit_plain1 %s, final %s, target %s"
          % (n, info["accepted"], info["fit_guess"], info["fit_plain1"], final, target), flush=True)
    assert np.all(final <= target * (1 + 1e-6))
    for t in range(T):
        one = im[t:t + 1].contiguous()
        X1, n1, info1 = li.deconvolve(one, None, 32, targetFit=target[t:t + 1])
        print("frame %d alone: iterations %d accepted %s fit_guess %.6f fit_plain1 %.6f"
              % (t, n1, info1["accepted"][0], info1["fit_guess"][0], info1["fit_plain1"][0]), flush=True)
        for k in ("fit_guess", "fit_plain1"):
            assert abs(info1[k][0] - info[k][t]) <= 1e-5 * info[k][t], (t, k)
        margin = abs(info["fit_guess"][t] - info["fit_plain1"][t]) / info["fit_plain1"][t]
        if margin > 1e-4:
            assert bool(info1["accepted"][0]) == bool(info["accepted"][t])
        assert n1 <= n
        e = float(torch.linalg.vector_norm((X1[:, 0] - Xb[:, t]).double()) / torch.linalg.vector_norm(Xb[:, t].double()))
        if n1 == n and bool(info1["accepted"][0]) == bool(info["accepted"][t]):
            assert e < 1e-4, (t, e)


def test_batch_frame0_vs_fp64_oracle_two_planes(projector, batch):
    """Frame 0's learned starting volume on planes z = -20 (345 px PSF) and
    z = 0 (15 px), Htf and one Prevedel iteration on the GPU against the FP64
    oracle from the same start."""
    from flfhip.lfdeconv import deconv_rl_device
    from oracle import projector as oproj
    from oracle import rl as orl
    w, hm, im, target, li = batch
    planes = [0, 20]
    Y, X = w["Y"], w["X"]
    projector.prepare(hm, 1, Y, X)
    im0 = im[:1].contiguous()
    g0, _, info = li.initial_volume(im0, projector.backward_device(hm, im0))
    start = g0[planes].clone()
    Htf = projector.backward_device(hm, im0, planes=planes)
    Xv = torch.zeros_like(g0)
    Xv[planes] = start
    deconv_rl_device(projector, hm, Htf, 1, Xv, planes=planes)
    Qf = [hm.Hcc(z, False) for z in planes]
    Qb = [hm.Hcc(z, True) for z in planes]
    Xr = orl.deconv_rl(lambda v: oproj.forward_project(v, Qf), lambda i: oproj.backward_project(i, Qb), None, 1,
                       Xguess=start.cpu().numpy().astype(np.float64), im=im0.cpu().numpy().astype(np.float64))
    Xg = Xv.cpu().numpy()
    for k, z in enumerate(planes):
        e = float(np.linalg.norm(Xg[z, 0] - Xr[k, 0]) / np.linalg.norm(Xr[k, 0]))
        print("frame 0 plane %d (PSF %d px): rel-L2 vs FP64 oracle %.2e" % (z, hm.PSFShape(z)[0], e), flush=True)
        assert e < 1e-4
