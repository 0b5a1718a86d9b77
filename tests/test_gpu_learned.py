"""The guarded learned start (LearnedInit) on the GPU: the guard falls back to
plain RL exactly, and the packaged trained network saves RL iterations."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _problem(proj, name, lenslets, seed, peak):
    from flfhip import synthetic, workloads
    hm, zs, sides = workloads.psf_set(name)
    N = hm.Nnum
    Y = X = N * lenslets
    proj.prepare(hm, 1, Y, X, frames_capacity=2)
    vol = synthetic.blob_volume(hm.numZ, Y, X, 150, seed=seed)
    clean = proj.forward_device(hm, proj.asnative(vol))
    im = proj.asnative(synthetic.poisson_image(proj.asnumpy(clean), peak, seed))
    Htf = proj.backward_device(hm, im)
    return hm, im, Htf


def _plain(proj, hm, Htf, im, iters):
    from flfhip.lfdeconv import deconv_rl_device
    X = torch.ones_like(Htf)
    fits = []
    deconv_rl_device(proj, hm, Htf, iters, X, im=im, fitOut=fits)
    return X, [float(f[0]) for f in fits]


def test_guard_rejects_bad_guess_and_matches_plain(projector):
    """A network whose guess is ~0 (head bias -30) fails the guard; the run
    then continues from the first plain iterate and reproduces plain RL."""
    from flfhip.lensletnet import LearnedInit, LensletNet
    from flfhip.lfdeconv import deconv_rl_device
    hm, im, Htf = _problem(projector, "rl300", 20, seed=11, peak=1000.0)
    net = LensletNet(hm.Nnum, hm.numZ, width=32, nblocks=1, seed=1)
    p = net.init_params(1)
    p[-hm.numZ:] = -30.0
    net.set_params(p)
    li = LearnedInit(net, projector, hm)
    Xp, fits_p = _plain(projector, hm, Htf, im, 4)
    # The guard projects guess and plain iterate as one two-frame call (two
    # frames per MAC workgroup), the plain run one frame at a time: the same
    # values up to FP32 summation order, hence the 1e-6 margins.
    Xl, n, info = li.deconvolve(im, Htf, 4, targetFit=fits_p[4] * (1 + 1e-6))
    print("guard", info, "plain fits", fits_p, flush=True)
    assert not info["accepted"][0]
    assert info["fit_guess"][0] > info["fit_plain1"][0]
    assert abs(info["fit_plain1"][0] - fits_p[1]) <= 1e-6 * fits_p[1]
    assert n == 4
    assert rel_l2(Xl.cpu().numpy(), Xp.cpu().numpy()) < 1e-5


def test_trained_network_saves_iterations(projector):
    """The packaged network (trained on 600x600 pairs through the 41-plane
    ring PSF set) on a held-out 600x600 frame: its guess passes the guard and
    RL from it reaches the data fit of 8 plain iterations in at most 3."""
    from flfhip.lensletnet import LearnedInit, LensletNet, default_weights_path
    hm, im, Htf = _problem(projector, "rl1455", 40, seed=12345, peak=2000.0)
    path = default_weights_path(hm.Nnum, hm.numZ, "rl1455")
    assert os.path.exists(path), path
    net = LensletNet.load(path)
    Xp, fits_p = _plain(projector, hm, Htf, im, 8)
    li = LearnedInit(net, projector, hm)
    fits_l = []
    Xl, n, info = li.deconvolve(im, Htf, 8, targetFit=fits_p[8], fitOut=fits_l)
    print("plain", fits_p, "learned", [float(f[0]) for f in fits_l], info, flush=True)
    assert info["accepted"][0]
    assert n <= 3
    assert float(fits_l[-1][0]) <= fits_p[8]


def test_heldout_evaluation_packaged_network(projector):
    """lensletnet.evaluate_heldout on 32 unseen 600x600 frames of the packaged
    network's own distribution (blob volumes, peaks 500..5000): the guard
    accepts the guess on most frames, RL from it reaches the 8-iteration
    plain fit in fewer iterations on average, and the per-frame counts agree
    with the summary."""
    from flfhip import workloads
    from flfhip.lensletnet import LensletNet, default_weights_path, evaluate_heldout
    hm, zs, sides = workloads.psf_set("rl1455")
    net = LensletNet.load(default_weights_path(hm.Nnum, hm.numZ, "rl1455"))
    per, summ = evaluate_heldout(net, projector, hm, 32, 40, seed=777, iters=8, batch=16)
    print(summ, flush=True)
    assert summ["frames"] == 32 and len(per["iterations"]) == 32
    assert summ["accepted_fraction"] >= 0.75
    assert summ["not_reached"] == 0
    assert summ["iterations_saved_mean"] > 2.0
    assert all(0 <= k <= 9 for k in per["iterations"])
    for t in range(32):
        assert per["learned_fit"][t] <= per["target_fit"][t] * (1 + 1e-6)
    assert np.isfinite(per["gt_learned"]).all() and np.isfinite(per["gt_plain"]).all()


def test_packaged_31_plane_network_heldout(projector):
    """The packaged 31-plane network (PSF set of the ragged time series,
    PSF 15..255 px; trained on 20k 600x600 pairs of seed 4 against
    8-iteration targets): on the first 32 of its 512 held-out frames (seed
    5) the guard keeps the guess and RL from it reaches the 8-iteration
    plain fit in at most 4 iterations."""
    from flfhip import workloads
    from flfhip.lensletnet import LensletNet, default_weights_path, evaluate_heldout
    hm, zs, sides = workloads.psf_set("ragged64")
    path = default_weights_path(hm.Nnum, hm.numZ, "ragged64")
    assert os.path.exists(path), path
    net = LensletNet.load(path)
    per, summ = evaluate_heldout(net, projector, hm, 32, 40, seed=5, iters=8, batch=32)
    print(summ, flush=True)
    assert summ["accepted_fraction"] >= 0.9
    assert summ["not_reached"] == 0 and max(per["iterations"]) <= 4
    assert summ["iterations_saved_mean"] >= 4.0


def test_guess_beside_htf_matches_sequential(projector):
    """Htf computed on the current stream while the network's guess runs on a
    second one: the same Htf, guess and RL result as one after the other,
    over repeated calls (the buffers of one call are reused by the next)."""
    from flfhip.lensletnet import LearnedInit, LensletNet
    hm, im, Htf = _problem(projector, "rl300", 20, seed=13, peak=1000.0)
    net = LensletNet(hm.Nnum, hm.numZ, width=32, nblocks=1, seed=2)
    li = LearnedInit(net, projector, hm)
    g_seq = net.volume_guess(im)
    for rep in range(3):
        out = torch.full_like(Htf, float("nan"))
        g, H2 = li.guess_and_htf(im, Htf_out=out)
        torch.cuda.synchronize()
        print("guess beside Htf, call", rep, flush=True)
        assert H2.data_ptr() == out.data_ptr()
        assert torch.equal(H2, Htf) and torch.equal(g, g_seq)
    Xa, na, ia = li.deconvolve(im, Htf, 3)
    Xb, nb, ib = li.deconvolve(im, None, 3)
    torch.cuda.synchronize()
    assert na == nb and torch.equal(Xa, Xb) and torch.equal(ib["Htf"], Htf)


@pytest.mark.parametrize("packaged", [True, False])
def test_two_frame_step_matches_single_frames(projector, packaged):
    """A step of two frames (the guard's pair projection then holds two
    frames per source) gives each frame the guard decision, data fits and
    volume of a one-frame step: with the packaged network (guess kept) and
    with one whose guess is ~0 (plain iterate kept)."""
    from flfhip import synthetic, workloads
    from flfhip.lensletnet import LearnedInit, LensletNet, default_weights_path
    hm, zs, sides = workloads.psf_set("rl1455")
    N = hm.Nnum
    Y = X = N * 40
    if packaged:
        net = LensletNet.load(default_weights_path(N, hm.numZ, "rl1455"))
    else:
        net = LensletNet(N, hm.numZ, width=32, nblocks=1, seed=3)
        p = net.init_params(3)
        p[-hm.numZ:] = -30.0
        net.set_params(p)
    li = LearnedInit(net, projector, hm)
    projector.prepare(hm, 2, Y, X, frames_capacity=4)
    vol = synthetic.blob_volume(hm.numZ, Y, X, 150, seed=31, T=2)
    clean = projector.forward_device(hm, projector.asnative(vol))
    im2 = projector.asnative(synthetic.poisson_image(projector.asnumpy(clean), 2000.0, 31))
    X2, n2, info2 = li.deconvolve(im2, None, 2)
    X2 = X2.clone()
    torch.cuda.synchronize()
    print("two-frame step", packaged, info2["accepted"], info2["fit_guess"], info2["fit_plain1"], flush=True)
    assert bool(np.all(info2["accepted"])) == packaged
    for t in range(2):
        projector.prepare(hm, 1, Y, X, frames_capacity=2)
        X1, n1, info1 = li.deconvolve(im2[t:t + 1].contiguous(), None, 2)
        torch.cuda.synchronize()
        assert bool(info1["accepted"][0]) == bool(info2["accepted"][t])
        for k in ("fit_guess", "fit_plain1"):
            assert abs(info1[k][0] - info2[k][t]) <= 1e-6 * info1[k][0], (t, k)
        assert n1 == n2
        assert rel_l2(X2[:, t].cpu().numpy(), X1[:, 0].cpu().numpy()) < 1e-5, t
