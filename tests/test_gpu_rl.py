"""Richardson-Lucy on the GPU (lfdeconv.DeconvRL through Projector_hip) against
the FP64 oracle: Prevedel and standard updates, with and without a trust
mask, several iterations and frames; 1e-4 relative L2 on the volume."""
import numpy as np
import pytest

from flfhip import synthetic
from oracle import projector as oproj
from oracle import rl as orl

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def problem():
    N = 15
    zs = [-2, 0, 3]
    sides = synthetic.psf_sides(zs, 15, 45)
    hm = synthetic.make_psf_set(N, zs, sides, seed=21)
    vol = synthetic.blob_volume(len(zs), 75, 90, 30, seed=4, T=2)
    fwd = lambda v: oproj.forward_project(v, [hm.Hcc(z, False) for z in range(len(zs))])
    bwd = lambda i: oproj.backward_project(i, [hm.Hcc(z, True) for z in range(len(zs))])
    clean = fwd(vol.astype(np.float64))
    im = np.stack([synthetic.poisson_image(clean[t:t + 1], 800.0, seed=t)[0] for t in range(2)]).astype(np.float32)
    return hm, im, fwd, bwd


@pytest.mark.parametrize("prevedel,use_trust,iters", [(True, False, 4), (False, False, 3), (False, True, 3)])
def test_deconv_rl_vs_oracle(projector, problem, prevedel, use_trust, iters):
    from flfhip import lfdeconv
    hm, im, fwd, bwd = problem
    print("RL prevedel=%s trust=%s iters=%d" % (prevedel, use_trust, iters), flush=True)
    trust = None
    im_used = im
    if use_trust:
        # lfdeconv.py:106-110: the image must be zero wherever the mask is.
        # The untrusted band (3 columns) is narrower than the smallest PSF
        # half-width (7 px), so every voxel sees trusted pixels and the
        # reference update X *= back / trustBack is well defined (a voxel with
        # trustBack = 0 would divide by zero in the reference too).
        trust = np.ones(im.shape[1:], np.float32)
        trust[:, :3] = 0
        im_used = im * trust[None]
        tb = bwd(trust[None].astype(np.float64))
        assert tb.min() > 1e-3 * tb.max()
    # Xguess None: the reference starts from Htf (lfdeconv.py:97-99)
    X = lfdeconv.DeconvRL(hm, None, iters, None, logPrint=False, projector=projector, im=im_used,
                          prevedel=prevedel, trustMask=trust)
    Xr = orl.deconv_rl(fwd, bwd, None, iters, Xguess=None, im=im_used, prevedel=prevedel, trust_mask=trust)
    err = rel_l2(X, Xr)
    print("rel-L2 vs FP64 oracle %.3e" % err, flush=True)
    assert X.shape == Xr.shape
    assert err < 1e-4


def test_fit_trace_matches_oracle(projector, problem):
    """The data fit reported per iteration (FP64 sums on the GPU) against the
    oracle's fit of the same iterates."""
    from flfhip import lfdeconv
    hm, im, fwd, bwd = problem
    fits = []
    lfdeconv.DeconvRL(hm, None, 3, None, logPrint=False, projector=projector, im=im, fitOut=fits)
    Xr, rec = orl.deconv_rl(fwd, bwd, None, 3, im=im, record_fit=True)
    # the GPU run also reports the fit of the last iterate (one more entry)
    got = np.array([np.asarray(f, np.float64) for f in fits[:3]])
    ref = np.array(rec)
    print("fits", got.tolist(), ref.tolist(), flush=True)
    assert got.shape == ref.shape
    assert np.max(np.abs(got - ref) / ref) < 1e-5


def test_cfg1_full_rl_vs_oracle(projector):
    """The CPU-runnable reference case at full size: 20 x 20 lenslets (300 x
    300), 11 planes z = -5..5 with ring PSFs 15..135 px, 200 blobs, Poisson
    peak 1000, 4 Prevedel iterations from a uniform start, against the FP64
    oracle (volume within 1e-4 relative L2, every iterate's data fit within
    1e-5)."""
    from flfhip import lfdeconv, workloads
    w = workloads.setup("rl300")
    hm = w["hm"]
    nz = hm.numZ
    Qf = [hm.Hcc(z, False) for z in range(nz)]
    Qb = [hm.Hcc(z, True) for z in range(nz)]
    fwd = lambda v: oproj.forward_project(v, Qf)
    bwd = lambda i: oproj.backward_project(i, Qb)
    clean = projector.ForwardProjectACC(hm, w["vol"], range(nz))
    im = synthetic.poisson_image(clean, w["peak"], w["seed"])
    X0 = np.ones((nz,) + im.shape, np.float32)
    fits = []
    X = lfdeconv.DeconvRL(hm, None, w["iters"], X0, logPrint=False, projector=projector, im=im, fitOut=fits)
    Xr, rec = orl.deconv_rl(fwd, bwd, None, w["iters"], Xguess=X0, im=im, record_fit=True)
    err = rel_l2(X, Xr)
    got = np.array([np.asarray(f, np.float64) for f in fits[:w["iters"]]])
    print("cfg1 RL(4) rel-L2 %.3e; fits gpu %s oracle %s" % (err, got.ravel().tolist(), np.ravel(rec).tolist()),
          flush=True)
    assert err < 1e-4
    assert np.max(np.abs(got - np.array(rec)) / np.array(rec)) < 1e-5


def test_graph_iterations_match_eager(projector):
    """RL iterations replayed from a captured HIP graph equal the eager ones
    bit for bit, on every replay, with other GPU work between replays."""
    import torch
    from flfhip.lfdeconv import RLGraph, deconv_rl_device
    N = 15
    zs = [-1, 0, 2]
    sides = synthetic.psf_sides(zs, 15, 31)
    hm = synthetic.make_psf_set(N, zs, sides, seed=5)
    vol = synthetic.blob_volume(len(zs), 90, 75, 20, seed=6)
    projector.prepare(hm, 1, 90, 75)
    im = projector.forward_device(hm, projector.asnative(vol))
    Htf = projector.backward_device(hm, im)
    X0 = Htf.clone()
    Xe = X0.clone()
    deconv_rl_device(projector, hm, Htf, 3, Xe)
    Xg = X0.clone()
    g = RLGraph(projector, hm, Htf, Xg)
    for rep in range(3):
        Xg.copy_(X0)
        g.run(3)
        junk = projector.forward_device(hm, projector.asnative(vol)) * 2.0      # eager work between replays
        torch.cuda.synchronize()
        diff = float((Xg - Xe).abs().max())
        print("replay", rep, "max abs diff", diff, flush=True)
        assert diff == 0.0
        del junk
