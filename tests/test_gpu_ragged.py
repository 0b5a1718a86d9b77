"""Frames of uneven size in one batch (flf_set_frame_dims, flfhip.ragged)
against the FP64 oracle run on each frame at its own size: projections
within 1e-5 relative L2 per frame, RL volumes within 1e-4."""
import ctypes

import numpy as np
import pytest

from flfhip import synthetic
from oracle import projector as oproj
from oracle import rl as orl

pytestmark = pytest.mark.gpu

SHAPES = [(90, 105), (60, 75), (52, 67), (75, 45)]


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def case():
    N, zs = 15, [-2, 0, 3]
    hm = synthetic.make_psf_set(N, zs, synthetic.psf_sides(zs, 15, 45), seed=21)
    Qf = [hm.Hcc(z, False) for z in range(3)]
    Qb = [hm.Hcc(z, True) for z in range(3)]
    fwd = lambda v: oproj.forward_project(v, Qf)
    bwd = lambda i: oproj.backward_project(i, Qb)
    vols = [synthetic.blob_volume(3, y, x, 12, seed=40 + t)[:, 0] for t, (y, x) in enumerate(SHAPES)]
    ims = [synthetic.poisson_image(fwd(v[:, None].astype(np.float64)), 800.0, seed=t)[0]
           for t, v in enumerate(vols)]
    return hm, fwd, bwd, vols, ims


def test_ragged_projections_vs_oracle(projector, case):
    from flfhip.ragged import RaggedBatch
    hm, fwd, bwd, vols, ims = case
    rb = RaggedBatch(hm, SHAPES, nbuckets=1, projectors=[projector])
    print("ragged projections, canvas", rb.buckets[0][0], flush=True)
    imgs = rb.forward(vols)
    back = rb.backward(ims)
    for t, (y, x) in enumerate(SHAPES):
        ef = rel_l2(imgs[t], fwd(vols[t][:, None].astype(np.float64))[0])
        eb = rel_l2(back[t], bwd(ims[t][None].astype(np.float64))[:, 0])
        print("frame %d %dx%d fwd %.2e bwd %.2e" % (t, y, x, ef, eb), flush=True)
        assert imgs[t].shape == (y, x) and back[t].shape == (3, y, x)
        assert ef < 1e-5 and eb < 1e-5


def test_ragged_outputs_zero_outside_frames(projector, case):
    """The raw batch: forward images and back-projected volumes are zero
    outside each frame's region, whatever the padding held."""
    import torch
    hm, fwd, bwd, vols, ims = case
    Yc, Xc = 90, 105
    projector.prepare(hm, len(SHAPES), Yc, Xc)
    vol = torch.rand((3, len(SHAPES), Yc, Xc), device=projector.device) + 0.5   # garbage in the padding
    img = torch.rand((len(SHAPES), Yc, Xc), device=projector.device) + 0.5
    projector.set_frame_dims(SHAPES)
    try:
        out_i = projector.forward_device(hm, vol)
        out_v = projector.backward_device(hm, img)
        torch.cuda.synchronize()
    finally:
        projector.set_frame_dims(None)
    for t, (y, x) in enumerate(SHAPES):
        m = torch.ones((Yc, Xc), dtype=torch.bool, device=projector.device)
        m[:y, :x] = False
        assert float(out_i[t][m].abs().max() if m.any() else 0.0) == 0.0
        assert float(out_v[:, t][:, m].abs().max() if m.any() else 0.0) == 0.0
        # inside: the projection of the frame's own region only
        vt = vol[:, t, :y, :x].cpu().numpy().astype(np.float64)[:, None]
        assert rel_l2(out_i[t, :y, :x].cpu().numpy(), fwd(vt)[0]) < 1e-5


def test_full_size_dims_are_uniform_batch(projector, case):
    """Frame sizes equal to the canvas give bit-identical results to no sizes."""
    import torch
    hm = case[0]
    projector.prepare(hm, 2, 90, 105)
    vol = torch.rand((3, 2, 90, 105), device=projector.device)
    a = projector.forward_device(hm, vol)
    projector.set_frame_dims([(90, 105), (90, 105)])
    b = projector.forward_device(hm, vol)
    projector.set_frame_dims(None)
    torch.cuda.synchronize()
    assert torch.equal(a, b)


@pytest.mark.parametrize("prevedel,nbuckets", [(True, 1), (False, 1), (True, 2)])
def test_ragged_rl_vs_oracle(projector, case, prevedel, nbuckets):
    from flfhip.ragged import RaggedBatch
    hm, fwd, bwd, vols, ims = case
    iters = 3
    # no per-bucket charge, so that two buckets are used when allowed
    rb = RaggedBatch(hm, SHAPES, nbuckets=nbuckets, projectors=[projector], spectrum_weight=0.0)
    print("ragged RL prevedel=%s buckets %s" % (prevedel, [(c, i) for c, i in rb.buckets]), flush=True)
    assert len(rb.buckets) == nbuckets
    fits = []
    Xs = rb.deconv(ims, iters, prevedel=prevedel, fitOut=fits)
    for t, (y, x) in enumerate(SHAPES):
        Xr, rec = orl.deconv_rl(fwd, bwd, None, iters, Xguess=None, im=ims[t][None], prevedel=prevedel,
                                record_fit=True)
        err = rel_l2(Xs[t], Xr[:, 0])
        print("frame %d %dx%d RL(%d) rel-L2 %.2e" % (t, y, x, iters, err), flush=True)
        assert Xs[t].shape == (3, y, x)
        assert err < 1e-4
        # per-iteration data fit of this frame, as the oracle reports it
        for idx, fl in fits:
            if t in idx:
                got = np.array([f[idx.index(t)] for f in fl[:iters]])
                ref = np.ravel(rec)[:iters]
                assert np.max(np.abs(got - ref) / ref) < 1e-5


def test_ragged_graph_matches_eager(projector, case):
    from flfhip.ragged import RaggedBatch
    hm, fwd, bwd, vols, ims = case
    rb = RaggedBatch(hm, SHAPES, nbuckets=1, projectors=[projector])
    a = rb.deconv(ims, 3)
    b = rb.deconv(ims, 3, graph=True)
    for t in range(len(SHAPES)):
        assert np.array_equal(a[t], b[t])


def test_set_frame_dims_rejects(projector, case):
    from flfhip import _lib
    hm = case[0]
    projector.prepare(hm, 2, 90, 105)
    lib, ctx = projector.lib, projector._ctx
    bad = [(0, 10), (91, 10), (10, 106), (10, 0)]
    for d in bad:
        arr = (ctypes.c_int * 4)(d[0], d[1], 90, 105)
        assert lib.flf_set_frame_dims(ctx, arr, 2, None) == _lib.FLF_EINVAL
    arr = (ctypes.c_int * 2)(60, 60)
    assert lib.flf_set_frame_dims(ctx, arr, 1, None) == _lib.FLF_EINVAL     # 1 size for 2 active frames
    assert lib.flf_set_frame_dims(ctx, None, 2, None) == 0


def test_ragged_series_frames_full_size_vs_oracle(projector):
    """The uneven time series at its real frame sizes: the first four frames
    of the 64-frame series (lenslet grids drawn from 20..60, i.e. 300..900
    px, by the series' own seeds), the series' ring PSFs of planes z = -15
    (255 px) and z = 0, two canvas buckets (the 570 px frame alone, the
    other three on a 795 px canvas); Htf and one Prevedel iteration per
    frame against the FP64 oracle at the frame's own size (1e-4)."""
    from flfhip import workloads
    from flfhip.psfmatrix import MemoryHMatrix
    from flfhip.ragged import RaggedBatch
    w = workloads.setup_ragged("ragged64", volume=False)
    shapes = w["shapes"][:4]
    full = w["hm"]
    planes = [0, w["zs"].index(0)]
    hm = MemoryHMatrix([full.Hcc(z, False) for z in planes], [full.Hcc(z, True) for z in planes])
    Qf = [hm.Hcc(z, False) for z in range(2)]
    Qb = [hm.Hcc(z, True) for z in range(2)]
    fwd = lambda v: oproj.forward_project(v, Qf)
    bwd = lambda i: oproj.backward_project(i, Qb)
    ims = []
    for t, (y, x) in enumerate(shapes):
        v = synthetic.blob_volume(2, y, x, 120, seed=300 + t)
        ims.append(synthetic.poisson_image(fwd(v.astype(np.float64)), 1000.0, seed=t)[0])
    rb = RaggedBatch(hm, shapes, nbuckets=2, spectrum_weight=0.5)
    print("series frames", shapes, "buckets", [(b[0], len(b[1])) for b in rb.buckets], flush=True)
    assert len(rb.buckets) == 2
    got = rb.deconv(ims, 1)
    for t, (y, x) in enumerate(shapes):
        ref = orl.deconv_rl(fwd, bwd, None, 1, im=ims[t][None].astype(np.float64))[:, 0]
        e = rel_l2(got[t], ref)
        print("frame %d %dx%d: RL (Htf + 1 iteration) rel-L2 %.2e" % (t, y, x, e), flush=True)
        assert got[t].shape == (2, y, x)
        assert e < 1e-4
