"""GPU tests of the C ABI's argument checks and of the fused standard-RL
ratio, the per-plane denominator broadcast, the data fit against fewer image
frames, and captured-graph invalidation (through the C ABI)."""
import ctypes

import numpy as np
import pytest

from flfhip import _lib, synthetic

pytestmark = pytest.mark.gpu


def _hm():
    N, zs = 15, [-2, 0, 3]
    return synthetic.make_psf_set(N, zs, synthetic.psf_sides(zs, 15, 45), seed=12)


def test_host_pointer_refused_before_launch(projector):
    """A host array handed to an entry point returns FLF_EINVAL; nothing is
    queued.  The check itself is asserted first (host-only call), so the
    projection below is only reached with a pointer the check refuses."""
    import torch
    lib = projector.lib
    hm = _hm()
    Y, X = 90, 105
    projector.prepare(hm, 1, Y, X)
    host = np.zeros((1, Y, X), np.float32)
    hp = host.ctypes.data_as(ctypes.c_void_p)
    dev = torch.zeros((hm.numZ, 1, Y, X), device=projector.device)
    print("check host pointer", flush=True)
    assert lib.flf_check_device_buffer(0, hp, host.nbytes) == _lib.FLF_EINVAL
    print("reason:", lib.flf_last_error(None).decode(), flush=True)
    assert lib.flf_check_device_buffer(0, ctypes.c_void_p(dev.data_ptr()), dev.numel() * 4) == _lib.FLF_OK
    # a span far past the allocation holding the tensor is refused
    big = torch.empty(1 << 26, dtype=torch.uint8, device=projector.device)
    torch.cuda.synchronize()
    assert lib.flf_check_device_buffer(0, ctypes.c_void_p(big.data_ptr()), 1 << 26) == _lib.FLF_OK
    assert lib.flf_check_device_buffer(0, ctypes.c_void_p(big.data_ptr()), (1 << 26) + (1 << 34)) == _lib.FLF_EINVAL
    print("reason:", lib.flf_last_error(None).decode(), flush=True)
    arr = (ctypes.c_int * 3)(0, 1, 2)
    print("flf_forward with a host image pointer", flush=True)
    rc = lib.flf_forward(projector._ctx, ctypes.c_void_p(dev.data_ptr()), hp, arr, 3,
                         ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == _lib.FLF_EINVAL
    assert b"image" in lib.flf_last_error(projector._ctx)
    torch.cuda.synchronize()
    # the context still works afterwards
    img = projector.forward_device(hm, dev)
    torch.cuda.synchronize()
    assert float(img.abs().max()) == 0.0


def test_forward_ratio_is_ieee_division(projector):
    """flf_forward_ratio writes the same image as flf_forward and
    ratio = im / image with IEEE float32 division (numpy's), 0/0 -> NaN and
    x/0 -> inf as in lfdeconv.py:191."""
    import torch
    hm = _hm()
    Y, X = 120, 135
    g = torch.Generator(device=projector.device)
    g.manual_seed(3)
    vol = torch.rand((3, 2, Y, X), generator=g, device=projector.device)
    im = torch.rand((2, Y, X), generator=g, device=projector.device) + 0.5
    projector.prepare(hm, 2, Y, X)
    ref = projector.forward_device(hm, vol)
    img, ratio = projector.forward_ratio_device(hm, vol, im)
    _, ratio2 = projector.forward_ratio_device(hm, vol, im, store_image=False)
    torch.cuda.synchronize()
    assert torch.equal(img, ref)
    a, r = img.cpu().numpy(), ratio.cpu().numpy()
    with np.errstate(divide="ignore", invalid="ignore"):
        want = im.cpu().numpy() / a
    np.testing.assert_array_equal(r, want)
    assert torch.equal(ratio, ratio2)
    # zero image: 0/0 and x/0
    z = torch.zeros_like(vol)
    im0 = im.clone()
    im0[0, 0, 0] = 0.0
    _, rz = projector.forward_ratio_device(hm, z, im0)
    torch.cuda.synchronize()
    rz = rz.cpu().numpy()
    assert np.isnan(rz[0, 0, 0]) and np.isposinf(rz[0, 0, 1])


def test_rl_scale_shared_denominator_and_periodic_fit(projector):
    """One denominator slab per plane serves every frame (v_frames = 1) and
    equals the update with the slab repeated; the data fit of 2T frames
    against T image frames equals the fit against the image repeated."""
    import torch
    g = torch.Generator(device=projector.device)
    g.manual_seed(9)
    nz, T, Y, X = 3, 4, 50, 60
    Htf = torch.rand((nz, T, Y, X), generator=g, device=projector.device)
    v = torch.rand((nz, 1, Y, X), generator=g, device=projector.device) + 0.1
    X0 = torch.rand((nz, T, Y, X), generator=g, device=projector.device)
    Xa, Xb = X0.clone(), X0.clone()
    projector.rl_scale_device(Xa, Htf, v)
    projector.rl_scale_device(Xb, Htf, v.expand(nz, T, Y, X).contiguous())
    R1 = torch.empty_like(Htf)
    R2 = torch.empty_like(Htf)
    projector.rl_ratio_device(R1, Htf, v)
    projector.rl_ratio_device(R2, Htf, v.expand(nz, T, Y, X).contiguous())
    torch.cuda.synchronize()
    assert torch.equal(Xa, Xb) and torch.equal(R1, R2)
    hm = _hm()
    projector.prepare(hm, 2 * T, Y, X)
    img = torch.rand((2 * T, Y, X), generator=g, device=projector.device)
    im = torch.rand((T, Y, X), generator=g, device=projector.device)
    a = projector.datafit_device(img, im)
    b = projector.datafit_device(img, torch.cat([im, im]))
    np.testing.assert_array_equal(a, b)


def test_graph_recaptured_after_subset_call_and_replan(projector):
    """A cached RLGraph replayed after an eager call on a plane subset and
    after the projector was planned for another shape still equals eager RL
    (the plane selection is in the captured kernel arguments; a re-plan
    triggers a new capture)."""
    import torch
    from flfhip.lfdeconv import RLGraph, deconv_rl_device
    hm = _hm()
    vol = synthetic.blob_volume(3, 90, 75, 20, seed=6)
    projector.prepare(hm, 1, 90, 75)
    im = projector.forward_device(hm, projector.asnative(vol))
    Htf = projector.backward_device(hm, im)
    X0 = Htf.clone()
    Xe = X0.clone()
    deconv_rl_device(projector, hm, Htf, 2, Xe)
    Xg = X0.clone()
    gph = RLGraph(projector, hm, Htf, Xg)
    gph.run(2)
    torch.cuda.synchronize()
    assert torch.equal(Xg, Xe)
    # eager work on a plane subset (rewrites the selection of every call)
    sub = projector.backward_device(hm, im, planes=[1])
    sub2 = projector.forward_device(hm, Htf, planes=[0, 2])
    Xg.copy_(X0)
    gph.run(2)
    torch.cuda.synchronize()
    print("after subset call: max abs diff %.3g" % float((Xg - Xe).abs().max()), flush=True)
    assert torch.equal(Xg, Xe)
    # plan another shape (frees the tables the graph was captured with)
    projector.prepare(hm, 1, 60, 45)
    other = projector.forward_device(hm, torch.ones((3, 1, 60, 45), device=projector.device))
    gen0 = projector.plan_generation
    Xg.copy_(X0)
    gph.run(2)
    torch.cuda.synchronize()
    assert projector.plan_generation == gen0 + 1          # planned back for the graph's shape
    print("after re-plan: max abs diff %.3g" % float((Xg - Xe).abs().max()), flush=True)
    assert torch.equal(Xg, Xe)
    del sub, sub2, other


def test_ragged_cached_graph_after_subset_call(projector):
    """RaggedBatch's cached graphs (out= buffers) stay equal to eager RL
    after eager calls on a plane subset between the steps."""
    import torch
    from flfhip.ragged import RaggedBatch
    hm = _hm()
    shapes = [(60, 60), (90, 75), (75, 90), (45, 60)]
    vols = [synthetic.blob_volume(3, y, x, 10, seed=40 + i)[:, 0] for i, (y, x) in enumerate(shapes)]
    rb = RaggedBatch(hm, shapes, nbuckets=2, projectors=[projector])
    clean = rb.forward(vols)
    ims = [synthetic.poisson_image(c[None], 500.0, 7 + i)[0] for i, c in enumerate(clean)]
    staged = rb.stage(ims)
    ref = [x.clone() for x in rb.deconv_staged(staged, 2)]
    bufs = [(torch.empty_like(r), torch.empty_like(r)) for r in ref]
    for rep in range(3):
        got = rb.deconv_staged(staged, 2, out=bufs, graph=True)
        torch.cuda.synchronize()
        for k in range(len(ref)):
            assert torch.equal(got[k], ref[k]), (rep, k)
        # eager subset work on every bucket's projector
        for k in range(len(rb.buckets)):
            with rb._bucket(k) as (p, idx, yc, xc):
                p.backward_device(hm, staged[k], planes=[2])
    assert len(rb._graphs) == len(rb.buckets)
