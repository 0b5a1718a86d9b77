"""flf_datafit_noalloc (the data fit with its partial sums in a fixed device
array, for callers that cannot allocate) against flf_datafit on the same
inputs, through the C ABI: bit-identical on two calls in a row, the second
with a smaller input, plus the refusals that return before any launch."""
import ctypes

import numpy as np
import pytest

from flfhip import _lib, synthetic

pytestmark = pytest.mark.gpu


def _noalloc(lib, img, im, out):
    import torch
    T, Y, X = img.shape
    return lib.flf_datafit_noalloc(ctypes.c_void_p(img.data_ptr()), ctypes.c_void_p(im.data_ptr()), T, im.shape[0],
                                   Y * X, ctypes.c_void_p(out.data_ptr()),
                                   ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))


def test_noalloc_datafit_matches_datafit(projector):
    import torch
    lib = projector.lib
    hm = synthetic.make_psf_set(15, [-1, 0, 1], synthetic.psf_sides([-1, 0, 1], 15, 31), seed=2)
    g = torch.Generator(device=projector.device)
    g.manual_seed(17)
    out = torch.empty(2 * 8, dtype=torch.float64, device=projector.device)
    for step, (T, Ti, Y, X) in enumerate([(8, 4, 150, 165), (3, 3, 60, 75)]):
        print("call %d: T=%d im_frames=%d %dx%d" % (step, T, Ti, Y, X), flush=True)
        img = torch.rand((T, Y, X), generator=g, device=projector.device)
        im = torch.rand((Ti, Y, X), generator=g, device=projector.device)
        projector.prepare(hm, T, Y, X)
        ref = projector.datafit_device(img, im)                  # flf_datafit (context buffers, host output)
        out.fill_(-1.0)
        assert _noalloc(lib, img, im, out) == _lib.FLF_OK, lib.flf_last_error(None)
        torch.cuda.synchronize()
        got = out[:2 * T].cpu().numpy().reshape(T, 2)
        np.testing.assert_array_equal(got, ref)
        if T < 8:
            assert float(out[2 * T:].min()) == -1.0                   # nothing written past 2T doubles
        # FP64 check of the sums themselves
        d = (img.double() - im.double().repeat(T // Ti, 1, 1)).reshape(T, -1)
        np.testing.assert_allclose(got[:, 0], (d * d).sum(1).cpu().numpy(), rtol=1e-12)


def test_noalloc_datafit_no_growth_and_refusals(projector):
    import torch
    lib = projector.lib
    img = torch.rand((2, 64, 64), device=projector.device)
    out = torch.empty(4, dtype=torch.float64, device=projector.device)
    assert _noalloc(lib, img, img, out) == _lib.FLF_OK
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for _ in range(5):
        assert _noalloc(lib, img, img, out) == _lib.FLF_OK
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] == free0
    np.testing.assert_array_equal(out.cpu().numpy()[0::2], [0.0, 0.0])
    # refusals (host-side, nothing queued): too many frames, host output
    big = torch.empty((513, 1, 4), device=projector.device)
    bo = torch.empty(2 * 513, dtype=torch.float64, device=projector.device)
    assert _noalloc(lib, big, big, bo) == _lib.FLF_EINVAL
    assert b"T must be" in lib.flf_last_error(None)
    host = np.zeros(4, np.float64)
    rc = lib.flf_datafit_noalloc(ctypes.c_void_p(img.data_ptr()), ctypes.c_void_p(img.data_ptr()), 2, 2, 64 * 64,
                                 host.ctypes.data_as(ctypes.c_void_p), None)
    assert rc == _lib.FLF_EINVAL
    assert b"out" in lib.flf_last_error(None)
