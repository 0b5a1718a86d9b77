"""CPU tests of the Python boundary's argument checks: every mismatch in
shape, dtype, layout or frame count is refused with the reference's
exception types (ValueError / TypeError, py_light_field.cpp:1425-1493,
lfdeconv.py broadcasting errors) before the library or the device is
touched.  The projector here is a Projector_hip whose library is a recorder:
a test fails if any library entry point is reached."""
import numpy as np
import pytest
import torch

from flfhip import lfdeconv, synthetic
from flfhip.projector import Projector_hip


class RecordingLib(object):
    def __init__(self):
        self.calls = []

    def __getattr__(self, name):
        if not name.startswith("flf_"):
            raise AttributeError(name)

        def fn(*a):
            self.calls.append(name)
            return 0
        return fn


def cpu_projector():
    p = Projector_hip.__new__(Projector_hip)
    p.lib = RecordingLib()
    p.device = torch.device("cpu")
    p.name = "recorder"
    p.storeVolumesOnGPU = False
    p.cacheFH = True
    p._ctx = None
    p._key = None
    p._tcap = p._frames = 0
    p._loaded = set()
    p.plan_generation = 0
    p._dims = None
    return p


@pytest.fixture
def setup():
    hm = synthetic.make_psf_set(5, [-1, 0, 1], [7, 5, 7], seed=1)
    return cpu_projector(), hm


def t(*shape, dtype=torch.float32):
    return torch.ones(shape, dtype=dtype)


def test_rl_update_refuses_mismatched_volumes(setup):
    p, hm = setup
    img = t(2, 20, 20)
    with pytest.raises(ValueError, match="X has shape"):
        p.rl_update_device(hm, img, t(3, 1, 20, 20), Htf=t(3, 2, 20, 20))    # X with 1 frame for 2
    with pytest.raises(ValueError, match="Htf has shape"):
        p.rl_update_device(hm, img, t(3, 2, 20, 20), Htf=t(3, 1, 20, 20))
    with pytest.raises(ValueError, match="trust has shape"):
        p.rl_update_device(hm, img, t(3, 2, 20, 20), trust=t(3, 2, 20, 20), mode=2)
    with pytest.raises(ValueError, match="needs Htf"):
        p.rl_update_device(hm, img, t(3, 2, 20, 20), mode=1)
    with pytest.raises(TypeError, match="float32"):
        p.rl_update_device(hm, img, t(3, 2, 20, 20, dtype=torch.float64), Htf=t(3, 2, 20, 20))
    with pytest.raises(ValueError, match="contiguous"):
        p.rl_update_device(hm, img, t(3, 2, 20, 40)[..., ::2], Htf=t(3, 2, 20, 20))
    with pytest.raises(ValueError, match="mode"):
        p.rl_update_device(hm, img, t(3, 2, 20, 20), mode=3)
    assert p.lib.calls == []


def test_backward_out_and_datafit_refused(setup):
    p, hm = setup
    with pytest.raises(ValueError, match="out has shape"):
        p.backward_device(hm, t(2, 20, 20), out=t(3, 1, 20, 20))
    with pytest.raises(ValueError, match="does not match"):
        p.datafit_device(t(2, 20, 20), t(2, 20, 21))
    with pytest.raises(ValueError, match="does not match"):
        p.datafit_device(t(3, 20, 20), t(2, 20, 20))        # 3 frames against 2: not a multiple
    with pytest.raises(ValueError, match="volume has 2 planes"):
        p.forward_device(hm, t(2, 1, 20, 20))
    with pytest.raises(ValueError, match="volB has shape"):
        p.forward_pair_device(hm, t(3, 1, 20, 20), t(3, 2, 20, 20))
    with pytest.raises(ValueError, match="im has shape"):
        p.forward_ratio_device(hm, t(3, 2, 20, 20), t(1, 20, 20))
    with pytest.raises(ValueError, match="v must be"):
        p.rl_scale_device(t(3, 2, 20, 20), t(3, 2, 20, 20), t(3, 3, 20, 20))
    assert p.lib.calls == []


def test_deconvrl_refuses_before_device_use(setup):
    p, hm = setup
    im = np.ones((2, 20, 20), np.float32)

    def no_device(*a, **k):
        raise AssertionError("reached the device")
    p.asnative = no_device
    with pytest.raises(ValueError, match="Xguess has shape"):
        lfdeconv.DeconvRL(hm, None, 1, np.ones((3, 1, 20, 20), np.float32), logPrint=False, projector=p, im=im)
    with pytest.raises(ValueError, match="Htf has shape"):
        lfdeconv.DeconvRL(hm, np.ones((3, 2, 20, 19), np.float32), 1, None, logPrint=False, projector=p, im=im)
    with pytest.raises(ValueError, match="trustMask has shape"):
        lfdeconv.DeconvRL(hm, None, 1, None, logPrint=False, projector=p, im=im, prevedel=False,
                          trustMask=np.ones((19, 20)))
    with pytest.raises(ValueError, match="im must be"):
        lfdeconv.DeconvRL(hm, None, 1, None, logPrint=False, projector=p, im=np.ones((20, 20), np.float32))
    assert p.lib.calls == []


def test_state_key_tracks_plan_and_frame_sizes(setup, monkeypatch):
    # A captured graph is replayed only while the plan's tables and the
    # per-frame sizes it reads are those of its capture.
    from flfhip import projector as pmod
    monkeypatch.setattr(pmod, "_stream_handle", lambda: None)
    p, hm = setup
    k0 = p.state_key()
    p._key = (id(hm), None, 3, 0, 5, 20, 20)
    p._frames = p._tcap = 2
    p.set_frame_dims([(10, 20), (20, 20)])
    k1 = p.state_key()
    assert k1 != k0
    p.set_frame_dims(None)
    p.set_frame_dims([(10, 20), (20, 20)])
    assert p.state_key() == k1                   # same sizes again: the same graph stays valid
    p.set_frame_dims([(20, 20), (20, 20)])       # full size is the same as no per-frame sizes
    assert p.state_key()[2] is None
    p.plan_generation += 1
    assert p.state_key() != k1
