"""The reference's own benchmark scenario (benchmarking.ipynb:14-28): the
example 1273 x 1463 light-field image of Nnum = 19 lenslets (the
reference's Data/02_Rectified/exampleData/Nils_test_LFImage.tif, kept here
as tests/golden/nils_lfimage_1273x1463.tif) with the PSF of M 22.222, NA 0.5,
ML pitch 125 um, f_ML 3125 um, n 1.33, 520 nm, synthesised on the GPU.  Two
planes (z = -60 um, PSF 191 px, and z = 0, PSF 77 px): back projection of the
image, forward projection of that volume and one Prevedel RL iteration, each
against the FP64 oracle."""
import os

import numpy as np
import pytest

from oracle import projector as oproj
from oracle import rl as orl

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

HERE = os.path.dirname(os.path.abspath(__file__))
IMAGE = os.path.join(HERE, "golden", "nils_lfimage_1273x1463.tif")


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


@pytest.fixture(scope="module")
def scenario(projector):
    from flfhip import lfimage, psfgen
    from flfhip.workloads import NILS
    im = lfimage.LoadLightFieldTiff(IMAGE)
    zs = psfgen.plane_depths(*NILS["optics"][8:])
    planes = [0, int(np.argmin(np.abs(zs)))]
    hm, info = psfgen.GeneratePSF(*NILS["optics"], planes=planes)
    print("Nils PSF planes z = %s um, sides %s" % ([round(z * 1e6) for z in info["zs"]],
                                                   [hm.PSFShape(k)[0] for k in range(hm.numZ)]), flush=True)
    return im, hm


def test_nils_psf_geometry(scenario):
    im, hm = scenario
    assert im.shape == (1273, 1463)
    assert [hm.PSFShape(k)[0] for k in range(2)] == [191, 77]
    # each PSF is normalised on the full 191 px side, then cropped to its
    # plane's CAindex window (generate_psf.py:606-631): the z = -60 um plane
    # keeps the whole side and sums to 1; the z = 0 plane (77 px) loses its
    # PSFs' tails beyond the window
    for k in range(2):
        H = hm.Hcc(k, False)
        assert H.shape[:2] == (10, 10)
        sums = H.sum(axis=(2, 3), dtype=np.float64)
        if k == 0:
            np.testing.assert_allclose(sums, 1.0, rtol=1e-5)
        else:
            assert np.all(sums <= 1.0 + 1e-5) and np.all(sums > 0.99)


def test_nils_projections_match_oracle(projector, scenario):
    im, hm = scenario
    back = projector.BackwardProjectACC(hm, im[None], [0, 1])
    assert projector.fshape(0) == (1520, 1710)
    for k in range(2):
        ref = oproj.project_special(im[None].astype(np.float64), hm.Hcc(k, True).astype(np.float64))
        e = rel_l2(back[k], ref)
        print("Nils image back projection, plane %d: rel-L2 %.2e" % (k, e), flush=True)
        assert e < 1e-5
    img = projector.ForwardProjectACC(hm, back, [0, 1])
    ref = sum(oproj.project_special(back[k].astype(np.float64), hm.Hcc(k, False).astype(np.float64))
              for k in range(2))
    e = rel_l2(img, ref)
    print("forward projection of that volume: rel-L2 %.2e" % e, flush=True)
    assert e < 1e-5


def test_nils_rl_iteration_matches_oracle(projector, scenario):
    """DeconvRL with Htf=None (the reference benchmark's call), one
    iteration on the two planes, against the FP64 oracle's RL."""
    from flfhip.lfdeconv import DeconvRL
    im, hm = scenario
    got = DeconvRL(hm, None, 1, None, im=im[None], projector=projector, logPrint=False)
    Qf = [hm.Hcc(k, False) for k in range(2)]
    Qb = [hm.Hcc(k, True) for k in range(2)]
    ref = orl.deconv_rl(lambda v: oproj.forward_project(v, Qf), lambda i: oproj.backward_project(i, Qb), None, 1,
                        im=im[None].astype(np.float64))
    for k in range(2):
        e = rel_l2(got[k, 0], ref[k, 0])
        print("RL iteration 1, plane %d: rel-L2 %.2e" % (k, e), flush=True)
        assert e < 1e-4
