"""The time-series driver end to end on the GPU: PSF cache files and TIFF
images on disk in, one deconvolved TIFF volume per timepoint out, each
equal to the FP64 oracle's RL (Xguess = Htf, deconvolve_time_series.py:
27-28) within 1e-4 relative L2; batches split by --batch-size; frames of
mixed shape with --ragged."""
import numpy as np
import pytest
from PIL import Image

from flfhip import lfimage, psfmatrix, synthetic
from flfhip import deconvolve_time_series as dts
from oracle import projector as oproj
from oracle import rl as orl

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def psf(tmp_path_factory):
    d = tmp_path_factory.mktemp("psf")
    N, zs = 15, [-2, 0, 3]
    hm = synthetic.make_psf_set(N, zs, synthetic.psf_sides(zs, 15, 45), seed=21)
    H = [hm.Hcc(z, False) for z in range(3)]
    Ht = [hm.Hcc(z, True) for z in range(3)]
    psfmatrix.SaveMatrix(str(d / "psf"), H, Ht)
    fwd = lambda v: oproj.forward_project(v, H)
    bwd = lambda i: oproj.backward_project(i, Ht)
    return str(d / "psf.mat"), fwd, bwd


def _images(tmp_path, fwd, shapes):
    files = []
    for t, (y, x) in enumerate(shapes):
        v = synthetic.blob_volume(3, y, x, 12, seed=60 + t)
        im = synthetic.poisson_image(fwd(v.astype(np.float64)), 600.0, seed=t)[0]
        f = str(tmp_path / ("frame%02d.tif" % t))
        Image.fromarray(im.astype(np.float32)).save(f)
        files.append(f)
    return files


@pytest.mark.parametrize("ragged", [False, True])
def test_driver_matches_oracle(projector, psf, tmp_path, ragged):
    psfmat, fwd, bwd = psf
    shapes = [(90, 105), (60, 75), (90, 105)] if ragged else [(90, 105)] * 3
    files = _images(tmp_path, fwd, shapes)
    out = tmp_path / "out"
    argv = ["-p", psfmat, "-d", str(out), "-i", "2", "-b", "2", "-t"] + files + (["--ragged"] if ragged else [])
    print("driver", argv[-5:], flush=True)
    written = dts.main(argv)
    assert len(written) == 3
    for t, f in enumerate(files):
        vol = lfimage.LoadVolumeTiff(written[t])
        im = lfimage.LoadLightFieldTiff(f)
        Xr = orl.deconv_rl(fwd, bwd, None, 2, Xguess=None, im=im[None].astype(np.float64))
        err = rel_l2(vol, Xr[:, 0])
        print("timepoint %d %s rel-L2 vs oracle %.2e" % (t, shapes[t], err), flush=True)
        assert vol.shape == (3,) + shapes[t]
        assert written[t].endswith("frame%02d_deconvolved.tif" % t)
        assert err < 1e-4


def test_driver_learned_start_guard(projector, psf, tmp_path):
    """--learned with an untrained network: the guard rejects its guess
    (worse data fit than one plain iteration), so every frame continues
    from the first plain iterate X1 = Htf / Ht(H 1); the written volume is
    then RL from a uniform start with one more iteration."""
    from flfhip.lensletnet import LensletNet
    psfmat, fwd, bwd = psf
    files = _images(tmp_path, fwd, [(90, 105)] * 2)
    w = str(tmp_path / "untrained.npz")
    LensletNet(15, 3, width=16, nblocks=1, seed=5).save(w)
    written = dts.main(["-p", psfmat, "-d", str(tmp_path / "out"), "-i", "2", "--learned", w, "-t"] + files)
    for t, f in enumerate(files):
        vol = lfimage.LoadVolumeTiff(written[t])
        im = lfimage.LoadLightFieldTiff(f)[None].astype(np.float64)
        Xr = orl.deconv_rl(fwd, bwd, None, 3, Xguess=np.ones((3, 1) + im.shape[1:]), im=im)
        err = rel_l2(vol, Xr[:, 0])
        print("learned-start timepoint %d rel-L2 vs oracle (uniform start, 3 iterations) %.2e" % (t, err), flush=True)
        assert err < 1e-4
