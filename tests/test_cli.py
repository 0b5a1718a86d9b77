"""CPU tests of the time-series driver (flfhip.deconvolve_time_series) and
its TIFF I/O (flfhip.lfimage): file formats, argument checks and the
reference's refusal of mixed image shapes (deconvolve_time_series.py:20-23,
65-70)."""
import numpy as np
import pytest
from PIL import Image

from flfhip import lfimage
from flfhip import deconvolve_time_series as dts


def test_volume_tiff_roundtrip(tmp_path):
    v = np.random.default_rng(0).random((4, 30, 45)).astype(np.float32)
    p = str(tmp_path / "v.tif")
    lfimage.SaveVolumeTiff(p, v)
    assert np.array_equal(lfimage.LoadVolumeTiff(p), v)


def test_load_image_types_and_tiling(tmp_path):
    u = (np.random.default_rng(1).random((12, 9)) * 60000).astype(np.uint16)
    p = str(tmp_path / "u.tif")
    Image.fromarray(u).save(p)
    im = lfimage.LoadLightFieldTiff(p)
    assert im.dtype == np.float32 and np.array_equal(im, u.astype(np.float32))
    big = lfimage.LoadLightFieldTiff(p, bigger=(2, 3))
    assert big.shape == (24, 27) and np.array_equal(big[12:, 18:], im)
    lfimage.SaveVolumeTiff(str(tmp_path / "stack.tif"), np.zeros((2, 5, 5), np.float32))
    with pytest.raises(ValueError):
        lfimage.LoadLightFieldTiff(str(tmp_path / "stack.tif"))      # a stack is not one camera image


def test_mixed_shapes_refused_without_ragged(tmp_path, capsys):
    files = []
    for i, shape in enumerate([(30, 45), (45, 45)]):
        f = str(tmp_path / ("im%d.tif" % i))
        Image.fromarray(np.ones(shape, np.float32)).save(f)
        files.append(f)
    with pytest.raises(SystemExit) as e:
        dts.ProcessBatch(None, None, files, str(tmp_path / "out"), 2)     # fails before any projection
    assert e.value.code == 1
    assert "same dimensions" in capsys.readouterr().out


@pytest.mark.parametrize("argv", [["-b", "0"], ["-i", "0"]])
def test_invalid_arguments_exit_1(tmp_path, argv):
    with pytest.raises(SystemExit) as e:
        dts.main(["-p", str(tmp_path / "psf.mat"), "-t", "x.tif"] + argv)
    assert e.value.code == 1
