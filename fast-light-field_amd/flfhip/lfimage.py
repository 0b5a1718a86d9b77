"""Light-field image and volume files (the role of lfimage.py:4-21 and of
the tifffile calls in deconvolve_time_series.py:19, 33).

TIFF is read and written through Pillow: a camera image is one page
(any integer or float sample type, returned as float32); a deconvolved
volume is written as one float32 page per plane.
"""
import numpy as np


def _pil():
    try:
        from PIL import Image, ImageSequence
    except ImportError as e:                       # pragma: no cover - Pillow ships with this image
        raise ImportError("TIFF I/O needs Pillow (PIL)") from e
    return Image, ImageSequence


def LoadLightFieldTiff(imagePath, bigger=(1, 1)):
    """lfimage.py:4-21: a 2-D image as float32, optionally tiled `bigger`
    times (the reference's stress-test option).  No transpose, as in the
    reference."""
    Image, ImageSequence = _pil()
    with Image.open(imagePath) as im:
        pages = [np.array(p) for p in ImageSequence.Iterator(im)]
    if len(pages) != 1 or pages[0].ndim != 2:
        raise ValueError("%s: expected one 2-D image, found %d page(s) of shape %s"
                         % (imagePath, len(pages), pages[0].shape if pages else None))
    return np.tile(pages[0].astype(np.float32), bigger)


def SaveVolumeTiff(path, volume):
    """Write a (numZ, Y, X) volume as numZ float32 pages (what
    tifffile.imsave does with deconvolvedResult[:, i] in
    deconvolve_time_series.py:33)."""
    Image, _ = _pil()
    vol = np.asarray(volume, dtype=np.float32)
    if vol.ndim == 2:
        vol = vol[np.newaxis]
    if vol.ndim != 3:
        raise ValueError("volume must be (numZ, Y, X)")
    pages = [Image.fromarray(np.ascontiguousarray(p)) for p in vol]
    pages[0].save(path, format="TIFF", save_all=True, append_images=pages[1:])


def LoadVolumeTiff(path):
    """Read a multi-page float TIFF back as (numZ, Y, X) float32."""
    Image, ImageSequence = _pil()
    with Image.open(path) as im:
        return np.stack([np.array(p, dtype=np.float32) for p in ImageSequence.Iterator(im)])
