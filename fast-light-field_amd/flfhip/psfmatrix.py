"""PSF-matrix containers compatible with the reference's HMatrix.

``HMatrix`` reads the reference's on-disk cache format unchanged
(psfmatrix.py:14-98, 169-188): one raw float32 memmap per plane,
``H{z:02d}.array`` / ``Ht{z:02d}.array`` of shape ``HReducedShape[z]`` =
(h, h, s2, s2), plus ``HReducedShape.npy``.  ``MemoryHMatrix`` holds the same
quadrants in memory.  Both expose the attributes and methods the projector
uses: ``Nnum``, ``numZ``, ``zStart``, ``iterableBRange``, ``HPathFormat``,
``Hcc(cc, useHt)`` and ``PSFShape(cc)``.

``adjoint_quadrant`` derives the back-projection PSF set Ht from H (the role
of calcHt_new, generate_psf.py:646-719) for any odd PSF side: by default the
exact adjoint of the forward projection, or, with ``reference_window=True``,
the truncated set calcHt_new itself writes.
"""
import os

import numpy as np


class HMatrix(object):
    """Memmap-backed PSF quadrants (psfmatrix.py:14-91)."""

    def __init__(self, HPathFormat, HtPathFormat, HReducedShape, numZ=None, zStart=0, cacheMMap=True, cacheH=False):
        self.HPathFormat = HPathFormat
        self.HtPathFormat = HtPathFormat
        self.HReducedShape = [tuple(int(v) for v in s) for s in HReducedShape]
        self.numZ = numZ if numZ is not None else len(self.HReducedShape)
        self.zStart = zStart
        self.Nnum = self.HReducedShape[0][0] * 2 - 1
        self.iterableBRange = range(self.HReducedShape[0][0])
        self.cacheMMap = cacheMMap
        self.cacheH = cacheH
        self.mappedH = dict()
        self.mappedHt = dict()

    def Hcc(self, cc, useHt):
        cache = self.mappedHt if useHt else self.mappedH
        if self.cacheMMap and cc in cache:
            return cache[cc]
        fmt = self.HtPathFormat if useHt else self.HPathFormat
        result = np.memmap(fmt.format(z=cc + self.zStart), dtype="float32", mode="r",
                           shape=self.HReducedShape[cc + self.zStart])
        if self.cacheMMap:
            cache[cc] = result
        return result

    def PSFShape(self, cc):
        s = self.HReducedShape[cc + self.zStart]
        return (s[2], s[3])

    def ClearCache(self):
        self.mappedH.clear()
        self.mappedHt.clear()


class MemoryHMatrix(object):
    """In-memory PSF quadrants with the HMatrix interface."""

    def __init__(self, H, Ht, identifier=None):
        assert len(H) == len(Ht) and len(H) > 0
        self._H = [np.ascontiguousarray(q, dtype=np.float32) for q in H]
        self._Ht = [np.ascontiguousarray(q, dtype=np.float32) for q in Ht]
        for q, qt in zip(self._H, self._Ht):
            if q.ndim != 4 or q.shape != qt.shape or q.shape[0] != q.shape[1] or q.shape[2] != q.shape[3]:
                raise ValueError("each plane must be an (h, h, s2, s2) quadrant, same for H and Ht")
        self.HReducedShape = [q.shape for q in self._H]
        self.numZ = len(self._H)
        self.zStart = 0
        self.Nnum = self._H[0].shape[0] * 2 - 1
        self.iterableBRange = range(self._H[0].shape[0])
        self.HPathFormat = identifier if identifier is not None else "memory:%x" % id(self)
        self.HtPathFormat = self.HPathFormat + ":t"

    def Hcc(self, cc, useHt):
        return (self._Ht if useHt else self._H)[cc + self.zStart]

    def PSFShape(self, cc):
        s = self.HReducedShape[cc + self.zStart]
        return (s[2], s[3])


def GetPathFormats(matPath):
    """psfmatrix.py:93-98."""
    mmapPath = os.path.splitext(matPath)[0]
    return mmapPath, mmapPath + "/H{z:02d}.array", mmapPath + "/Ht{z:02d}.array"


def LoadMatrix(matPath, numZ=None, zStart=0, **_ignored):
    """psfmatrix.py:169-188 for an existing memmap cache.  Converting a
    Matlab .mat file needs h5py, which this package does not use."""
    mmapPath, hfmt, htfmt = GetPathFormats(matPath)
    shape_file = mmapPath + "/HReducedShape.npy"
    if not os.path.exists(shape_file):
        raise FileNotFoundError("%s not found (convert the .mat with the reference tools first)" % shape_file)
    return HMatrix(hfmt, htfmt, np.load(shape_file), numZ=numZ, zStart=zStart)


def SaveMatrix(dirPath, H, Ht):
    """Write quadrants in the reference's memmap cache format
    (psfmatrix.py:145-165) and return the HMatrix reading them."""
    os.makedirs(dirPath, exist_ok=True)
    shapes = []
    for z, (q, qt) in enumerate(zip(H, Ht)):
        for fmt, arr in (("/H{z:02d}.array", q), ("/Ht{z:02d}.array", qt)):
            m = np.memmap(dirPath + fmt.format(z=z), dtype="float32", mode="w+", shape=arr.shape)
            m[...] = arr
            m.flush()
            del m
        shapes.append(q.shape)
    np.save(dirPath + "/HReducedShape.npy", np.array(shapes))
    np.save(dirPath + "/HtReducedShape.npy", np.array(shapes))
    return HMatrix(dirPath + "/H{z:02d}.array", dirPath + "/Ht{z:02d}.array", shapes)


def _fold(i, nnum):
    c = nnum // 2
    return (i, False) if i <= c else (nnum - 1 - i, True)


def full_psf(Q):
    """All N*N PSFs from a quadrant by the symmetry the reference assumes
    (py_light_field.cpp:1514-1562): transpose for (a, b), row flip for
    N-1-b, column flip for N-1-a."""
    h = Q.shape[0]
    nnum = 2 * h - 1
    out = np.empty((nnum, nnum) + Q.shape[2:], dtype=Q.dtype)
    for b in range(nnum):
        bq, fb = _fold(b, nnum)
        for a in range(nnum):
            aq, fa = _fold(a, nnum)
            M = Q[bq][aq] if bq <= aq else Q[aq][bq].T
            if fb:
                M = M[::-1, :]
            if fa:
                M = M[:, ::-1]
            out[b, a] = M
    return out


def adjoint_quadrant(Q, reference_window=False):
    """Quadrant of the back-projection PSF set Ht for the forward set Q:
        Ht_a[e] = H_{(a + e - c2) mod N}[2 c2 - e]        (2-D indices)
    which makes back-projection the exact adjoint of forward projection.

    calcHt_new (generate_psf.py:646-719) stamps the same values, but its
    s2 x s2 cut around the image centre and the following im_shift2 by
    (h-1-a, h-1-b) pixels (generate_psf.py:711-716) zero the first h-1-b rows
    and h-1-a columns of Ht[b][a].  reference_window=True reproduces that
    truncation, so that the result equals the Ht the reference saves (for the
    PSF sides s2 = 2kN+1 it accepts, generate_psf.py:680); the default keeps
    those border values, which are zero for a PSF whose outermost h-1 rows
    and columns are zero."""
    Q = np.asarray(Q)
    h = Q.shape[0]
    nnum = 2 * h - 1
    s2 = Q.shape[2]
    if s2 % 2 != 1:
        raise ValueError("PSF side must be odd")
    c2 = (s2 - 1) // 2
    H = full_psf(Q)
    e = np.arange(s2)
    flip = (2 * c2 - e)
    out = np.empty_like(Q)
    for b in range(h):
        sb = (b + e - c2) % nnum
        for a in range(h):
            sa = (a + e - c2) % nnum
            out[b, a] = H[sb[:, None], sa[None, :], flip[:, None], flip[None, :]]
            if reference_window:
                out[b, a, :h - 1 - b, :] = 0
                out[b, a, :, :h - 1 - a] = 0
    return out
