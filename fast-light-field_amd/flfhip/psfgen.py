"""Light-field PSF synthesis on the GPU (the role of generate_psf.GeneratePSF,
generate_psf.py:63-798).

``GeneratePSF(M, NA, MLPitch, Nnum, OSR, n, fml, lam, zmin, zmax, zspacing)``
takes the reference's arguments (SI units) and returns the PSF set as a
``MemoryHMatrix``: per plane the quadrant H[b][a][q][p] the reference stores
(cropped to the plane's CAindex window, generate_psf.py:624-631 with
psfmatrix.py:141-147) and its Ht, built as calcHt_new builds it
(generate_psf.py:646-719; psfmatrix.adjoint_quadrant with the reference's
window).  ``SaveMatrix`` writes the reference's memmap cache format.

Every numerical stage runs in libflf_hip.so (FP64 kernels, flf_psf_*): the
Debye integrals of the wave PSF, the Fresnel propagation of every lenslet
offset through the microlens array, the intensity binning, clipping and
per-PSF normalisation.  This module holds the generator's bookkeeping (grid
sizes, crop windows, the PSF-size estimate from the computed intensity
line) and moves results to host memory.

Not supported: normalisePSF=False with reproduceMaxBug=True (the
reference's reproduction of a Matlab bug, generate_psf.py:328-333).
"""
import ctypes
import math

import numpy as np
import torch

from . import _lib
from .psfmatrix import MemoryHMatrix, SaveMatrix, adjoint_quadrant


class Optics(ctypes.Structure):
    _fields_ = [("M", ctypes.c_double), ("NA", ctypes.c_double), ("MLPitch", ctypes.c_double),
                ("n", ctypes.c_double), ("fml", ctypes.c_double), ("lam", ctypes.c_double),
                ("Nnum", ctypes.c_int), ("OSR", ctypes.c_int)]


# Named parameter sets of the reference's command line (generate_psf.py:848-851).
PRESETS = {
    "prevedelM40Partial": (40, 0.95, 150e-6, 15, 3, 1.0, 3000e-6, 520e-9, -26e-6, 0, 2e-6),
    "prevedelM22": (22.2, 0.5, 125e-6, 19, 3, 1.33, 3125e-6, 520e-9, -156e-6, 156e-6, 4e-6),
}


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _check(lib, rc, what):
    if rc != 0:
        msg = lib.flf_last_error(None)
        text = "%s failed (%d): %s" % (what, rc, msg.decode() if msg else "")
        raise ValueError(text) if rc == _lib.FLF_EINVAL else _lib.FLFError(text)


def plane_depths(zmin, zmax, zspacing):
    """x3objspace (generate_psf.py:150)."""
    return np.arange(zmin, zmax + 0.1 * zspacing, zspacing)


class Geometry(object):
    """Grid and window sizes of the generator for a PSF size estimate
    imgsize_ref (lenslets): generate_psf.py:207-221, 352-367, 517-519,
    624-631."""

    def __init__(self, Nnum, OSR, imgsize_ref, p3max):
        self.Nnum, self.OSR, self.imgsize_ref, self.p3max = Nnum, OSR, imgsize_ref, p3max
        self.half = max(Nnum * (imgsize_ref + 1), 2 * Nnum)        # IMG_HALFWIDTH
        self.L = 2 * self.half * OSR + 1
        self.centre = (self.L + 1) // 2                              # 1-based
        self.s2 = 2 * Nnum * imgsize_ref + 1                         # len(CP_p)

    def lo(self, p3):
        """First wedge row of the plane at depth p3 (centerArea_IL_p[0])."""
        ref_il = math.ceil(self.imgsize_ref * (abs(p3) / self.p3max)) if self.p3max > 0 else 0
        hw = max(self.Nnum * ref_il * self.OSR, 2 * self.Nnum * self.OSR)
        return int(max(self.centre - hw, 1) - 1)

    def ca_window(self, p3):
        """CAindex of the plane, as a 0-based half-open slice of the s2 side."""
        c = int(math.ceil(self.s2 / 2))
        ref_il = math.ceil(self.imgsize_ref * (abs(p3) / self.p3max)) if self.p3max > 0 else 0
        hw = max(self.Nnum * ref_il, 2 * self.Nnum)
        return int(max(c - hw, 1)) - 1, int(min(c + hw, self.s2))


def psf_line(lib, optics, p3max, device):
    """The normalised intensity line at the deepest plane and the PSF size in
    lenslets it implies (generate_psf.py:158-167)."""
    npts = optics.Nnum * optics.OSR * 20 + 1
    out = torch.empty(npts, dtype=torch.float64, device=device)
    _check(lib, lib.flf_psf_line(ctypes.byref(optics), ctypes.c_double(p3max), npts, _ptr(out), _stream()),
           "flf_psf_line")
    line = out.cpu().numpy()
    line = line / line.max()
    below = np.nonzero(line < 0.04)[0]
    if len(below) == 0:
        raise ValueError("Estimated PSF size exceeds the limit")
    return line, int(math.ceil(below[0] / (optics.OSR * optics.Nnum)))


def plane_psfs(lib, optics, geo, p3, device, normalise=True, clip=0.005):
    """(h, h, s2, s2) float32 light-field PSFs of one plane (uncropped),
    clipped and (normalise) divided by their sums."""
    h = (optics.Nnum + 1) // 2
    L = geo.L
    lo = geo.lo(p3)
    pat = torch.empty((L, L, 2), dtype=torch.float64, device=device)
    _check(lib, lib.flf_psf_wave(ctypes.byref(optics), ctypes.c_double(p3), geo.half, lo, _ptr(pat), _stream()),
           "flf_psf_wave")
    Hd = torch.empty((h, h, geo.s2, geo.s2), dtype=torch.float64, device=device)
    _check(lib, lib.flf_psf_lightfield(ctypes.byref(optics), _ptr(pat), geo.half, geo.imgsize_ref, lo, _ptr(Hd),
                                       _stream()), "flf_psf_lightfield")
    del pat
    out = torch.empty((h, h, geo.s2, geo.s2), dtype=torch.float32, device=device)
    _check(lib, lib.flf_psf_normalise(_ptr(Hd), h, geo.s2, ctypes.c_double(clip), ctypes.c_int(1 if normalise else 0),
                                      _ptr(out), _stream()), "flf_psf_normalise")
    return out.cpu().numpy()


def GeneratePSF(M, NA, MLPitch, Nnum, OSR, n, fml, lam, zmin, zmax, zspacing, normalisePSF=True,
                reproduceMaxBug=False, matrixSaveDir=None, verbose=False, device=0, planes=None):
    """generate_psf.GeneratePSF (generate_psf.py:63-798) on the GPU.  Returns
    (MemoryHMatrix, info); with matrixSaveDir the quadrants are also written
    in the reference's memmap cache format there.  planes: indices into
    x3objspace (default all)."""
    if Nnum % 2 != 1:
        raise AssertionError("Nnum must be an odd number")
    if OSR % 2 != 1:
        raise AssertionError("OSR must be an odd number")
    if reproduceMaxBug:
        raise NotImplementedError("reproduceMaxBug is not supported")
    if not torch.cuda.is_available():
        raise RuntimeError("GeneratePSF needs a ROCm GPU")
    lib = _lib.load()
    dev = torch.device("cuda", device)
    optics = Optics(float(M), float(NA), float(MLPitch), float(n), float(fml), float(lam), int(Nnum), int(OSR))
    zs = plane_depths(zmin, zmax, zspacing)
    if len(zs) < 2:
        raise AssertionError("at least two planes are needed (generate_psf.py:155)")
    p3max = float(np.max(np.abs(zs)))
    with torch.cuda.device(dev):
        line, ref = psf_line(lib, optics, p3max, dev)
        geo = Geometry(Nnum, OSR, ref, p3max)
        if verbose:
            print("PSF size ~ %d lenslets, grid %d x %d, PSF side %d" % (ref, geo.L, geo.L, geo.s2))
        planes = range(len(zs)) if planes is None else list(planes)
        full = [plane_psfs(lib, optics, geo, float(zs[c]), dev, normalise=normalisePSF) for c in planes]
    if not normalisePSF:
        # the reference's alternative: one division by the maximum over all planes
        hmax = max(float(q.max()) for q in full)
        full = [q / np.float32(hmax) for q in full]
    H, Ht = [], []
    for q, c in zip(full, planes):
        lo, hi = geo.ca_window(float(zs[c]))
        qt = adjoint_quadrant(q, reference_window=True)           # calcHt_new on the uncropped set
        H.append(np.ascontiguousarray(q[:, :, lo:hi, lo:hi]))
        Ht.append(np.ascontiguousarray(qt[:, :, lo:hi, lo:hi]))
    hm = MemoryHMatrix(H, Ht, identifier="generated:M%gNA%gMLPitch%gfml%gNnum%dlambda%gn%g:%s"
                       % (M, NA, MLPitch * 1e6, fml * 1e6, Nnum, lam * 1e9, n, ",".join("%g" % (zs[c] * 1e6)
                                                                                       for c in planes)))
    info = {"imgsize_ref": ref, "grid": geo.L, "psf_side": geo.s2, "zs": [float(zs[c]) for c in planes],
            "line": line}
    if matrixSaveDir is not None:
        SaveMatrix(matrixSaveDir, H, Ht)
    return hm, info


def main(argv=None):
    """python -m flfhip.psfgen NAME [--out DIR]: the reference's named
    parameter sets (generate_psf.py:842-856)."""
    import argparse
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("name", choices=sorted(PRESETS))
    ap.add_argument("--out", default=None, help="directory for the memmap cache (H{z:02d}.array ...)")
    args = ap.parse_args(argv)
    hm, info = GeneratePSF(*PRESETS[args.name], matrixSaveDir=args.out, verbose=True)
    print("generated %d planes, PSF size %d lenslets, sides %s" % (hm.numZ, info["imgsize_ref"],
                                                                   [s[2] for s in hm.HReducedShape]))


if __name__ == "__main__":
    main()
