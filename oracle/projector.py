"""FP64 projectors (oracle; test use only).

Two independent algorithms for the same per-plane operator

    out_t[p] = sum_{q} src_t[q] * K_{q mod N}[p - q + c2]        ('same' crop)

where K is H (forward) or Ht (backward):

* ``project_direct``  - restates the reference's definitive "old" code,
  ForwardProjectForZ_old / BackwardProjectForZ_old (projector.py:1084-1126):
  for every lenslet offset (b, a), zero everything except src[b::N, a::N]
  and fftconvolve(..., 'same') with the full PSF K[b][a].
* ``project_special`` - restates the fast path the reference actually runs,
  ProjectForZList + InverseRFFTList + CropAfterFFT
  (py_light_field.cpp:1417-1740, 711-1102, 284-341; projector.py:103-151,
  792-896): sub-sampled small FFTs tiled up to the full grid, the quadrant
  spectra F(Q), F(Q^T), the mirrorY copy and the conj-mirrored row access,
  accumulated per offset and inverse-transformed once.
"""
import numpy as np
from scipy.signal import fftconvolve

from . import shapes
from .psf import expand_quadrant


def project_direct(src, Kfull):
    """projector.py:1084-1126 in FP64.  src (T, Y, X); Kfull (N, N, s2, s2)."""
    src = np.asarray(src, dtype=np.float64)
    single = src.ndim == 2
    if single:
        src = src[None]
    nnum = Kfull.shape[0]
    out = np.zeros(src.shape, dtype=np.float64)
    for t in range(src.shape[0]):
        for b in range(nnum):
            for a in range(nnum):
                tmp = np.zeros(src.shape[1:], dtype=np.float64)
                tmp[b::nnum, a::nnum] = src[t, b::nnum, a::nnum]
                out[t] += fftconvolve(tmp, np.asarray(Kfull[b, a], np.float64), "same")
    return out[0] if single else out


def project_bruteforce(src, Kfull):
    """Explicit shift-and-add of the operator definition (tiny inputs only)."""
    src = np.asarray(src, dtype=np.float64)
    T, Y, X = src.shape
    nnum = Kfull.shape[0]
    s2 = Kfull.shape[2]
    c2 = (s2 - 1) // 2
    out = np.zeros((T, Y, X))
    for t in range(T):
        for q0 in range(Y):
            for q1 in range(X):
                v = src[t, q0, q1]
                if v == 0:
                    continue
                K = Kfull[q0 % nnum, q1 % nnum]
                for e0 in range(s2):
                    p0 = q0 + e0 - c2
                    if p0 < 0 or p0 >= Y:
                        continue
                    for e1 in range(s2):
                        p1 = q1 + e1 - c2
                        if 0 <= p1 < X:
                            out[t, p0, p1] += v * K[e0, e1]
    return out


def _multipliers(fshape, s2, nnum):
    """projector.py:137-151, evaluated exactly in complex128.
    Returns (mirrorX[fY] rows, mirrorY[fX] cols, expandX[N, fX], expandY[N, fY])."""
    fY, fX = fshape
    kY = np.arange(fY)
    kX = np.arange(fX)
    mirrorX = np.exp(1j * (1 + fY - s2) * 2 * np.pi / fY * kY)
    mirrorY = np.exp(1j * (1 + fX - s2) * 2 * np.pi / fX * kX)
    expandX = np.exp(-2j * np.pi / fX * np.outer(np.arange(nnum), kX))
    expandY = np.exp(-2j * np.pi / fY * np.outer(np.arange(nnum), kY))
    return mirrorX, mirrorY, expandX, expandY


def project_special_fourier(src, Q, fshape):
    """ProjectForZList for one plane (py_light_field.cpp:1417-1635):
    returns the accumulated half-plane spectrum (T, fY, fX//2+1)."""
    src = np.asarray(src, dtype=np.float64)
    T, Y, X = src.shape
    h = Q.shape[0]
    nnum = 2 * h - 1
    c = nnum // 2
    s2 = Q.shape[2]
    fY, fX = fshape
    assert fY % nnum == 0 and fX % nnum == 0
    rY, rX = fY // nnum, fX // nnum
    rfX = fX // 2 + 1
    mirrorX, mirrorY, expandX, expandY = _multipliers(fshape, s2, nnum)
    yMult = (expandY, expandY * mirrorX[None, :])        # projector.py:151
    acc = np.zeros((T, fY, rfX), dtype=np.complex128)
    rows = np.arange(fY) % rY
    mirror_cols = (fX - np.arange(rfX)) % fX              # fbu[trueLength - x], x=0 -> 0

    def part1(bb, aa):
        # special_fftconvolve_part1 (py_light_field.cpp:1001-1027)
        sub = np.zeros((T, rY, rX))
        s = src[:, bb::nnum, aa::nnum]
        sub[:, :s.shape[1], :s.shape[2]] = s
        F = np.fft.fft2(sub)
        reps = -(-fX // rX)
        return np.tile(F, (1, 1, reps))[:, :, :fX] * expandX[aa][None, None, :]

    def convolve4(bb, aa, fh, mirror_x):
        # ConvolvePart4Nomirror / ConvolvePart4Mirror (py_light_field.cpp:1029-1087)
        nonlocal acc
        P = part1(bb, aa)[:, rows, :rfX]
        acc += P * fh[None, :, :rfX] * yMult[0][bb][None, :, None]          # CalculateRow
        if mirror_x:
            bb2 = nnum - 1 - bb
            P2 = part1(bb2, aa)[:, rows, :rfX]
            K2 = np.conj(fh[:, mirror_cols])                               # CalculateRowMirror
            acc += P2 * K2[None] * yMult[1][bb2][None, :, None]

    def mirror_y_array(fh):
        # MirrorYArray (py_light_field.cpp:284-301)
        return np.conj(fh[(-np.arange(fY)) % fY, :]) * mirrorY[None, :]

    def part2(bb, aa, fh, mirror_y, mirror_x):
        # ConvolvePart2 (py_light_field.cpp:1294-1317)
        convolve4(bb, aa, fh, mirror_x)
        if mirror_y:
            convolve4(bb, nnum - 1 - aa, mirror_y_array(fh), mirror_x)

    for bb in range(h):
        for aa in range(bb, h):
            mX = bb != c
            mY = aa != c
            tr = (aa != bb) and (aa != nnum - bb - 1)
            Qs = np.asarray(Q[bb][aa], np.float64)
            f1 = np.fft.fft2(Qs, s=fshape)                        # FHWorkItem
            part2(bb, aa, f1, mY, mX)
            if tr:
                f2 = np.fft.fft2(Qs.T, s=fshape)                  # transposed FH
                part2(aa, bb, f2, mX, mY)                         # roles swapped
    return acc


def inverse_and_crop(acc, fshape, s1, s2):
    """IFFTWorkItem + CropAfterFFT (py_light_field.cpp:925-950;
    projector.py:792-796): c2r at fshape, then the 'same' crop."""
    full = np.fft.irfft2(acc, s=tuple(fshape), axes=(-2, -1))
    shape = tuple(int(a + b - 1) for a, b in zip(s1, (s2, s2)))
    full = full[..., :shape[0], :shape[1]]
    return shapes.centered(full, tuple(s1))


def project_special(src, Q, fshape=None, pad="5smooth"):
    """One plane through the special-FFT path, real-space result (T, Y, X)."""
    src = np.asarray(src, dtype=np.float64)
    single = src.ndim == 2
    if single:
        src = src[None]
    nnum = 2 * Q.shape[0] - 1
    s2 = Q.shape[2]
    if fshape is None:
        fshape, _, _ = shapes.convolution_shape(src.shape[1:], (s2, s2), nnum, pad)
    acc = project_special_fourier(src, Q, fshape)
    out = inverse_and_crop(acc, fshape, src.shape[1:], s2)
    return out[0] if single else out


def forward_project(vol, Qs, planes=None, method="special", pad="5smooth"):
    """Projector_allC.ForwardProjectACC semantics (projector.py:851-896):
    vol (nz, T, Y, X) -> image (T, Y, X) = sum over planes."""
    vol = np.asarray(vol, np.float64)
    planes = range(len(Qs)) if planes is None else planes
    out = None
    for cc in planes:
        if method == "special":
            r = project_special(vol[cc], Qs[cc], pad=pad)
        else:
            r = project_direct(vol[cc], expand_quadrant(np.asarray(Qs[cc], np.float64)))
        out = r if out is None else out + r
    return out


def backward_project(img, Qts, planes=None, method="special", pad="5smooth"):
    """Projector_allC.BackwardProjectACC semantics (projector.py:806-849):
    image (T, Y, X) -> volume (nz, T, Y, X); planes not listed stay zero."""
    img = np.asarray(img, np.float64)
    nz = len(Qts)
    out = np.zeros((nz,) + img.shape)
    planes = range(nz) if planes is None else planes
    for cc in planes:
        if method == "special":
            out[cc] = project_special(img, Qts[cc], pad=pad)
        else:
            out[cc] = project_direct(img, expand_quadrant(np.asarray(Qts[cc], np.float64)))
    return out
