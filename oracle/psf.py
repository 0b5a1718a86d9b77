"""PSF-set symmetry and adjoint (oracle; test use only).

The reference stores one quadrant ``Q[b][a]`` (b <= a < h, h = (N+1)/2) per
plane and derives the other offsets by mirror and transpose
(py_light_field.cpp:1514-1562; projector.py:159-189):

* transpose, pixel (a, b):      Q[b][a].T
* mirrorX, pixel (N-1-b, .):    row flip (flipud)
* mirrorY, pixel (., N-1-a):    column flip (fliplr)

``Q[a][b]`` for a > b is never read by the reference (py_light_field.cpp:
1514-1521), so ``expand_quadrant`` defines the full set from the upper
triangle only.
"""
import numpy as np


def fold(i, nnum):
    """Index of the quadrant row/column an offset maps to, and whether a flip
    is applied (the mirror rule of py_light_field.cpp:1516-1519)."""
    c = nnum // 2
    return (i, False) if i <= c else (nnum - 1 - i, True)


def expand_quadrant(Q):
    """Full PSF set H[b][a] (N, N, s2, s2) from the quadrant Q (h, h, s2, s2)."""
    h = Q.shape[0]
    nnum = 2 * h - 1
    s2 = Q.shape[2]
    H = np.zeros((nnum, nnum, s2, s2), dtype=Q.dtype)
    for b in range(nnum):
        bq, fb = fold(b, nnum)
        for a in range(nnum):
            aq, fa = fold(a, nnum)
            M = Q[bq][aq] if bq <= aq else Q[aq][bq].T
            if fb:
                M = M[::-1, :]
            if fa:
                M = M[:, ::-1]
            H[b, a] = M
    return H


def quadrant_of(H):
    """The (h, h, s2, s2) quadrant the reference stores; entries with a < b
    are filled with the transpose of Q[a][b] so the array is fully defined
    (the reference never reads them)."""
    nnum = H.shape[0]
    h = (nnum + 1) // 2
    Q = np.zeros((h, h) + H.shape[2:], dtype=H.dtype)
    for b in range(h):
        for a in range(h):
            Q[b, a] = H[b, a]
    return Q


def adjoint_psf(H):
    """PSF set Ht such that back-projection with Ht is the exact adjoint of
    forward projection with H.

    Forward (projector.py:1084-1104):  img[p] = sum_q vol[q] H_{q mod N}[p-q+c2]
    Backward (projector.py:1106-1126): bk[q]  = sum_p img[p] Ht_{p mod N}[q-p+c2]
    Equality of <F x, y> and <x, B y> for all x, y requires
        Ht_a[e] = H_{(a + e - c2) mod N}[2*c2 - e]      (2-D indices).
    This is the same relation the reference's calcHt_new builds by stamping
    rotated PSFs (generate_psf.py:646-719); the formula here holds for any
    odd s2, including sizes that are not 2kN+1.
    """
    nnum = H.shape[0]
    s2 = H.shape[2]
    assert s2 % 2 == 1, "PSF side must be odd"
    c2 = (s2 - 1) // 2
    e = np.arange(s2)
    Ht = np.zeros_like(H)
    for b in range(nnum):
        src_b = (b + e - c2) % nnum            # per output row e0
        for a in range(nnum):
            src_a = (a + e - c2) % nnum        # per output column e1
            # Ht[b,a][e0,e1] = H[src_b[e0], src_a[e1]][2c2-e0, 2c2-e1]
            Ht[b, a] = H[src_b[:, None], src_a[None, :], (2 * c2 - e)[:, None], (2 * c2 - e)[None, :]]
    return Ht


def mat_layout(Q):
    """Quadrant in the generator's own layout from the HMatrix layout.

    generate_psf.py:533 builds H[p, q, aa, bb, z] and saves ``H.T`` to HDF5
    (generate_psf.py:777); h5py reads it back as _H[z, bb, aa, q, p] and
    LoadRawMatrixData keeps _H[z, :h, :h] (psfmatrix.py:141-146), so
    Hcc[bb][aa][q][p] == H[p, q, aa, bb].  Returns H[p, q, aa, bb] for one
    plane."""
    return np.asarray(Q).transpose(3, 2, 1, 0)


def hcc_layout(Hpq):
    """Inverse of ``mat_layout``."""
    return np.asarray(Hpq).transpose(3, 2, 1, 0)


def calc_ht_stamp(Hpq):
    """Restatement of calcHt_new with backwardProject_new
    (generate_psf.py:646-719), FP64, one plane.

    Hpq is one plane in the generator's layout, H[p, q, aa, bb] with
    aa, bb < h (the full quadrant, generate_psf.py:533).  For each target
    offset (a0, b0) the reference back-projects a single lit pixel by
    stamping the 180-degree-rotated PSF of every lenslet offset (aa, bb) at
    the positions the lenslet sub-sampling would read (generate_psf.py:
    661-691), cuts an s2 x s2 window around the image centre (711-712) and
    shifts it by (h-1-a0, h-1-b0) pixels with zero fill (im_shift2,
    generate_psf.py:457-478; both shifts are >= 0 here).  Valid only for
    PSF sides s2 = 2kN+1, the reference's own assertion (generate_psf.py:
    680)."""
    Hpq = np.asarray(Hpq, dtype=np.float64)
    s2, _, h, _ = Hpq.shape
    nnum = 2 * h - 1
    c2 = s2 // 2
    if c2 % nnum != 0:
        raise ValueError("calcHt_new needs s2 = 2kN+1 (generate_psf.py:680)")
    blocks = -(-s2 // nnum)
    side = (blocks + 2) * nnum if blocks % 2 == 1 else (blocks + 3) * nnum
    centre = -(-side // 2)
    origin = centre - (-(-nnum // 2))
    rot = Hpq[::-1, ::-1]
    out = np.zeros_like(Hpq)
    for a0 in range(h):
        for b0 in range(h):
            canvas = np.zeros((side, side))
            for aa in range(nnum):
                K = rot[:, :, aa, :] if aa <= nnum // 2 else rot[::-1, :, nnum - 1 - aa, :]
                for bb in range(nnum):
                    Kab = K[:, :, bb] if bb <= nnum // 2 else K[:, ::-1, nnum - 1 - bb]
                    sa = (aa - a0) % nnum
                    sb = (bb - b0) % nnum
                    sub = Kab[sa::nnum, sb::nnum]
                    da = origin + a0 - c2 + sa
                    db = origin + b0 - c2 + sb
                    canvas[da:da + sub.shape[0] * nnum:nnum, db:db + sub.shape[1] * nnum:nnum] = sub
            lo, hi = centre - c2 - 1, centre + c2
            cut = canvas[lo:hi, lo:hi]
            sx, sy = h - 1 - a0, h - 1 - b0
            shifted = np.zeros_like(cut)
            shifted[sx:, sy:] = cut[:s2 - sx, :s2 - sy]
            out[:, :, a0, b0] = shifted
    return out
