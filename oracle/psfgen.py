"""Wave-optics light-field PSF synthesis in FP64 (oracle; test use only).

Restates GeneratePSF (generate_psf.py:63-798, Prevedel et al.'s model as the
reference implements it) for the parameter sets the tests use, in numpy /
scipy:

* ``debye``        the Debye integral of one point, I0(u, v), by adaptive
                   quadrature (scipy.integrate.quad, the reference's method and
                   tolerance, generate_psf.py:264-308 with the integrand of
                   light_field_integrands.pyx:12-39, J0 from scipy.special).
* ``psf_line``     the in-plane intensity profile used to size the PSF
                   (calcPSFFT, generate_psf.py:103-142, 160-167).
* ``ml_array``     the microlens phase array (calcML, generate_psf.py:169-205).
* ``wave_psf``     the wave PSF of one plane: an octant wedge of Debye values
                   completed by mirror and rotations (calcPSF_p,
                   generate_psf.py:310-350).
* ``fresnel2d``    angular-spectrum Fresnel propagation (generate_psf.py:245-259).
* ``pixel_binning`` (generate_psf.py:480-513).
* ``generate``     the light-field PSF of every lenslet offset of the quadrant
                   and plane (ComputeLFpsf, generate_psf.py:533-592), the clip
                   to zero and the per-PSF normalisation (594-613), the
                   per-plane crop (CAindex, 624-631 with psfmatrix.py:141-147).

Layout: ``generate`` returns one quadrant per plane in the HMatrix layout
Hcc[bb][aa][q][p] == H[p, q, aa, bb] of the generator (the .mat file stores
H.T, generate_psf.py:777, and psfmatrix.py:141-147 reads it back).

Pins (tests/test_psfgen.py): the small-aperture in-focus Debye integral
against the Airy closed form 2 J1(v)/v and the on-axis defocus against
sin(u/4)/(u/4); the microlens array against its closed form; the wave PSF's
8-fold symmetry; the Fresnel propagator against a direct quadratic-phase
convolution; binning against block sums.
"""
import numpy as np
import scipy.integrate
import scipy.signal
import scipy.special


class Optics(object):
    """Derived constants of GeneratePSF (generate_psf.py:81-96, 147-150)."""

    def __init__(self, M, NA, MLPitch, Nnum, OSR, n, fml, lam, zmin, zmax, zspacing):
        assert Nnum % 2 == 1 and OSR % 2 == 1
        self.M, self.NA, self.MLPitch, self.Nnum, self.OSR = M, NA, MLPitch, Nnum, OSR
        self.n, self.fml, self.lam = n, fml, lam
        self.k = 2 * np.pi * n / lam
        self.k0 = 2 * np.pi / lam
        self.fobj = 200e-3 / M
        self.alpha = np.arcsin(NA / n)
        self.pixelPitch = MLPitch / Nnum
        self.zs = np.arange(zmin, zmax + 0.1 * zspacing, zspacing)
        self.p3max = float(np.max(np.abs(self.zs)))


def debye(u, v, alpha, tol=1e-12, limit=180):
    """I0(u, v) = int_0^alpha sqrt(cos t) (1 + cos t) exp(-i (u/2) sin^2(t/2) /
    sin^2(alpha/2)) J0(v sin t / sin alpha) sin t dt (generate_psf.py:283-300,
    light_field_integrands.pyx:12-39)."""
    aF = np.sin(alpha / 2) ** (-2)
    vF = v / np.sin(alpha)

    def common(t):
        return np.sqrt(np.cos(t)) * (1 + np.cos(t)) * scipy.special.j0(np.sin(t) * vF) * np.sin(t)

    def re(t):
        return common(t) * np.cos(-(u / 2) * np.sin(t / 2) ** 2 * aF)

    def im(t):
        return common(t) * np.sin(-(u / 2) * np.sin(t / 2) ** 2 * aF)

    r, _ = scipy.integrate.quad(re, 0, alpha, limit=limit, epsabs=tol, epsrel=tol)
    i, _ = scipy.integrate.quad(im, 0, alpha, limit=limit, epsabs=tol, epsrel=tol)
    return r + 1j * i


def point_field(o, x1, x2, p3, tol=1e-12):
    """Koi * I0 at image point (x1, x2) for a point source on the axis at
    depth p3 (generate_psf.py:272-305 with p1 = p2 = 0)."""
    r = np.sqrt(x1 ** 2 + x2 ** 2) / o.M
    v = o.k * r * np.sin(o.alpha)
    u = 4 * o.k * p3 * np.sin(o.alpha / 2) ** 2
    Koi = o.M / (o.fobj * o.lam) ** 2 * np.exp(-1j * u / (4 * np.sin(o.alpha / 2) ** 2))
    return Koi * debye(u, v, o.alpha, tol=tol)


def psf_line(o):
    """Normalised intensity along x1 at the deepest plane, and the PSF size
    in lenslets it implies (generate_psf.py:158-167; the reference integrates
    this line with quad's default tolerance)."""
    xs = (o.pixelPitch / o.OSR) * np.arange(0, o.Nnum * o.OSR * 20 + 1)
    line = np.array([abs(point_field(o, x, 0.0, o.p3max, tol=1.49e-8)) ** 2 for x in xs])
    line = line / line.max()
    out = np.nonzero(line < 0.04)[0]
    if len(out) == 0:
        raise ValueError("Estimated PSF size exceeds the limit")
    return line, int(np.ceil(out[0] / (o.OSR * o.Nnum)))


class Grid(object):
    """Oversampled grid and crop windows (generate_psf.py:207-221, 352-355,
    517-519)."""

    def __init__(self, o, imgsize_ref):
        N, OSR = o.Nnum, o.OSR
        self.imgsize_ref = imgsize_ref
        self.half = max(N * (imgsize_ref + 1), 2 * N)                   # IMG_HALFWIDTH
        self.x = (o.pixelPitch / OSR) * np.arange(-self.half * OSR, self.half * OSR + 0.1, 1)
        self.L = len(self.x)
        self.centre = (self.L + 1) // 2                                 # 1-based (centerPT_m)
        self.hw = N * imgsize_ref * OSR                                  # halfWidth
        self.xml = (o.pixelPitch / OSR) * np.arange(-(N * OSR - 1) / 2, (N * OSR - 1) / 2 + 0.1, 1)
        self.cp = np.arange((self.centre - 1) / OSR - self.hw / OSR, (self.centre - 1) / OSR + self.hw / OSR + 0.1,
                            1).astype(int)                               # CP_p

    def plane_lo(self, o, p3):
        """First row of the wedge computed for plane p3 (generate_psf.py:363-367)."""
        ref_il = np.ceil(self.imgsize_ref * (abs(p3) / o.p3max))
        hw_il = max(o.Nnum * (ref_il + 0) * o.OSR, 2 * o.Nnum * o.OSR)
        return int(max(self.centre - hw_il, 1) - 1)


def ml_array(o, g):
    """calcML (generate_psf.py:169-205) with k0: lens phase exp(-i k0 |x|^2 /
    (2 fml)) around every lenslet centre, by 'same' convolution of the centre
    grid with one lens."""
    pat = np.exp(-1j * o.k0 / (2 * o.fml) * (g.xml[:, None] ** 2 + g.xml[None, :] ** 2))
    c0 = int(np.nonzero(g.x == 0)[0][0])
    step = len(g.xml)
    cs = np.sort(np.append(np.arange(c0, -1, -step), np.arange(c0 + step, g.L, step)))
    centres = np.zeros((g.L, g.L))
    centres[np.ix_(cs, cs)] = 1
    return scipy.signal.fftconvolve(centres.astype(complex), pat, "same")


def wave_psf(o, g, p3, lo=None):
    """calcPSF_p (generate_psf.py:310-350): Debye values on rows lo..centre-1
    and columns a..centre-1, mirrored to the right half of the top rows and
    completed by the 90/180/270-degree rotations where still zero."""
    L, C = g.L, g.centre
    if lo is None:
        lo = g.plane_lo(o, p3)
    pat = np.zeros((L, L), complex)
    for a in range(lo, C):
        for b in range(a, C):
            pat[a, b] = point_field(o, g.x[a], g.x[b], p3)
    top = pat[:C, :C]
    p0 = pat.copy()
    p0[:C, C - 1:] = np.fliplr(top)
    out = p0.copy()
    for k in (-1, -2, -3):
        r = np.rot90(p0, k)
        z = out == 0
        out[z] = r[z]
    return out


def fresnel2d(f0, dx0, z, lam):
    """generate_psf.py:245-259 (vacuum wavenumber)."""
    Nx, Ny = f0.shape
    k = 2 * np.pi / lam
    u = np.append(np.arange(0, np.ceil(Nx / 2)), np.arange(np.ceil(-Nx / 2), 0)) / (Nx * dx0)
    v = np.append(np.arange(0, np.ceil(Ny / 2)), np.arange(np.ceil(-Ny / 2), 0)) / (Ny * dx0)
    H = np.exp(-1j * 2 * np.pi ** 2 * (u[:, None] ** 2 + v[None, :] ** 2) * z / k)
    return np.exp(1j * k * z) * np.fft.ifft2(np.fft.fft2(f0) * H)


def shift_zero(img, sx, sy):
    """im_shift2 (generate_psf.py:457-478): integer shift with zero fill."""
    out = np.zeros_like(img)
    X, Y = img.shape
    src = img[max(0, -sx):X - max(0, sx), max(0, -sy):Y - max(0, sy)]
    out[max(0, sx):X - max(0, -sx), max(0, sy):Y - max(0, -sy)] = src
    return out


def pixel_binning(S, OSR):
    """generate_psf.py:480-513: crop around the centre and sum OSR x OSR
    blocks.  Returns (binned, shift0, shift1)."""
    n0, n1 = S.shape
    out = []
    crops = []
    for n in (n0, n1):
        c = (n - 1) // 2 + 1
        ci = c - (OSR - 1) // 2
        init = ci - (ci // OSR) * OSR
        shift = 0
        if init < 1:
            init += OSR
            shift = 1
        hw = c - init
        crops.append((init - 1, c + hw))
        out.append(shift)
    C = S[crops[0][0]:crops[0][1], crops[1][0]:crops[1][1]]
    m, k = C.shape
    return C.reshape(m // OSR, OSR, k // OSR, OSR).sum(axis=(1, 3)), out[0], out[1]


def lf_psf(o, g, psf, ml, a, b):
    """ComputeLFpsf for lenslet offset (a, b) of one plane (generate_psf.py:
    541-567): shift, multiply by the lens array, propagate by fml, shift
    back, aperture, |.|^2, bin, crop to CP_p.  Returns H[:, :, a, b]."""
    h = (o.Nnum + 1) // 2
    sa, sb = o.OSR * (a + 1 - h), o.OSR * (b + 1 - h)
    f1 = fresnel2d(shift_zero(psf, sa, sb) * ml, o.pixelPitch / o.OSR, o.fml, o.lam)
    f1 = shift_zero(f1, -sa, -sb)
    c0 = g.centre - 1
    lo, hi = max(c0 - g.hw, 0), min(c0 + g.hw, g.L - 1)
    ap = np.zeros_like(f1)
    ap[lo:hi + 1, lo:hi + 1] = f1[lo:hi + 1, lo:hi + 1]
    binned, s0, s1 = pixel_binning(np.abs(ap ** 2), o.OSR)
    return binned[np.ix_(g.cp - s0, g.cp - s1)]


def generate(o, imgsize_ref=None, planes=None, clip=0.005):
    """The normalised light-field PSF quadrants of the listed planes
    (default all), each cropped to its CAindex window, in the HMatrix layout
    (list of (h, h, s_z, s_z) float64 arrays), plus the grid."""
    if imgsize_ref is None:
        imgsize_ref = psf_line(o)[1]
    g = Grid(o, imgsize_ref)
    ml = ml_array(o, g)
    h = (o.Nnum + 1) // 2
    planes = range(len(o.zs)) if planes is None else planes
    s2 = len(g.cp)
    centre_cp = int(np.ceil(s2 / 2))
    out = []
    for c in planes:
        p3 = o.zs[c]
        psf = wave_psf(o, g, p3)
        H = np.zeros((s2, s2, h, h))
        for a in range(h):
            for b in range(h):
                H[:, :, a, b] = lf_psf(o, g, psf, ml, a, b)
        H[H < clip * H.max()] = 0                                   # generate_psf.py:594-599 (per plane)
        H = H / H.sum(axis=(0, 1), keepdims=True)                   # generate_psf.py:606-613 (per PSF)
        ref_il = np.ceil(imgsize_ref * (abs(p3) / o.p3max))
        hw_il = max(o.Nnum * (ref_il + 0), 2 * o.Nnum)
        ca0 = int(max(centre_cp - hw_il, 1))
        ca1 = int(min(centre_cp + hw_il, s2))
        H = H[ca0 - 1:ca1, ca0 - 1:ca1]
        out.append(H.transpose(3, 2, 1, 0).copy())                   # Hcc[bb][aa][q][p]
    return out, g
