"""Pins of the FP64 PSF-synthesis oracle (CPU) and of the product's host-side
geometry.  Each check is a closed form or an independent computation, chosen
so that a plausible mistake in the restatement (a dropped factor of the
Debye integrand, a wrong rotation or shift direction, a transposed window)
fails it."""
import numpy as np
import pytest
import scipy.special

from oracle import psfgen as pg

SMALL = (20, 0.3, 100e-6, 5, 3, 1.0, 2000e-6, 520e-9, -4e-6, 4e-6, 2e-6)


def test_debye_in_focus_is_airy_at_small_aperture():
    # u = 0, small alpha: I0 -> 2 sin^2(alpha) J1(v)/v (paraxial limit of
    # generate_psf.py:283-300), relative corrections O(alpha^2)
    alpha = 0.02
    for v in [0.3, 1.0, 2.5, 3.8317, 5.0, 7.0156, 9.0]:
        got = pg.debye(0.0, v, alpha)
        want = 2 * np.sin(alpha) ** 2 * scipy.special.j1(v) / v
        assert abs(got.imag) < 1e-12 * abs(got.real) + 1e-18
        assert abs(got.real - want) < 1e-3 * 2 * np.sin(alpha) ** 2 * 0.5, (v, got, want)


def test_debye_on_axis_defocus_is_sinc():
    # v = 0, small alpha: |I0(u, 0) / I0(0, 0)| -> |sin(u/4) / (u/4)|
    alpha = 0.02
    i00 = abs(pg.debye(0.0, 0.0, alpha))
    for u in [1.0, 4.0, 8.0, 4 * np.pi, 20.0]:
        r = abs(pg.debye(u, 0.0, alpha)) / i00
        assert abs(r - abs(np.sin(u / 4) / (u / 4))) < 2e-3, (u, r)


def test_debye_planted_error_fails():
    # dropping the sin(t) measure factor of the integrand breaks the Airy limit
    alpha, v = 0.02, 2.5
    aF = np.sin(alpha / 2) ** (-2)
    import scipy.integrate
    bad = scipy.integrate.quad(lambda t: np.sqrt(np.cos(t)) * (1 + np.cos(t)) * scipy.special.j0(np.sin(t) * v /
                                                                                             np.sin(alpha)), 0, alpha)[0]
    want = 2 * np.sin(alpha) ** 2 * scipy.special.j1(v) / v
    assert abs(bad - want) > 0.5 * abs(want)
    assert abs(pg.debye(0.0, v, alpha).real - want) < 1e-3 * abs(want) + 1e-9


def test_ml_array_closed_form():
    # calcML's 'same' convolution of the centre grid with one lens equals the
    # nearest-centre lens phase exp(-i k0 |d dx|^2 / (2 fml))
    o = pg.Optics(*SMALL)
    g = pg.Grid(o, 1)
    ml = pg.ml_array(o, g)
    K = o.Nnum * o.OSR
    dx = o.pixelPitch / o.OSR
    i = np.arange(g.L)
    d = (i + (K - 1) // 2) % K - (K - 1) // 2
    want = np.exp(-1j * o.k0 / (2 * o.fml) * dx * dx * (d[:, None] ** 2 + d[None, :] ** 2))
    assert np.max(np.abs(ml - want)) < 1e-12


def test_wave_psf_symmetry_and_values():
    # the wedge + mirror + rotation fill gives a field with the 8-fold
    # symmetry of a radial function, equal to the Debye value at every point
    o = pg.Optics(*SMALL)
    g = pg.Grid(o, 1)
    p3 = o.zs[0]
    psf = pg.wave_psf(o, g, p3)
    s = np.max(np.abs(psf))
    for T in (psf.T, psf[::-1], psf[:, ::-1], np.rot90(psf)):
        assert np.max(np.abs(T - psf)) < 1e-12 * s
    rng = np.random.default_rng(0)
    for _ in range(6):
        a, b = rng.integers(0, g.L, 2)
        assert abs(psf[a, b] - pg.point_field(o, g.x[a], g.x[b], p3)) < 1e-10 * s


def test_wave_psf_rows_outside_the_wedge_stay_zero():
    o = pg.Optics(*SMALL)
    g = pg.Grid(o, 1)
    psf = pg.wave_psf(o, g, o.zs[0], lo=5)
    assert np.all(psf[:5] == 0) and np.all(psf[:, :5] == 0) and np.all(psf[-5:] == 0) and np.all(psf[:, -5:] == 0)
    assert np.all(psf[5:-5, 5:-5] != 0)


def test_fresnel_is_two_circular_convolutions():
    # angular-spectrum propagation with the separable even transfer function
    # equals exp(i k z) P f0 P^T with p = IDFT of the 1-D transfer function
    rng = np.random.default_rng(1)
    L, dx, z, lam = 9, 1e-6, 5e-5, 5e-7
    f0 = rng.random((L, L)) + 1j * rng.random((L, L))
    got = pg.fresnel2d(f0, dx, z, lam)
    k = 2 * np.pi / lam
    f = np.array([j if j < (L + 1) // 2 else j - L for j in range(L)]) / (L * dx)
    Hu = np.exp(-1j * 2 * np.pi ** 2 * f ** 2 * z / k)
    p = np.array([np.sum(Hu * np.exp(2j * np.pi * np.arange(L) * m / L)) / L for m in range(L)])
    P = np.array([[p[(m - n) % L] for n in range(L)] for m in range(L)])
    want = np.exp(1j * k * z) * P @ f0 @ P.T
    assert np.max(np.abs(got - want)) < 1e-12 * np.max(np.abs(want))


def test_shift_and_binning_definitions():
    a = np.arange(35.0).reshape(5, 7)
    for sx, sy in [(2, -1), (-3, 0), (0, 4), (-1, -2)]:
        out = pg.shift_zero(a, sx, sy)
        for i in range(5):
            for j in range(7):
                src = (i - sx, j - sy)
                want = a[src] if 0 <= src[0] < 5 and 0 <= src[1] < 7 else 0
                assert out[i, j] == want
    # OSR 3 on a 2*3*H + 1 grid: crop from index 2, 3 x 3 block sums, shift 1
    S = np.arange(31.0 * 31).reshape(31, 31)
    B, s0, s1 = pg.pixel_binning(S, 3)
    assert (s0, s1) == (1, 1) and B.shape == (9, 9)
    assert B[0, 0] == S[2:5, 2:5].sum() and B[4, 4] == S[14:17, 14:17].sum()
    B1, t0, t1 = pg.pixel_binning(S, 1)
    assert (t0, t1) == (1, 1) and B1.shape == (31, 31)


def test_generate_small_normalised_and_symmetric():
    o = pg.Optics(*SMALL)
    H, g = pg.generate(o, imgsize_ref=1, planes=[0, 4])
    for Q in H:
        np.testing.assert_allclose(Q.sum(axis=(2, 3)), 1.0, rtol=1e-12)
        h = Q.shape[0]
        # the quadrant's transpose symmetry Q[a][b] = Q[b][a].T (radial wave PSF)
        for b in range(h):
            for a in range(h):
                assert np.max(np.abs(Q[a, b] - Q[b, a].T)) < 1e-9 * Q.max()
    # the microlens array breaks the +-z conjugate symmetry of the wave PSF:
    # the light-field PSFs above and below focus differ
    assert np.max(np.abs(H[0] - H[1])) > 1e-3 * H[0].max()


def test_product_geometry_matches_oracle():
    from flfhip import psfgen
    o = pg.Optics(*SMALL)
    for ref in (1, 2, 4):
        g = pg.Grid(o, ref)
        geo = psfgen.Geometry(o.Nnum, o.OSR, ref, o.p3max)
        assert (geo.half, geo.L, geo.centre, geo.s2) == (g.half, g.L, g.centre, len(g.cp))
        for p3 in o.zs:
            assert geo.lo(p3) == g.plane_lo(o, p3)
            lo, hi = geo.ca_window(p3)
            c = int(np.ceil(len(g.cp) / 2))
            ref_il = np.ceil(ref * abs(p3) / o.p3max)
            hw = max(o.Nnum * ref_il, 2 * o.Nnum)
            assert (lo, hi) == (int(max(c - hw, 1)) - 1, int(min(c + hw, len(g.cp))))
    np.testing.assert_allclose(psfgen.plane_depths(-4e-6, 4e-6, 2e-6), o.zs)
