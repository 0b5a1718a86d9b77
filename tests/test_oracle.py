"""Pins of the FP64 oracle (CPU).  Each check is chosen so that a plausible
mistake in the oracle (dropped term, wrong sign/index, transposed operand,
wrong crop offset) fails it:

* the reference's own known answers for the Fourier shapes
  (py_light_field.cpp:1904-1925);
* the special-FFT restatement against the direct per-offset convolution (two
  different algorithms), square / non-square / odd / even grids;
* the direct convolution against an explicit shift-and-add loop;
* the adjoint identity <F x, y> = <x, B y> for the derived Ht;
* the symmetry expansion producing each of the N*N offsets from the right
  quadrant entry with the right flip;
* RL closed forms.
"""
import numpy as np
import pytest

from oracle import projector as oproj
from oracle import psf as opsf
from oracle import rl as orl
from oracle import shapes as oshapes


def test_shape_known_answers():
    # py_light_field.cpp:1904-1925: (450,675) image, 391 and 61 px PSFs, minimal padding.
    f, fs, s1 = oshapes.convolution_shape((450, 675), (391, 391), 15, "minimal")
    assert list(f) == [840, 1065]
    assert [s.stop for s in fs] == [840, 1065]
    f, _, _ = oshapes.convolution_shape((450, 675), (61, 61), 15, "minimal")
    assert list(f) == [510, 735]
    # 7-smooth (CPU default): 56 = 2^3*7 stays, 71 (prime) -> 72.
    f, _, _ = oshapes.convolution_shape((450, 675), (391, 391), 15, "7smooth")
    assert list(f) == [840, 1080]
    # 5-smooth: 56 -> 60, 71 -> 72.
    f, _, _ = oshapes.convolution_shape((450, 675), (391, 391), 15, "5smooth")
    assert list(f) == [900, 1080]
    # Batched input shape (T, Y, X) drops the leading axis.
    f2, _, _ = oshapes.convolution_shape((3, 450, 675), (391, 391), 15, "minimal")
    assert list(f2) == [840, 1065]


def test_centered_crop():
    a = np.arange(7 * 9).reshape(7, 9)
    c = oshapes.centered(a, (3, 5))
    assert c.shape == (3, 5) and c[0, 0] == a[2, 2]


def _rand_q(rng, N, s2):
    h = (N + 1) // 2
    return rng.random((h, h, s2, s2))


@pytest.mark.parametrize("N,Y,X,s2,pad", [
    (3, 12, 12, 5, "minimal"), (5, 20, 25, 7, "5smooth"), (5, 25, 20, 9, "7smooth"),
    (15, 60, 45, 31, "5smooth"), (15, 45, 45, 15, "minimal"), (7, 28, 35, 11, "5smooth")])
def test_special_matches_direct(N, Y, X, s2, pad):
    rng = np.random.default_rng(N * 1000 + Y)
    Q = _rand_q(rng, N, s2)
    src = rng.random((2, Y, X))
    d = oproj.project_direct(src, opsf.expand_quadrant(Q))
    s = oproj.project_special(src, Q, pad=pad)
    assert np.max(np.abs(d - s)) < 1e-10 * np.max(np.abs(d))


def test_direct_matches_bruteforce():
    rng = np.random.default_rng(3)
    for N, Y, X, s2 in [(3, 9, 12, 5), (5, 10, 10, 3)]:
        Q = _rand_q(rng, N, s2)
        H = opsf.expand_quadrant(Q)
        src = rng.random((1, Y, X))
        assert np.max(np.abs(oproj.project_direct(src, H) - oproj.project_bruteforce(src, H))) < 1e-12


def test_special_detects_planted_error():
    # A wrong mirror phase (plant: drop the +1 of the mirror multiplier) must
    # break agreement: the special path is sensitive to the flip conventions.
    rng = np.random.default_rng(5)
    N, Y, X, s2 = 5, 20, 20, 7
    Q = _rand_q(rng, N, s2)
    src = rng.random((1, Y, X))
    good = oproj.project_special(src, Q, fshape=(30, 30))
    orig = oproj._multipliers

    def bad(fshape, s2_, nnum):
        mX, mY, eX, eY = orig(fshape, s2_, nnum)
        return mX * np.exp(-2j * np.pi * np.arange(fshape[0]) / fshape[0]), mY, eX, eY

    oproj._multipliers = bad
    try:
        wrong = oproj.project_special(src, Q, fshape=(30, 30))
    finally:
        oproj._multipliers = orig
    d = oproj.project_direct(src, opsf.expand_quadrant(Q))
    assert np.max(np.abs(good - d)) < 1e-10
    assert np.max(np.abs(wrong - d)) > 1e-3


def test_adjoint_identity():
    rng = np.random.default_rng(7)
    for N, Y, X, s2 in [(5, 20, 25, 7), (15, 45, 30, 31)]:
        H = opsf.expand_quadrant(_rand_q(rng, N, s2))
        Ht = opsf.adjoint_psf(H)
        x = rng.random((1, Y, X))
        y = rng.random((1, Y, X))
        lhs = np.sum(oproj.project_direct(x, H) * y)
        rhs = np.sum(x * oproj.project_direct(y, Ht))
        assert abs(lhs - rhs) < 1e-11 * abs(lhs)


def test_expand_quadrant_rules():
    N, s2 = 7, 5
    h = (N + 1) // 2
    Q = np.arange(h * h * s2 * s2, dtype=np.float64).reshape(h, h, s2, s2)
    H = opsf.expand_quadrant(Q)
    c = N // 2
    seen = set()
    for b in range(h):
        for a in range(b, h):
            np.testing.assert_array_equal(H[b, a], Q[b, a])
            if b != c:   # mirrorX (py_light_field.cpp:1518)
                np.testing.assert_array_equal(H[N - 1 - b, a], Q[b, a][::-1, :])
            if a != c:   # mirrorY (py_light_field.cpp:1519)
                np.testing.assert_array_equal(H[b, N - 1 - a], Q[b, a][:, ::-1])
            if a != c and b != c:
                np.testing.assert_array_equal(H[N - 1 - b, N - 1 - a], Q[b, a][::-1, ::-1])
            if a != b and a != c:   # the reference transposes only off-diagonal entries
                np.testing.assert_array_equal(H[a, b], Q[b, a].T)
                np.testing.assert_array_equal(H[N - 1 - a, b], Q[b, a].T[::-1, :])
            for (bb, aa) in [(b, a), (a, b), (N - 1 - b, a), (b, N - 1 - a), (N - 1 - b, N - 1 - a),
                             (N - 1 - a, b), (a, N - 1 - b), (N - 1 - a, N - 1 - b)]:
                seen.add((bb, aa))
    assert len(seen) == N * N
    assert c == h - 1


def test_adjoint_of_symmetric_set_is_symmetric():
    # A radially symmetric PSF set keeps the 8-fold symmetry in Ht, so the
    # reference's quadrant storage of Ht is valid for it.
    from flfhip import synthetic
    rng = np.random.default_rng(0)
    Q = synthetic.ring_quadrant(7, 21, 4.0, 1.0, rng).astype(np.float64)
    H = opsf.expand_quadrant(Q)
    Ht = opsf.adjoint_psf(H)
    assert np.max(np.abs(opsf.expand_quadrant(opsf.quadrant_of(Ht)) - Ht)) < 1e-15


def test_rl_identity_psf_closed_form():
    # PSF = centred delta for every offset -> F = B = identity on one plane;
    # one Prevedel iteration from X0 gives X1 = X0 * Htf / X0 = im.
    N, s2 = 5, 3
    h = (N + 1) // 2
    Q = np.zeros((h, h, s2, s2))
    Q[:, :, 1, 1] = 1.0
    rng = np.random.default_rng(2)
    im = rng.random((1, 10, 10)) + 0.5
    fwd = lambda v: oproj.forward_project(v, [Q])
    bwd = lambda i: oproj.backward_project(i, [Q])
    X0 = rng.random((1, 1, 10, 10)) + 0.1
    X1 = orl.deconv_rl(fwd, bwd, None, 1, Xguess=X0, im=im)
    np.testing.assert_allclose(X1[0], im, rtol=1e-12)


def test_rl_nan_to_zero_and_fit():
    X = np.array([0.0, 1.0, 2.0])
    out = orl.rl_update_prevedel(X.copy(), np.array([0.0, 1.0, 1.0]), np.array([0.0, 0.0, 2.0]))
    assert out[0] == 0.0 and np.isinf(out[1]) and out[2] == 1.0
    fit = orl.data_fit(np.ones((2, 3, 3)), 2 * np.ones((2, 3, 3)))
    np.testing.assert_allclose(fit, [0.5, 0.5])


def test_rl_reduces_misfit():
    from flfhip import synthetic
    N = 5
    zs = [-1, 0, 1]
    hm = synthetic.make_psf_set(N, zs, [7, 5, 7], seed=1)
    Qs = [hm.Hcc(z, False).astype(np.float64) for z in range(3)]
    Qts = [hm.Hcc(z, True).astype(np.float64) for z in range(3)]
    vol = synthetic.blob_volume(3, 30, 30, 6, seed=1, sigma=(1.0, 2.0)).astype(np.float64)
    im = oproj.forward_project(vol, Qs)
    fwd = lambda v: oproj.forward_project(v, Qs)
    bwd = lambda i: oproj.backward_project(i, Qts)
    _, fits = orl.deconv_rl(fwd, bwd, None, 6, Xguess=np.ones_like(vol), im=im, record_fit=True)
    fits = [f[0] for f in fits]
    assert fits[-1] < fits[0] * 0.7


def test_rl_standard_identity_psf_closed_form():
    # Standard update (lfdeconv.py:151-164, 191-193) with F = B = identity:
    # X1 = X0 * B(im / F X0) = X0 * im / X0 = im for any positive X0; two
    # frames share the one plane.
    N, s2 = 5, 3
    h = (N + 1) // 2
    Q = np.zeros((h, h, s2, s2))
    Q[:, :, 1, 1] = 1.0
    rng = np.random.default_rng(12)
    im = rng.random((2, 10, 10)) + 0.5
    fwd = lambda v: oproj.forward_project(v, [Q])
    bwd = lambda i: oproj.backward_project(i, [Q])
    X0 = rng.random((1, 2, 10, 10)) + 0.1
    X1 = orl.deconv_rl(fwd, bwd, None, 1, Xguess=X0, im=im, prevedel=False)
    np.testing.assert_allclose(X1[0], im, rtol=1e-12)
    # a second iteration is a fixed point
    X2 = orl.deconv_rl(fwd, bwd, None, 2, Xguess=X0, im=im, prevedel=False)
    np.testing.assert_allclose(X2[0], im, rtol=1e-12)


def _loops_project(src, Kfull):
    """Per-plane operator out[p] = sum_q src[q] K_{q mod N}[p - q + c2]
    written with explicit loops (no library convolution)."""
    Y, X = src.shape
    nnum, s2 = Kfull.shape[0], Kfull.shape[2]
    c2 = s2 // 2
    out = np.zeros((Y, X))
    for q0 in range(Y):
        for q1 in range(X):
            for p0 in range(max(0, q0 - c2), min(Y, q0 + c2 + 1)):
                for p1 in range(max(0, q1 - c2), min(X, q1 + c2 + 1)):
                    out[p0, p1] += src[q0, q1] * Kfull[q0 % nnum, q1 % nnum, p0 - q0 + c2, p1 - q1 + c2]
    return out


def _trust_case():
    # N = 3, one plane, 9 x 9 image, 5 px PSF; a trust mask with a zero
    # block, and an image that is zero there (the reference asserts this,
    # lfdeconv.py:106-109).  The block is small enough that B(trust) > 0
    # everywhere (every pixel has trusted neighbours inside the PSF), so no
    # 0/0 arises whose FFT rounding would decide the result.
    N, s2, Y = 3, 5, 9
    h = (N + 1) // 2
    rng = np.random.default_rng(31)
    Q = rng.random((h, h, s2, s2)) + 0.05
    H = opsf.expand_quadrant(Q)
    # back projection applies the stored Ht quadrant expanded by the same
    # symmetry rules (projector.py:806-849)
    Ht = opsf.expand_quadrant(opsf.quadrant_of(opsf.adjoint_psf(H)))
    trust = np.ones((Y, Y))
    trust[1:3, 5:7] = 0.0
    im = (rng.random((1, Y, Y)) + 0.2) * trust
    X0 = rng.random((1, 1, Y, Y)) + 0.3
    return Q, H, Ht, trust, im, X0


def _trust_expected(H, Ht, trust, im, X0):
    Y, X = im.shape[1:]
    HX = _loops_project(X0[0, 0], H)
    err = np.zeros((Y, X))
    for i in range(Y):
        for j in range(X):
            err[i, j] = im[0, i, j] / HX[i, j]
    Berr = _loops_project(err, Ht)
    Btrust = _loops_project(trust, Ht)
    X1 = np.zeros((Y, X))
    for i in range(Y):
        for j in range(X):
            v = X0[0, 0, i, j] * Berr[i, j] / Btrust[i, j]
            X1[i, j] = 0.0 if np.isnan(v) else v
    return X1


def test_rl_standard_trust_mask_hand_built():
    # deconv_rl(prevedel=False, trust_mask=...) against X0 * B(im / F X0) /
    # B(trust) with every operator and elementwise step as explicit loops
    # (lfdeconv.py:106-110, 151-164, 191-193).
    Q, H, Ht, trust, im, X0 = _trust_case()
    Qt = opsf.quadrant_of(Ht)
    fwd = lambda v: oproj.forward_project(v, [Q])
    bwd = lambda i: oproj.backward_project(i, [Qt])
    got = orl.deconv_rl(fwd, bwd, None, 1, Xguess=X0, im=im, prevedel=False, trust_mask=trust)
    want = _trust_expected(H, Ht, trust, im, X0)
    assert np.max(np.abs(got[0, 0] - want)) < 1e-12 * np.max(np.abs(want))
    assert np.all(want[1:3, 5:7] > 0)      # the untrusted pixels are still updated


def test_rl_standard_planted_double_trust_fails():
    # Planting a plausible mistake in the oracle (the trust normalisation
    # applied twice) must break the hand-built check above.
    Q, H, Ht, trust, im, X0 = _trust_case()
    Qt = opsf.quadrant_of(Ht)
    fwd = lambda v: oproj.forward_project(v, [Q])
    bwd = lambda i: oproj.backward_project(i, [Qt])
    orig = orl.rl_update_standard

    def twice(X, back, trust_back=None):
        X = orig(X, back, trust_back)
        return X / trust_back if trust_back is not None else X

    orl.rl_update_standard = twice
    try:
        bad = orl.deconv_rl(fwd, bwd, None, 1, Xguess=X0, im=im, prevedel=False, trust_mask=trust)
    finally:
        orl.rl_update_standard = orig
    want = _trust_expected(H, Ht, trust, im, X0)
    assert np.max(np.abs(bad[0, 0] - want)) > 1e-3 * np.max(np.abs(want))


def test_rl_standard_division_semantics():
    # lfdeconv.py:156-164: X *= back; X /= trustBack; NaN -> 0.  So
    # 0 * b / 0 = NaN -> 0, x / 0 = inf (kept), and a NaN in back is zeroed.
    X = np.array([0.0, 1.0, 2.0, 1.0, 3.0])
    back = np.array([1.0, 1.0, 0.0, 2.0, np.nan])
    tb = np.array([0.0, 0.0, 1.0, 4.0, 1.0])
    out = orl.rl_update_standard(X.copy(), back, tb)
    assert out[0] == 0.0 and np.isposinf(out[1]) and out[2] == 0.0 and out[3] == 0.5 and out[4] == 0.0
    out2 = orl.rl_update_standard(X.copy(), back)
    np.testing.assert_array_equal(out2, [0.0, 1.0, 0.0, 2.0, 0.0])


@pytest.mark.parametrize("N,s2", [(3, 7), (5, 11), (3, 13), (7, 29)])
def test_calc_ht_stamp_matches_adjoint_quadrant(N, s2):
    # calcHt_new (generate_psf.py:646-719) restated in the generator's own
    # layout, converted to the HMatrix layout, equals the product's
    # adjoint_quadrant with the reference's window truncation exactly; and
    # without the truncation the two agree wherever the reference keeps a
    # value.  The quadrant keeps the transpose symmetry the projector
    # assumes (its lower triangle is never read, py_light_field.cpp:1514-1521).
    from flfhip.psfmatrix import adjoint_quadrant
    rng = np.random.default_rng(N * 100 + s2)
    h = (N + 1) // 2
    Q = opsf.quadrant_of(opsf.expand_quadrant(rng.random((h, h, s2, s2)) + 0.1))
    ref = opsf.hcc_layout(opsf.calc_ht_stamp(opsf.mat_layout(Q)))
    np.testing.assert_array_equal(adjoint_quadrant(Q, reference_window=True), ref)
    exact = adjoint_quadrant(Q)
    for b in range(h):
        for a in range(h):
            kept = ref[b, a] != 0
            np.testing.assert_array_equal(exact[b, a][kept], ref[b, a][kept])
            assert kept[h - 1 - b:, h - 1 - a:].all()


def test_calc_ht_stamp_planted_flip_fails():
    # Dropping the 180-degree rotation of the stamped PSF (a plausible
    # mistake) must break agreement with the adjoint.
    from flfhip.psfmatrix import adjoint_quadrant
    rng = np.random.default_rng(4)
    N, s2 = 5, 11
    h = (N + 1) // 2
    Q = opsf.quadrant_of(opsf.expand_quadrant(rng.random((h, h, s2, s2)) + 0.1))
    Hpq = opsf.mat_layout(Q)
    bad = opsf.hcc_layout(opsf.calc_ht_stamp(Hpq[::-1, ::-1].copy()))
    assert np.max(np.abs(bad - adjoint_quadrant(Q, reference_window=True))) > 1e-2


def test_calc_ht_stamp_rejects_unsupported_side():
    with pytest.raises(ValueError):
        opsf.calc_ht_stamp(np.zeros((9, 9, 3, 3)))
