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
