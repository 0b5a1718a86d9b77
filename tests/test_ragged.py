"""CPU pins for batches of frames of uneven size: the oracle shows that a
frame zero-padded into a larger canvas projects (forward and back) to the
same values as at its own size, which is what makes one plan per canvas
exact; and the bucket planner is checked against brute force."""
import itertools

import numpy as np
import pytest

from flfhip import synthetic
from flfhip.ragged import plan_buckets
from oracle import projector as oproj


@pytest.fixture(scope="module")
def psfs():
    N, zs = 15, [-2, 0, 3]
    hm = synthetic.make_psf_set(N, zs, synthetic.psf_sides(zs, 15, 45), seed=21)
    return ([hm.Hcc(z, False) for z in range(3)], [hm.Hcc(z, True) for z in range(3)])


@pytest.mark.parametrize("yv,xv", [(60, 75), (45, 105), (90, 60), (52, 67), (7, 16)])
def test_zero_padding_keeps_projections(psfs, yv, xv):
    Qf, Qb = psfs
    Y, X = 90, 105
    v = synthetic.blob_volume(3, yv, xv, 10, seed=3).astype(np.float64)
    vp = np.zeros((3, 1, Y, X))
    vp[:, :, :yv, :xv] = v
    a = oproj.forward_project(v, Qf)
    b = oproj.forward_project(vp, Qf)
    assert np.abs(a - b[:, :yv, :xv]).max() <= 1e-12 * np.abs(a).max()
    im = np.random.default_rng(1).random((1, yv, xv))
    ip = np.zeros((1, Y, X))
    ip[:, :yv, :xv] = im
    c = oproj.backward_project(im, Qb)
    d = oproj.backward_project(ip, Qb)
    assert np.abs(c - d[:, :, :yv, :xv]).max() <= 1e-12 * np.abs(c).max()


def test_padding_mistake_is_caught(psfs):
    """A frame placed one lenslet away from the corner projects differently
    (the lenslet phase matters): the check above is not vacuous."""
    Qf, _ = psfs
    v = synthetic.blob_volume(3, 60, 75, 10, seed=3).astype(np.float64)
    vp = np.zeros((3, 1, 90, 105))
    vp[:, :, 1:61, :75] = v
    a = oproj.forward_project(v, Qf)
    b = oproj.forward_project(vp, Qf)[:, 1:61, :75]
    assert np.abs(a - b).max() > 1e-3 * np.abs(a).max()


def _cost(groups, shapes, w):
    c = 0.0
    for g in groups:
        yc = max(shapes[i][0] for i in g)
        xc = max(shapes[i][1] for i in g)
        c += (len(g) + w) * yc * xc
    return c


def _set_partitions(items):
    if not items:
        yield []
        return
    first, rest = items[0], items[1:]
    for part in _set_partitions(rest):
        yield [[first]] + part
        for k in range(len(part)):
            yield part[:k] + [[first] + part[k]] + part[k + 1:]


@pytest.mark.parametrize("seed", range(6))
def test_bucket_plan_optimal_for_square_frames(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 8))
    shapes = [(int(s), int(s)) for s in 15 * rng.integers(20, 61, n)]
    for nb in (1, 2, 3):
        plan = plan_buckets(shapes, nb, spectrum_weight=4.0)
        got = _cost([idx for _, idx in plan], shapes, 4.0)
        best = min(_cost(p, shapes, 4.0) for p in _set_partitions(list(range(n))) if len(p) <= nb)
        assert got == pytest.approx(best, rel=1e-12)


def test_bucket_plan_covers_every_frame():
    rng = np.random.default_rng(9)
    shapes = [(int(a), int(b)) for a, b in zip(rng.integers(100, 900, 40), rng.integers(100, 900, 40))]
    for nb in (1, 2, 4, 40, 100):
        plan = plan_buckets(shapes, nb)
        assert 1 <= len(plan) <= min(nb, 40)
        seen = sorted(itertools.chain.from_iterable(idx for _, idx in plan))
        assert seen == list(range(40))
        for (yc, xc), idx in plan:
            assert idx == sorted(idx)
            assert all(shapes[i][0] <= yc and shapes[i][1] <= xc for i in idx)
            assert yc == max(shapes[i][0] for i in idx) and xc == max(shapes[i][1] for i in idx)
    assert plan_buckets([], 3) == []
