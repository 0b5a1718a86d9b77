"""CPU pins for the network restatement in oracle/lensletnet.py."""
import numpy as np
import pytest
import torch

from oracle import lensletnet as onet


def _params(shapes, rng, scale=0.2):
    out = []
    for ci, co in shapes:
        out.append(rng.normal(0, scale / np.sqrt(9 * ci), co * 9 * ci))
        out.append(rng.normal(0, 0.05, co))
    return np.concatenate(out)


def test_unshuffle_matches_loop():
    rng = np.random.default_rng(0)
    N, Ly, Lx, T = 3, 4, 5, 2
    img = rng.random((T, Ly * N, Lx * N))
    u = onet.unshuffle(torch.tensor(img), N).numpy()
    ref = np.zeros((T, N * N, Ly, Lx))
    for t in range(T):
        for ly in range(Ly):
            for lx in range(Lx):
                for b in range(N):
                    for a in range(N):
                        ref[t, b * N + a, ly, lx] = img[t, ly * N + b, lx * N + a]
    assert np.array_equal(u, ref)


def test_weight_layout_is_tap_major():
    # one nonzero weight at tap (ky, kx) = (0, 2) moves the input by (-1, +1)
    N, nz, width, nb = 3, 1, 16, 0
    shapes = onet.layer_shapes(N, nz, width, nb)
    p = torch.zeros(sum(co * 9 * ci + co for ci, co in shapes), dtype=torch.float64)
    layers = onet.split_params(p, shapes)
    ci, co = shapes[0]
    flat = np.zeros(co * 9 * ci)
    flat[(0 * 9 + 2) * ci + 4] = 1.0        # co 0, tap 2, ci 4
    w = onet.split_params(torch.tensor(np.concatenate([flat, np.zeros(co)] + [
        np.zeros(c2 * 9 * c1 + c2) for c1, c2 in shapes[1:]])), shapes)[0][0]
    x = torch.zeros(1, ci, 5, 5, dtype=torch.float64)
    x[0, 4, 2, 3] = 1.0
    y = torch.nn.functional.conv2d(x, w, padding=1)
    # y[p] = W[tap] x[p + (ky-1, kx-1)] = x[p + (-1, +1)]  ->  nonzero at p = (3, 2)
    assert y[0, 0, 3, 2] == 1.0 and float(y.abs().sum()) == 1.0
    assert len(layers) == len(shapes)


def test_grad_matches_finite_differences():
    rng = np.random.default_rng(1)
    N, nz, width, nb = 3, 2, 16, 1
    shapes = onet.layer_shapes(N, nz, width, nb)
    p = _params(shapes, rng, scale=1.0)
    img = rng.random((2, 4 * N, 3 * N)) + 0.1
    tgt = rng.random((2, nz, 4, 3))
    L, g = onet.loss_and_grad(p, img, tgt, N, nz, width, nb)
    idx = rng.choice(p.size, 40, replace=False)
    # include a bias of every layer
    off = 0
    for ci, co in shapes:
        off += co * 9 * ci
        idx = np.append(idx, off)
        off += co
    eps = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        Lp, _ = onet.loss_and_grad(pp, img, tgt, N, nz, width, nb)
        Lm, _ = onet.loss_and_grad(pm, img, tgt, N, nz, width, nb)
        fd = (Lp - Lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_zero_loss_at_own_output():
    rng = np.random.default_rng(2)
    N, nz, width, nb = 3, 2, 16, 1
    shapes = onet.layer_shapes(N, nz, width, nb)
    p = _params(shapes, rng)
    img = rng.random((1, 2 * N, 2 * N)) + 0.1
    tgt = onet.guess(torch.tensor(p), torch.tensor(img), N, nz, width, nb).numpy()
    L, g = onet.loss_and_grad(p, img, tgt, N, nz, width, nb)
    assert L < 1e-28 and np.abs(g).max() < 1e-14


def test_planted_sign_error_is_caught():
    """A dropped residual term changes the output: the check above is not
    vacuous for the block structure."""
    rng = np.random.default_rng(3)
    N, nz, width, nb = 3, 2, 16, 1
    shapes = onet.layer_shapes(N, nz, width, nb)
    p = torch.tensor(_params(shapes, rng, scale=1.0))
    img = torch.tensor(rng.random((1, 2 * N, 2 * N)) + 0.1)
    o, s = onet.forward(p, img, N, nz, width, nb)
    layers = onet.split_params(p, shapes)
    F = torch.nn.functional
    x = onet.unshuffle(torch.sqrt(img / s[:, None, None]), N)
    h = F.relu(F.conv2d(x, layers[0][0], layers[0][1], padding=1))
    t = F.relu(F.conv2d(h, layers[1][0], layers[1][1], padding=1))
    h_wrong = F.conv2d(t, layers[2][0], layers[2][1], padding=1)          # residual dropped
    o_wrong = F.conv2d(h_wrong, layers[3][0], layers[3][1], padding=1)
    assert float((o - o_wrong).abs().max()) > 1e-3
