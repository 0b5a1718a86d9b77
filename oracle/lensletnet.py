"""CPU restatement of the learned initialiser (lenslet-aware conv net).

TEST INFRASTRUCTURE ONLY (see the package header).  The reference has no
learned component, so this follows the network definition of the HIP library
(flf_net_* in libflf_hip.so) instead of a reference file; the projector and RL
that produce its training pairs are the ones restated in ``projector``/``rl``.

Network, all 3x3 convolutions with zero padding 1 at lenslet resolution:
    s   = mean(image)                       per frame
    x   = pixel_unshuffle(sqrt(max(image, 0) / s), N)   N*N channels, channel b*N + a
    h0  = relu(conv0(x))
    h_i = h_{i-1} + conv_b(relu(conv_a(h_{i-1})))      i = 1..nblocks
    o   = conv_head(h_n)
    y   = softplus(o) * s / nz              volume guess at lenslet resolution
    L   = mean over (frame, plane, lenslet) of (softplus(o) - target nz / s)^2
Parameters are one flat float32 vector, per layer W[Co][tap][Ci] (tap =
3 ky + kx, offset (ky - 1, kx - 1)) followed by b[Co].

``bf16=True`` rounds the quantities the GPU stores in bfloat16 (input x,
weights, h_i, t_i) at the same points, so the comparison tolerance is set by
the accumulation order and not by the storage format.

Pins (tests/test_oracle_net.py): autograd gradients against central finite
differences in FP64; ``unshuffle`` against an explicit loop; zero loss and
zero gradient at a target that equals the output.
"""
import numpy as np
import torch
import torch.nn.functional as F


def layer_shapes(nnum, nz, width, nblocks):
    return [(nnum * nnum, width)] + [(width, width)] * (2 * nblocks) + [(width, nz)]


def split_params(p, shapes):
    """flat vector -> list of (W [Co, Ci, 3, 3], b [Co]) torch tensors."""
    out, off = [], 0
    for ci, co in shapes:
        w = p[off:off + co * 9 * ci].reshape(co, 9, ci).permute(0, 2, 1).reshape(co, ci, 3, 3)
        off += co * 9 * ci
        b = p[off:off + co]
        off += co
        out.append((w, b))
    assert off == p.numel()
    return out


def unshuffle(img, nnum):
    """(T, Ly*N, Lx*N) -> (T, N*N, Ly, Lx), channel b*N + a."""
    T, Y, X = img.shape
    return F.pixel_unshuffle(img[:, None], nnum)


def _r(x, on):
    return x.to(torch.bfloat16).to(x.dtype) if on else x


def forward(p, img, nnum, nz, width, nblocks, bf16=False):
    """Returns (o, s): head pre-activation (T, nz, Ly, Lx) and frame means."""
    layers = split_params(p, layer_shapes(nnum, nz, width, nblocks))
    s = img.reshape(img.shape[0], -1).mean(dim=1).clamp_min(1e-12)
    x = _r(unshuffle(torch.sqrt(img.clamp_min(0) / s[:, None, None]), nnum), bf16)
    w, b = layers[0]
    h = _r(F.relu(F.conv2d(x, _r(w, bf16), b, padding=1)), bf16)
    for i in range(nblocks):
        wa, ba = layers[1 + 2 * i]
        wb, bb = layers[2 + 2 * i]
        t = _r(F.relu(F.conv2d(h, _r(wa, bf16), ba, padding=1)), bf16)
        h = _r(h + F.conv2d(t, _r(wb, bf16), bb, padding=1), bf16)
    w, b = layers[-1]
    o = F.conv2d(h, _r(w, bf16), b, padding=1)
    return o, s


def guess(p, img, nnum, nz, width, nblocks, bf16=False):
    o, s = forward(p, img, nnum, nz, width, nblocks, bf16)
    return F.softplus(o) * (s / nz)[:, None, None, None]


def loss(o, s, target):
    nz = o.shape[1]
    return ((F.softplus(o) - target * nz / s[:, None, None, None]) ** 2).mean()


def loss_and_grad(p_np, img_np, tgt_np, nnum, nz, width, nblocks, bf16=False, dtype=torch.float64):
    """Loss and d loss / d params (flat, same layout) by autograd on the CPU."""
    p = torch.tensor(np.asarray(p_np), dtype=dtype, requires_grad=True)
    img = torch.tensor(np.asarray(img_np), dtype=dtype)
    tgt = torch.tensor(np.asarray(tgt_np), dtype=dtype)
    o, s = forward(p, img, nnum, nz, width, nblocks, bf16)
    L = loss(o, s, tgt)
    L.backward()
    return float(L.detach()), p.grad.detach().numpy().copy()
