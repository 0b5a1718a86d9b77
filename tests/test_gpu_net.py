"""GPU checks of the learned initialiser's BF16 MFMA kernels (through the C ABI)
against the CPU restatement in oracle/lensletnet.py."""
import ctypes

import numpy as np
import pytest

from oracle import lensletnet as onet

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def lib():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from flfhip import _lib
    return _lib.load()


@pytest.mark.parametrize("M,N,K", [(77, 150, 48), (64, 128, 16), (300, 41, 1152), (33, 257, 2160)])
def test_mfma_gemm_layout(lib, M, N, K):
    """C = A B^T on v_mfma_f32_32x32x16_bf16: exact bf16 products summed in
    fp32, so 1e-5 relative against an FP64 product of the same bf16 values."""
    print("gemm", M, N, K, flush=True)
    g = torch.Generator().manual_seed(M * 7 + N)
    A = torch.randn(M, K, generator=g).to(torch.bfloat16)
    B = torch.randn(N, K, generator=g).to(torch.bfloat16)
    Ad, Bd = A.cuda(), B.cuda()
    C = torch.empty(M, N, dtype=torch.float32, device="cuda")
    rc = lib.flf_mfma_gemm_nt(ctypes.c_void_p(Ad.data_ptr()), ctypes.c_void_p(Bd.data_ptr()),
                              ctypes.c_void_p(C.data_ptr()), M, N, K,
                              ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0
    torch.cuda.synchronize()
    ref = A.double().numpy() @ B.double().numpy().T
    got = C.cpu().numpy()
    assert rel_l2(got, ref) < 1e-5
    assert np.max(np.abs(got - ref)) < 1e-4 * np.sqrt(K)


def test_mfma_gemm_rejects_bad_k(lib):
    assert lib.flf_mfma_gemm_nt(ctypes.c_void_p(1), ctypes.c_void_p(1), ctypes.c_void_p(1), 4, 4, 24, None) == -1


CASES = [
    # (Nnum, nz, width, nblocks, T, Ly, Lx)
    (5, 7, 32, 2, 2, 6, 7),
    (15, 41, 128, 1, 1, 9, 8),
]


def _make(case, seed):
    from flfhip.lensletnet import LensletNet
    N, nz, width, nb, T, Ly, Lx = case
    net = LensletNet(N, nz, width, nb, seed=seed)
    rng = np.random.default_rng(seed)
    p = net.init_params(seed)
    # larger residual/head weights than the initialiser so every layer matters
    p = p * 3.0 + rng.normal(0, 0.02, p.size).astype(np.float32)
    p = p.astype(np.float32)
    net.set_params(p)
    img = (rng.random((T, Ly * N, Lx * N)) * 100 + 5).astype(np.float32)
    tgt = (rng.random((T, nz, Ly, Lx)) * 2).astype(np.float32)
    return net, p, img, tgt


@pytest.mark.parametrize("case", CASES)
def test_net_forward_vs_oracle(case):
    """Inference: GPU guess against the CPU network with the same bf16
    storage points; tolerance 1e-2 relative L2 (fp32 accumulation order vs
    FP64 through bf16 re-rounding of ~2 nblocks + 2 activations)."""
    print("forward", case, flush=True)
    net, p, img, tgt = _make(case, 11)
    N, nz, width, nb, T, Ly, Lx = case
    got = net.forward(torch.tensor(img).cuda()).cpu().numpy()
    ref = onet.guess(torch.tensor(p, dtype=torch.float64), torch.tensor(img, dtype=torch.float64), N, nz, width, nb,
                     bf16=True).numpy()
    err = rel_l2(got, ref)
    print("rel-L2 %.3e" % err, flush=True)
    assert got.shape == (T, nz, Ly, Lx)
    assert err < 1e-2


@pytest.mark.parametrize("case", CASES)
def test_net_gradients_vs_oracle(case):
    """Training kernels: loss and every layer's weight and bias gradient from
    flf_net_train_step (lr = 0: no update) against CPU autograd of the same
    network.  Gradients pass through bf16 storage on the GPU (dO, dh, dt), so
    the bar is 5e-2 relative L2 per tensor and cosine >= 0.998 overall."""
    print("grads", case, flush=True)
    net, p, img, tgt = _make(case, 12)
    N, nz, width, nb, T, Ly, Lx = case
    loss = net.train_step(torch.tensor(img).cuda(), torch.tensor(tgt).cuda(), 0.0)
    g = net.get_grads()
    assert np.array_equal(net.get_params(), p)          # lr = 0 leaves the weights alone
    Lref, gref = onet.loss_and_grad(p, img, tgt, N, nz, width, nb, bf16=True)
    print("loss gpu %.6e cpu %.6e" % (loss, Lref), flush=True)
    assert abs(loss - Lref) <= 1e-2 * abs(Lref)
    off = 0
    for li, (ci, co) in enumerate(onet.layer_shapes(N, nz, width, nb)):
        nW = co * 9 * ci
        eW = rel_l2(g[off:off + nW], gref[off:off + nW])
        eb = rel_l2(g[off + nW:off + nW + co], gref[off + nW:off + nW + co])
        print("layer %d dW %.3e db %.3e" % (li, eW, eb), flush=True)
        assert eW < 5e-2 and eb < 5e-2, li
        off += nW + co
    cos = float(np.dot(g, gref) / (np.linalg.norm(g) * np.linalg.norm(gref)))
    assert cos > 0.998


def test_adam_first_step():
    """First Adam step: m^ = g, v^ = g^2, so w1 = w0 - lr g / (|g| + 1e-8)."""
    case = CASES[0]
    net, p, img, tgt = _make(case, 13)
    lr = 1e-3
    net.train_step(torch.tensor(img).cuda(), torch.tensor(tgt).cuda(), 0.0)
    g = net.get_grads().astype(np.float64)
    net.set_params(p)                                     # resets the Adam state
    net.train_step(torch.tensor(img).cuda(), torch.tensor(tgt).cuda(), lr)
    w1 = net.get_params().astype(np.float64)
    ref = p - lr * g / (np.abs(g) + 1e-8)
    assert np.max(np.abs(w1 - ref)) < 2e-6


def test_training_reduces_loss():
    """Fitting a fixed small batch: the loss falls below a fifth of its start."""
    from flfhip.lensletnet import LensletNet
    N, nz, width, nb = 5, 4, 32, 1
    net = LensletNet(N, nz, width, nb, seed=3)
    rng = np.random.default_rng(3)
    img = torch.tensor((rng.random((4, 8 * N, 8 * N)) * 50 + 1).astype(np.float32)).cuda()
    tgt = torch.tensor(rng.random((4, nz, 8, 8)).astype(np.float32) * 0.5 + 0.5).cuda()
    first = net.train_step(img, tgt, 3e-3)
    for _ in range(150):
        last = net.train_step(img, tgt, 3e-3)
    print("loss %.4e -> %.4e" % (first, last), flush=True)
    assert last < 0.2 * first


def test_upsample_exact(lib):
    from flfhip.lensletnet import LensletNet  # noqa: F401  (loads the library)
    T, nz, Ly, Lx, N = 2, 3, 4, 5, 3
    g = torch.arange(T * nz * Ly * Lx, dtype=torch.float32).reshape(T, nz, Ly, Lx).cuda()
    vol = torch.empty((nz, T, Ly * N, Lx * N), dtype=torch.float32, device="cuda")
    rc = lib.flf_net_upsample(ctypes.c_void_p(g.data_ptr()), T, nz, Ly, Lx, N, ctypes.c_void_p(vol.data_ptr()),
                              ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0
    ref = np.repeat(np.repeat(g.cpu().numpy(), N, axis=2), N, axis=3).transpose(1, 0, 2, 3)
    assert np.array_equal(vol.cpu().numpy(), ref)


def test_upsample_full_size_rows(lib):
    """Upsampling at the 1455^2, 41-plane size with two frames: more output
    rows (119310) than workgroups, so the row loop wraps."""
    from flfhip.lensletnet import LensletNet  # noqa: F401  (loads the library)
    T, nz, Ly, Lx, N = 2, 41, 97, 97, 15
    g = torch.rand((T, nz, Ly, Lx), dtype=torch.float32, device="cuda")
    vol = torch.empty((nz, T, Ly * N, Lx * N), dtype=torch.float32, device="cuda")
    rc = lib.flf_net_upsample(ctypes.c_void_p(g.data_ptr()), T, nz, Ly, Lx, N, ctypes.c_void_p(vol.data_ptr()),
                              ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0
    ref = g.repeat_interleave(N, dim=2).repeat_interleave(N, dim=3).permute(1, 0, 2, 3)
    assert torch.equal(vol, ref)
