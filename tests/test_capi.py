"""The C ABI (CPU): the library loads, exports every symbol include/flf_hip.h
declares, the bindings declare the same set, host-only planning queries give
the expected shapes and table invariants, and GPU entry points fail with a
status code (no crash) where no GPU is present."""
import ctypes
import os
import re

import numpy as np
import pytest

from flfhip import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "flf_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(flf_[a-z_0-9]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libflf_hip.so not built")
    return _lib.load()


def test_header_symbols_exported(lib):
    names = declared_functions()
    assert len(names) >= 14
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)


def test_version(lib):
    assert lib.flf_version() == 1


def test_fourier_len(lib):
    # 5-smooth reduced length: (1455+345-1)/15 -> 120; (1455+15-1)/15 = 98 -> 100.
    assert lib.flf_fourier_len(1455, 345, 15) == 1800
    assert lib.flf_fourier_len(1455, 15, 15) == 1500
    assert lib.flf_fourier_len(300, 135, 15) == 450
    assert lib.flf_fourier_len(450, 391, 15) == 900


@pytest.mark.parametrize("N,fY,fX", [(15, 1800, 1800), (15, 1620, 1620), (15, 1500, 1500), (15, 450, 450),
                                     (15, 225, 300), (15, 300, 225), (5, 30, 40), (3, 24, 24), (15, 225, 225),
                                     (7, 105, 84)])
def test_grid_tables(lib, N, fY, fX):
    st = (ctypes.c_longlong * 6)()
    assert lib.flf_grid_check(N, fY, fX, st) == 0
    chunks, items, max_res, nu, slots, max_items = list(st)
    h = (N + 1) // 2
    rfX = fX // 2 + 1
    if fY == fX:
        assert nu == h * (h + 1) // 2 and slots == 4
        assert max_res <= 8
    else:
        assert nu == h * h and slots == 2
        assert max_res <= 4
    assert max_items <= 256
    # Every half-plane bin is one slot of one item (checked inside); items
    # cover at least bins/slots of them.
    assert items * slots >= fY * rfX


def test_grid_check_rejects(lib):
    st = (ctypes.c_longlong * 6)()
    assert lib.flf_grid_check(14, 280, 280, st) == _lib.FLF_EINVAL      # even Nnum
    assert lib.flf_grid_check(15, 100, 100, st) == _lib.FLF_EINVAL      # not a multiple of Nnum


def test_create_without_gpu_fails_cleanly(lib):
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("GPU present")
    except Exception:
        pass
    ctx = ctypes.c_void_p()
    rc = lib.flf_create(0, ctypes.byref(ctx))
    assert rc != 0 and not ctx.value


def test_null_context_errors(lib):
    assert lib.flf_plan(None, 15, 10, 10, 1, 1, None) == _lib.FLF_EINVAL
    assert lib.flf_forward(None, None, None, None, 0, None) == _lib.FLF_EINVAL
    # context-free calls report through flf_last_error(NULL), per host thread
    assert lib.flf_rl_scale(None, None, None, 0, 1, 10, 1, None) == _lib.FLF_EINVAL
    assert b"bad nz" in lib.flf_last_error(None)


def test_host_buffers_refused(lib):
    # Host memory is never accepted as a kernel argument: the check runs on
    # the host before anything is queued (here without a GPU at all).
    a = np.zeros(64, np.float32)
    assert lib.flf_check_device_buffer(0, a.ctypes.data_as(ctypes.c_void_p), a.nbytes) == _lib.FLF_EINVAL
    assert b"device memory" in lib.flf_last_error(None) or b"host" in lib.flf_last_error(None)
    assert lib.flf_check_device_buffer(0, None, 4) == _lib.FLF_EINVAL
    assert b"null pointer" in lib.flf_last_error(None)
    p = a.ctypes.data_as(ctypes.c_void_p)
    # refused: FLF_EINVAL on a GPU box, FLF_EHIP here where no device exists
    assert lib.flf_rl_scale(p, p, p, 1, 1, 64, 1, None) in (_lib.FLF_EINVAL, _lib.FLF_EHIP)


def test_error_mapping():
    class FakeLib:
        def flf_last_error(self, ctx):
            return b"msg"
    _lib.check(FakeLib(), None, 0, "x")
    with pytest.raises(ValueError):
        _lib.check(FakeLib(), None, _lib.FLF_EINVAL, "x")
    with pytest.raises(MemoryError):
        _lib.check(FakeLib(), None, _lib.FLF_ENOMEM, "x")
    with pytest.raises(RuntimeError):
        _lib.check(FakeLib(), None, _lib.FLF_ESTATE, "x")
    with pytest.raises(_lib.FLFError):
        _lib.check(FakeLib(), None, _lib.FLF_EHIP, "x")


def test_product_psf_helpers_match_oracle():
    from flfhip import psfmatrix
    from oracle import psf as opsf
    rng = np.random.default_rng(0)
    Q = rng.random((4, 4, 9, 9)).astype(np.float32)
    np.testing.assert_array_equal(psfmatrix.full_psf(Q), opsf.expand_quadrant(Q))
    np.testing.assert_array_equal(psfmatrix.adjoint_quadrant(Q),
                                  opsf.quadrant_of(opsf.adjoint_psf(opsf.expand_quadrant(Q))))


def test_memmap_roundtrip(tmp_path):
    from flfhip import psfmatrix, synthetic
    hm = synthetic.make_psf_set(5, [-1, 0, 1], [9, 5, 9], seed=3)
    H = [hm.Hcc(z, False) for z in range(3)]
    Ht = [hm.Hcc(z, True) for z in range(3)]
    disk = psfmatrix.SaveMatrix(str(tmp_path / "psf"), H, Ht)
    again = psfmatrix.LoadMatrix(str(tmp_path / "psf.mat"))
    for z in range(3):
        np.testing.assert_array_equal(disk.Hcc(z, False), H[z])
        np.testing.assert_array_equal(again.Hcc(z, True), Ht[z])
        assert again.PSFShape(z) == H[z].shape[2:]
    assert again.Nnum == 5 and again.numZ == 3


def test_synthetic_psf_normalised_and_symmetric():
    from flfhip import synthetic, psfmatrix
    rng = np.random.default_rng(1)
    Q = synthetic.ring_quadrant(15, 61, 12.0, 1.5, rng)
    np.testing.assert_allclose(Q.sum(axis=(2, 3)), 1.0, rtol=1e-5)
    H = psfmatrix.full_psf(Q)
    # Physical symmetry: H[b][a] for every offset equals the ring centred on
    # that offset's lenslet centre (checked on the centre offset and a corner).
    np.testing.assert_allclose(H[7, 7], H[7, 7][::-1, :], atol=1e-7)
    np.testing.assert_allclose(H[0, 0], H[0, 0].T, atol=1e-7)


def test_option_entry_points_reject_null():
    lib = _lib.load()
    assert lib.flf_set_option(None, b"fwd_split", 1) == -1
    v = ctypes.c_int()
    assert lib.flf_get_option(None, b"fwd_split", ctypes.byref(v)) == -1


def _plan(lib, L, wide, order):
    r = (ctypes.c_int * 16)()
    n = ctypes.c_int()
    rc = lib.flf_radix_plan(L, wide, order, r, ctypes.byref(n))
    return rc, list(r[:n.value])


@pytest.mark.parametrize("L", [1800, 1620, 1500, 900, 120, 225, 66, 99, 1520, 1710, 1539, 65, 78, 56, 2048, 7])
def test_radix_plan_orders(lib, L):
    """Every pass order of a plan multiplies out to L, keeps an odd prime
    factor above 5 as the last pass, uses the radices its kernels hold
    (2..5 block-level; also 8 and 6 wave-level), and differs from the
    default order as documented: 1 reverses the 5-smooth passes, 2 takes no
    radix 8, 3 takes the odd radices first; block-level plans change only
    under order 1."""
    import math
    for wide in (0, 1):
        base = None
        for order in range(4):
            rc, r = _plan(lib, L, wide, order)
            assert rc == 0, (L, wide, order)
            assert math.prod(r) == L
            smooth = [x for x in r if x in (2, 3, 4, 5, 6, 8)]
            primes = [x for x in r if x not in (2, 3, 4, 5, 6, 8)]
            assert len(primes) <= 1 and r[len(smooth):] == primes          # the prime pass is last
            assert set(smooth) <= ({2, 3, 4, 5, 6, 8} if wide else {2, 3, 4, 5})
            if order == 0:
                base = smooth
            elif order == 1:
                assert smooth == base[::-1]
            elif not wide:
                assert smooth == base
            elif order == 2:
                assert 8 not in smooth
            else:
                odd = [x for x in smooth if x in (3, 5)]
                assert smooth[:len(odd)] == odd
    assert _plan(lib, L, 1, 4)[0] == _lib.FLF_EINVAL


def test_radix_plan_rejects_unsupported(lib):
    assert _plan(lib, 7 * 11, 1, 0)[0] == _lib.FLF_EINVAL       # two primes above 5
    assert _plan(lib, 37, 0, 0)[0] == _lib.FLF_EINVAL           # prime above 31
