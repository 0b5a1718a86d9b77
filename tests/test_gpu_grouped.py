"""Grouped launches (option grouped_launch, on by default): the small FFTs,
the multiply-accumulate and the column transforms of every active Fourier
grid group of a projection each go out as one dispatch.  Every workgroup
does the same arithmetic as in the group-by-group launches, so results are
bit-identical; the MAC profile shows one launch per projection instead of
one per grid."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _hm(N, sides, seed):
    rng = np.random.default_rng(seed)
    h = (N + 1) // 2
    from flfhip.psfmatrix import MemoryHMatrix
    return MemoryHMatrix([rng.random((h, h, s, s)).astype(np.float32) for s in sides],
                         [rng.random((h, h, s, s)).astype(np.float32) for s in sides])


def _pair():
    from flfhip.projector import Projector_hip
    pg, pp = Projector_hip(), Projector_hip()
    pg.set_option("grouped_launch", 1)
    pp.set_option("grouped_launch", 0)
    return pg, pp


@pytest.mark.parametrize("N,Y,X,sides", [(15, 150, 150, [15, 45, 75, 105]), (15, 120, 165, [15, 61, 91]),
                                         (13, 130, 130, [27, 53, 79])])
def test_grouped_bit_identical(N, Y, X, sides):
    hm = _hm(N, sides, seed=N + len(sides))
    pg, pp = _pair()
    T = 3
    g = torch.Generator(device="cuda")
    g.manual_seed(7)
    vol = torch.rand((hm.numZ, T, Y, X), generator=g, device="cuda")
    vol2 = torch.rand((hm.numZ, T, Y, X), generator=g, device="cuda")
    img = torch.rand((T, Y, X), generator=g, device="cuda") + 0.1
    Htf = torch.rand((hm.numZ, T, Y, X), generator=g, device="cuda") + 0.1
    for p in (pg, pp):
        p.prepare(hm, T, Y, X, frames_capacity=2 * T)
    ngrids = len(set(pg.fshape(z) for z in range(hm.numZ)))
    print("grids", sorted(set(pg.fshape(z) for z in range(hm.numZ))), flush=True)
    assert ngrids >= 3
    pg.profile(True)
    fg = pg.forward_device(hm, vol)
    _, launches, _ = pg.profile_read()
    pg.profile(False)
    assert launches == 1                       # one MAC dispatch for every grid
    fp = pp.forward_device(hm, vol)
    assert torch.equal(fg, fp)
    assert torch.equal(pg.backward_device(hm, img), pp.backward_device(hm, img))
    sub = [0, hm.numZ - 1]
    assert torch.equal(pg.forward_device(hm, vol, sub), pp.forward_device(hm, vol, sub))
    assert torch.equal(pg.backward_device(hm, img, planes=sub), pp.backward_device(hm, img, planes=sub))
    assert torch.equal(pg.forward_pair_device(hm, vol, vol2), pp.forward_pair_device(hm, vol, vol2))
    Xg, Xp = vol.clone(), vol.clone()
    pg.rl_update_device(hm, fg, Xg, Htf=Htf, mode=1)
    pp.rl_update_device(hm, fp, Xp, Htf=Htf, mode=1)
    assert torch.equal(Xg, Xp)
    # frames of uneven size inside the canvas
    dims = [(Y, X), (Y - 2 * N, X - N), (Y - N, X - 3 * N)]
    for p in (pg, pp):
        p.set_frame_dims(dims)
    assert torch.equal(pg.forward_device(hm, vol), pp.forward_device(hm, vol))
    assert torch.equal(pg.backward_device(hm, img), pp.backward_device(hm, img))


def test_grouped_graph_replay_matches_eager():
    """An RL iteration captured as a graph with grouped launches replays to
    the eager result (the argument tables exist before the capture)."""
    from flfhip.lfdeconv import RLGraph, deconv_rl_device
    hm = _hm(15, [15, 45, 105], seed=4)
    pg, _ = _pair()
    Y = X = 150
    g = torch.Generator(device="cuda")
    g.manual_seed(3)
    img = torch.rand((2, Y, X), generator=g, device="cuda") + 0.1
    pg.prepare(hm, 2, Y, X)
    Htf = pg.backward_device(hm, img)
    Xe = Htf.clone()
    deconv_rl_device(pg, hm, Htf, 3, Xe)
    Xg = Htf.clone()
    graph = RLGraph(pg, hm, Htf, Xg)
    graph.run(1)
    torch.cuda.synchronize()
    graph.run(2)
    torch.cuda.synchronize()
    assert torch.equal(Xg, Xe)
