"""GPU parity of the HIP projector against the FP64 oracle (through the C ABI)."""
import numpy as np
import pytest

from flfhip import synthetic
from oracle import projector as oproj
from oracle import psf as opsf

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def _hm_random(N, s2_list, seed):
    """Random non-negative symmetric-by-construction quadrants (a quadrant is
    any array; the symmetry expansion defines the full set)."""
    rng = np.random.default_rng(seed)
    h = (N + 1) // 2
    H = [rng.random((h, h, s, s)).astype(np.float32) for s in s2_list]
    Ht = [rng.random((h, h, s, s)).astype(np.float32) for s in s2_list]
    from flfhip.psfmatrix import MemoryHMatrix
    return MemoryHMatrix(H, Ht)


# The reference self-test shapes (projector.py:1178-1179): square and non-square.
@pytest.mark.parametrize("shape", [(2, 150, 150), (2, 150, 300), (2, 300, 150)])
@pytest.mark.parametrize("backward", [False, True])
def test_selftest_shapes(projector, shape, backward):
    hm = _hm_random(15, [61], seed=0)
    rng = np.random.default_rng(0)
    x = rng.random(shape).astype(np.float32)
    print("case", shape, backward, flush=True)
    if backward:
        got = projector.BackwardProjectACC(hm, x, [0])[0]
        ref = oproj.project_special(x, hm.Hcc(0, True))
    else:
        got = projector.ForwardProjectACC(hm, x[None], [0])
        ref = oproj.project_special(x, hm.Hcc(0, False))
    err = rel_l2(got, ref)
    mx = float(np.max(np.abs(got - ref)))
    print("rel-L2 %.3e max-abs %.3e" % (err, mx), flush=True)
    assert err < 1e-5
    assert mx < 1e-4 * max(1.0, float(np.max(np.abs(ref))))


@pytest.mark.parametrize("N,Y,X,s2", [(5, 20, 25, 7), (15, 45, 60, 15), (15, 105, 105, 31), (3, 12, 12, 5)])
def test_small_direct(projector, N, Y, X, s2):
    hm = _hm_random(N, [s2], seed=N + Y)
    rng = np.random.default_rng(1)
    x = rng.random((1, 1, Y, X)).astype(np.float32)
    got = projector.ForwardProjectACC(hm, x, [0])
    ref = oproj.project_direct(x[0], opsf.expand_quadrant(hm.Hcc(0, False).astype(np.float64)))
    assert rel_l2(got, ref) < 1e-5


# Several frames per call: the multiply-accumulate kernel serves two frames per
# workgroup from one read of the spectra; T = 3 leaves a ragged last block.
@pytest.mark.parametrize("T,Y,X", [(3, 150, 150), (3, 150, 300), (2, 105, 120)])
def test_multiframe(projector, T, Y, X):
    hm = _hm_random(15, [31, 45], seed=T + Y)
    rng = np.random.default_rng(5)
    vol = rng.random((2, T, Y, X)).astype(np.float32)
    img = rng.random((T, Y, X)).astype(np.float32)
    print("multiframe", T, Y, X, flush=True)
    fwd = projector.ForwardProjectACC(hm, vol, [0, 1])
    bwd = projector.BackwardProjectACC(hm, img, [0, 1])
    for t in range(T):
        ref_f = sum(oproj.project_special(vol[z, t][None], hm.Hcc(z, False))[0] for z in range(2))
        assert rel_l2(fwd[t], ref_f) < 1e-5, t
        for z in range(2):
            ref_b = oproj.project_special(img[t][None], hm.Hcc(z, True))[0]
            assert rel_l2(bwd[z, t], ref_b) < 1e-5, (t, z)


def test_frame_capacity(projector):
    """One plan with room for 3 frames serves calls with 3, 1 and 2 frames
    (flf_set_frames); every result matches the oracle."""
    hm = _hm_random(15, [31, 45], seed=9)
    rng = np.random.default_rng(9)
    Y, X = 120, 135
    projector.prepare(hm, 1, Y, X, frames_capacity=3)
    for T in (3, 1, 2, 1):
        vol = rng.random((2, T, Y, X)).astype(np.float32)
        img = rng.random((T, Y, X)).astype(np.float32)
        print("frames", T, flush=True)
        fwd = projector.ForwardProjectACC(hm, vol, [0, 1])
        bwd = projector.BackwardProjectACC(hm, img, [0, 1])
        for t in range(T):
            ref_f = sum(oproj.project_special(vol[z, t][None], hm.Hcc(z, False))[0] for z in range(2))
            assert rel_l2(fwd[t], ref_f) < 1e-5, (T, t)
            ref_b = oproj.project_special(img[t][None], hm.Hcc(1, True))[0]
            assert rel_l2(bwd[1, t], ref_b) < 1e-5, (T, t)


@pytest.mark.parametrize("T,Y,X,ranges", [(3, 150, 150, 1), (8, 150, 150, 2), (10, 105, 120, 3), (2, 150, 300, 1),
                                          (9, 90, 150, 2)])
def test_multiframe_matrix_core_kernel(T, Y, X, ranges):
    """The matrix-core MAC (k_mac_tc, 8 frames per workgroup; option
    mac_mfma, also FLF_MAC_MFMA=1 when a projector is created) against the
    oracle: partial and full frame blocks (2, 3, 8, 9, 10 frames), square
    and non-square grids, 1 to 3 forward plane ranges and 2 planes per
    back-projection workgroup."""
    from flfhip.projector import Projector_hip
    proj = Projector_hip()
    proj.set_option("mac_mfma", 1)
    proj.set_option("fwd_split", ranges)
    proj.set_option("bwd_planes", 2)
    hm = _hm_random(15, [31, 45, 61], seed=T + Y + 1)
    rng = np.random.default_rng(7)
    vol = rng.random((3, T, Y, X)).astype(np.float32)
    img = rng.random((T, Y, X)).astype(np.float32)
    print("matrix-core multiframe", T, Y, X, ranges, flush=True)
    fwd = proj.ForwardProjectACC(hm, vol, [0, 1, 2])
    bwd = proj.BackwardProjectACC(hm, img, [0, 1, 2])
    for t in range(T):
        ref_f = sum(oproj.project_special(vol[z, t][None], hm.Hcc(z, False))[0] for z in range(3))
        assert rel_l2(fwd[t], ref_f) < 1e-5, t
        for z in range(3):
            ref_b = oproj.project_special(img[t][None], hm.Hcc(z, True))[0]
            assert rel_l2(bwd[z, t], ref_b) < 1e-5, (t, z)


def test_matrix_core_full_size_matches_vector_kernel():
    """At the 1455^2, 41-plane geometry with 8 and 10 frames, the
    matrix-core MAC's forward and backward projections equal the vector
    kernel's (FP32, other summation order) to 1e-6 relative L2 per frame."""
    import torch
    from flfhip import workloads
    from flfhip.projector import Projector_hip
    hm, zs, sides = workloads.psf_set("rl1455")
    Y = X = 1455
    pv, pm = Projector_hip(), Projector_hip()
    pm.set_option("mac_mfma", 1)
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    for T in (8, 10):
        vol = torch.rand((41, T, Y, X), generator=g, device="cuda")
        img = torch.rand((T, Y, X), generator=g, device="cuda")
        for p in (pv, pm):
            p.prepare(hm, T, Y, X)
        fv, fm = pv.forward_device(hm, vol), pm.forward_device(hm, vol)
        bv, bm = pv.backward_device(hm, img), pm.backward_device(hm, img)
        for t in range(T):
            ef = float(torch.linalg.vector_norm((fm[t] - fv[t]).double()) / torch.linalg.vector_norm(fv[t].double()))
            eb = float(torch.linalg.vector_norm((bm[:, t] - bv[:, t]).double()) /
                       torch.linalg.vector_norm(bv[:, t].double()))
            assert ef < 1e-6 and eb < 1e-6, (T, t, ef, eb)
        print("T %d: matrix-core vs vector kernel within 1e-6 on every frame" % T, flush=True)
        del vol, img, fv, fm, bv, bm


def test_autotune_and_options_keep_parity(projector):
    """The autotuner's choices, and every single option value, leave the
    projections within the parity bar."""
    from flfhip import autotune
    hm = _hm_random(15, [31, 45, 61], seed=31)
    rng = np.random.default_rng(8)
    T, Y, X = 2, 120, 150
    vol = rng.random((3, T, Y, X)).astype(np.float32)
    ref = [sum(oproj.project_special(vol[z, t][None], hm.Hcc(z, False))[0] for z in range(3)) for t in range(T)]
    saved = {k: projector.get_option(k) for k in ("fwd_split", "mac_frames", "ifft_col_batch", "bwd_planes")}
    chosen, timings = autotune.tune(projector, hm, T, Y, X, reps=1)
    print("autotune", chosen, flush=True)
    assert set(chosen) >= {"fwd_split", "ifft_col_batch", "mac_frames", "bwd_planes"}
    settings = [dict(chosen)] + [{"fwd_split": v} for v in (1, 4)] + [{"ifft_col_batch": 2}, {"mac_frames": 1}]
    for opts in settings:
        for k, v in opts.items():
            projector.set_option(k, v)
        fwd = projector.ForwardProjectACC(hm, vol, [0, 1, 2])
        for t in range(T):
            assert rel_l2(fwd[t], ref[t]) < 1e-5, (opts, t)
    for k, v in saved.items():
        if k == "ifft_col_batch" and v == 0:
            continue                     # read before any plan: every new plan sets its own
        projector.set_option(k, v)
    with pytest.raises(Exception):
        projector.set_option("no_such_option", 1)


def test_multiframe_full_size_equals_single_frames(projector):
    """At the full 1455 x 1455, 41-plane geometry (three Fourier grids, PSFs
    up to 345 px) a two-frame forward and backward projection equal two
    one-frame projections for every frames-per-workgroup setting and plane
    split: frames must not leak into each other's tables or accumulators."""
    import torch
    from flfhip import workloads
    hm, zs, sides = workloads.psf_set("rl1455")
    Y = X = 1455
    nz = len(zs)
    projector.prepare(hm, 1, Y, X, frames_capacity=2)
    g = torch.Generator(device=projector.device)
    g.manual_seed(3)
    vol = torch.rand((nz, 2, Y, X), generator=g, device=projector.device)
    img = torch.rand((2, Y, X), generator=g, device=projector.device)
    bwd0 = projector.get_option("bwd_planes")
    try:
        projector.set_option("mac_frames", 1)
        projector.set_option("fwd_split", 1)
        projector.set_option("bwd_planes", 1)
        ref_f = [projector.forward_device(hm, vol[:, t:t + 1].contiguous())[0].clone() for t in range(2)]
        ref_b = [projector.backward_device(hm, img[t:t + 1].contiguous())[:, 0].clone() for t in range(2)]
        for fb in (2, 3, 4):
            for sp in (1, 3):
                projector.set_option("mac_frames", fb)
                projector.set_option("fwd_split", sp)
                projector.set_option("bwd_planes", sp)
                two_f = projector.forward_device(hm, vol)
                two_b = projector.backward_device(hm, img)
                torch.cuda.synchronize()
                for t in range(2):
                    ef = float((two_f[t] - ref_f[t]).norm() / ref_f[t].norm())
                    eb = float((two_b[:, t] - ref_b[t]).norm() / ref_b[t].norm())
                    print("fb %d split %d frame %d: forward %.2e backward %.2e" % (fb, sp, t, ef, eb), flush=True)
                    assert ef < 1e-6 and eb < 1e-6
    finally:
        projector.set_option("mac_frames", 2)
        projector.set_option("fwd_split", 0)
        projector.set_option("bwd_planes", bwd0)


def test_one_item_per_thread_matches_lane_pairs(projector):
    """The MAC with one item per thread (FLF_ITEMS_PER_THREAD=1, read when a
    context is created) and the default lane pairs give the same projections
    for one and three frames."""
    import os
    import torch
    from flfhip import synthetic
    from flfhip.projector import Projector_hip
    N, zs = 15, [-3, 0, 2, 5]
    hm = synthetic.make_psf_set(N, zs, synthetic.psf_sides(zs, 15, 61), seed=8)
    old = os.environ.get("FLF_ITEMS_PER_THREAD")
    os.environ["FLF_ITEMS_PER_THREAD"] = "1"
    try:
        p1 = Projector_hip()
    finally:
        if old is None:
            del os.environ["FLF_ITEMS_PER_THREAD"]
        else:
            os.environ["FLF_ITEMS_PER_THREAD"] = old
    g = torch.Generator(device=projector.device)
    g.manual_seed(4)
    for T in (1, 3):
        vol = torch.rand((len(zs), T, 150, 165), generator=g, device=projector.device)
        img = torch.rand((T, 150, 165), generator=g, device=projector.device)
        a_f, a_b = projector.forward_device(hm, vol), projector.backward_device(hm, img)
        b_f, b_b = p1.forward_device(hm, vol), p1.backward_device(hm, img)
        torch.cuda.synchronize()
        ef = float((a_f - b_f).norm() / b_f.norm())
        eb = float((a_b - b_b).norm() / b_b.norm())
        print("T=%d lane pairs vs one item per thread: forward %.2e backward %.2e" % (T, ef, eb), flush=True)
        assert ef < 1e-6 and eb < 1e-6
    del p1


@pytest.mark.parametrize("grouped", [1, 0])
def test_bwd_frame_group_bit_identical(grouped):
    """The back projection with a chunk's frame blocks dispatched in
    super-blocks of 1, 2, 3, 5 or 64 (option bwd_frame_group; 5 frame blocks
    of 2 frames here, so 2 and 3 leave a shorter last super-block, 1 is the
    frame-block-slowest order and 64 clamps to all) equals the default order
    bit for bit: only the dispatch order changes."""
    import torch
    from flfhip.projector import Projector_hip
    hm = synthetic.make_psf_set(15, [0, 1, 2], [15, 45, 61], seed=71)
    T, Y, X = 9, 150, 165
    p = Projector_hip()
    p.prepare(hm, T, Y, X)
    p.set_option("grouped_launch", grouped)
    p.set_option("mac_frames", 2)
    g = torch.Generator(device="cuda")
    g.manual_seed(71 + grouped)
    img = torch.rand((T, Y, X), generator=g, device="cuda")
    ref = p.backward_device(hm, img).clone()
    for fg in (1, 2, 3, 5, 64):
        p.set_option("bwd_frame_group", fg)
        assert p.get_option("bwd_frame_group") == fg
        got = p.backward_device(hm, img)
        torch.cuda.synchronize()
        print("grouped %d bwd_frame_group %d: identical %s" % (grouped, fg, bool(torch.equal(got, ref))), flush=True)
        assert torch.equal(got, ref)
    with pytest.raises(ValueError):
        p.set_option("bwd_frame_group", 65)
    del p


@pytest.mark.parametrize("grouped", [1, 0])
@pytest.mark.parametrize("N,Y,X,sides", [(15, 150, 165, [15, 45, 61]), (13, 130, 143, [27, 53])])
def test_lane_quads_match_lane_pairs(grouped, N, Y, X, sides):
    """Four items per lane group (mac_ipt 4: every P/M table read serves four
    items and the quad adds its partial sums in two shuffle steps) and the
    lane pairs give the same projections for one to four frames per
    workgroup, in grouped and per-group launches; the one-frame case also
    against the FP64 oracle."""
    import torch
    from flfhip import synthetic
    from flfhip.projector import Projector_hip
    zs = list(range(len(sides)))
    hm = synthetic.make_psf_set(N, zs, sides, seed=N + 40)
    p2, p4 = Projector_hip(), Projector_hip()
    p2.set_option("mac_ipt", 2)
    p4.set_option("mac_ipt", 4)
    T = 3
    for p in (p2, p4):
        p.prepare(hm, T, Y, X)
        p.set_option("grouped_launch", grouped)
    assert p4.get_option("mac_ipt") == 4
    g = torch.Generator(device="cuda")
    g.manual_seed(N + grouped)
    vol = torch.rand((len(zs), T, Y, X), generator=g, device="cuda")
    img = torch.rand((T, Y, X), generator=g, device="cuda")
    ref_f, ref_b = p2.forward_device(hm, vol).clone(), p2.backward_device(hm, img).clone()
    for fb in (1, 2, 3, 4):
        p4.set_option("mac_frames", fb)
        got_f, got_b = p4.forward_device(hm, vol), p4.backward_device(hm, img)
        torch.cuda.synchronize()
        ef = float((got_f - ref_f).norm() / ref_f.norm())
        eb = float((got_b - ref_b).norm() / ref_b.norm())
        print("grouped %d N %d fb %d: lane quads vs pairs forward %.2e backward %.2e" % (grouped, N, fb, ef, eb),
              flush=True)
        assert ef < 1e-6 and eb < 1e-6
    x1 = vol[:, :1].cpu().numpy()
    got1 = p4.ForwardProjectACC(hm, x1, zs)
    ref1 = sum(oproj.project_special(x1[z], hm.Hcc(z, False)) for z in range(len(zs)))
    e1 = rel_l2(got1, ref1)
    print("lane quads one frame vs oracle: %.2e" % e1, flush=True)
    assert e1 < 1e-5
    del p2, p4


def test_forward_pair_equals_concatenated_frames(projector):
    """flf_forward_pair (frames from two buffers) gives the same image as
    flf_forward on the two volumes concatenated along frames, bit for bit;
    flf_rl_ratio equals flf_rl_scale on X = 1."""
    import torch
    from flfhip import synthetic
    N, zs = 15, [-2, 0, 3]
    hm = synthetic.make_psf_set(N, zs, synthetic.psf_sides(zs, 15, 45), seed=12)
    g = torch.Generator(device=projector.device)
    g.manual_seed(6)
    for S in (1, 2):
        A = torch.rand((3, S, 120, 135), generator=g, device=projector.device)
        B = torch.rand((3, S, 120, 135), generator=g, device=projector.device)
        projector.prepare(hm, 2 * S, 120, 135)
        ref = projector.forward_device(hm, torch.cat([A, B], dim=1))
        got = projector.forward_pair_device(hm, A, B)
        torch.cuda.synchronize()
        print("S=%d pair vs concatenated: max abs diff %.3g" % (S, float((got - ref).abs().max())), flush=True)
        assert torch.equal(got, ref)
    Htf = torch.rand((3, 1, 64, 64), generator=g, device=projector.device)
    v = torch.rand((3, 1, 64, 64), generator=g, device=projector.device)
    v[0, 0, 0, :4] = 0.0                                  # 0/0 and x/0 entries
    Htf[0, 0, 0, :2] = 0.0
    X1 = torch.ones_like(Htf)
    projector.rl_scale_device(X1, Htf, v)
    X2 = torch.empty_like(Htf)
    projector.rl_ratio_device(X2, Htf, v)
    torch.cuda.synchronize()
    assert torch.equal(X1, X2)
    assert float(X2[0, 0, 0, 0]) == 0.0                   # NaN -> 0
    assert np.isposinf(float(X2[0, 0, 0, 2]))             # x/0 -> inf


@pytest.mark.parametrize("planes", [None, [0, 2, 3]])
def test_backward_planes_per_workgroup(projector, planes):
    """Back projection with 1..8 planes per MAC workgroup (bwd_planes): every
    setting gives bit-identical planes (the per-plane arithmetic does not
    depend on the grouping) within 1e-5 of the oracle; planes left out of
    the call stay zero."""
    import torch
    hm = _hm_random(15, [15, 31, 45, 31, 61], seed=41)
    rng = np.random.default_rng(9)
    T, Y, X = 2, 105, 120
    img = rng.random((T, Y, X)).astype(np.float32)
    sel = list(range(5)) if planes is None else planes
    ref = {z: np.stack([oproj.project_special(img[t][None], hm.Hcc(z, True))[0] for t in range(T)]) for z in sel}
    saved = projector.get_option("bwd_planes")
    base = None
    try:
        for bp in (1, 2, 3, 4, 8):
            projector.set_option("bwd_planes", bp)
            vol = projector.backward_device(hm, projector.asnative(img), planes=planes)
            torch.cuda.synchronize()
            v = vol.cpu().numpy()
            if base is None:
                base = v
            assert np.array_equal(v, base), bp
            for z in range(5):
                if z in sel:
                    assert rel_l2(v[z], ref[z]) < 1e-5, (bp, z)
                else:
                    assert not v[z].any(), (bp, z)
    finally:
        projector.set_option("bwd_planes", saved)


@pytest.mark.parametrize("side_stream", [False, True])
def test_group_streams_bit_identical(projector, side_stream):
    """Fourier-grid groups run side by side on streams of their own
    (group_streams = 1) give the same bits as one stream, for forward,
    backward, two-source forward, plane subsets (one active group) and fused
    RL iterations, called from the default or a created stream, with new
    inputs on every call (a missing join would read the previous call's
    group results); the forward stays within 1e-5 of the oracle."""
    import torch
    from flfhip.lfdeconv import deconv_rl_device
    sides = [15, 31, 45, 61, 91]
    hm = _hm_random(15, sides, seed=43)
    T, Y, X = 2, 105, 120
    projector.prepare(hm, T, Y, X)
    assert len({projector.fshape(z) for z in range(len(sides))}) >= 2
    saved = projector.get_option("group_streams")
    g = torch.Generator(device=projector.device)
    g.manual_seed(17)
    vols = [torch.rand((5, T, Y, X), generator=g, device=projector.device) for _ in range(3)]
    imgs = [torch.rand((T, Y, X), generator=g, device=projector.device) + 0.1 for _ in range(3)]
    stream = torch.cuda.Stream(device=projector.device) if side_stream else torch.cuda.current_stream()

    def run():
        out = []
        with torch.cuda.stream(stream):
            for v, im in zip(vols, imgs):
                out.append(projector.forward_device(hm, v).clone())
                out.append(projector.backward_device(hm, im).clone())
                out.append(projector.backward_device(hm, im, planes=[1, 3]).clone())
                out.append(projector.forward_device(hm, v, planes=[0]).clone())
                out.append(projector.forward_pair_device(hm, v[:, :1].contiguous(), v[:, 1:].contiguous()).clone())
            Htf = projector.backward_device(hm, imgs[0])
            Xr = torch.ones_like(Htf)
            deconv_rl_device(projector, hm, Htf, 3, Xr, im=imgs[0])
            out.append(Xr)
        stream.synchronize()
        return out

    try:
        projector.set_option("group_streams", 0)
        one = run()
        projector.set_option("group_streams", 1)
        many = run()
    finally:
        projector.set_option("group_streams", saved)
    for i, (a, b) in enumerate(zip(one, many)):
        assert torch.equal(a, b), i
    v0 = vols[0].cpu().numpy().astype(np.float64)
    ref = oproj.forward_project(v0, [hm.Hcc(z, False) for z in range(len(sides))])
    for t in range(T):
        assert rel_l2(many[0][t].cpu().numpy(), ref[t]) < 1e-5


# Lenslet pitches with an odd prime factor above 5: the full-grid transforms
# end with a radix-P pass.  The reference self-test shapes (projector.py:
# 1178-1179) at Nnum = 19, both directions; 77 px PSFs give the even grid
# 228 = 19 x 12, 21 px PSFs the odd grid 171 = 19 x 9 (full-length c2r rows).
@pytest.mark.parametrize("shape", [(2, 150, 150), (2, 150, 300), (2, 300, 150)])
@pytest.mark.parametrize("s2", [77, 21])
@pytest.mark.parametrize("backward", [False, True])
def test_selftest_shapes_nnum19(projector, shape, s2, backward):
    hm = _hm_random(19, [s2], seed=s2)
    rng = np.random.default_rng(19)
    x = rng.random(shape).astype(np.float32)
    print("Nnum 19 case", shape, s2, backward, flush=True)
    if backward:
        got = projector.BackwardProjectACC(hm, x, [0])[0]
        ref = oproj.project_special(x, hm.Hcc(0, True))
    else:
        got = projector.ForwardProjectACC(hm, x[None], [0])
        ref = oproj.project_special(x, hm.Hcc(0, False))
    print("grid", projector.fshape(0), "rel-L2 %.3e" % rel_l2(got, ref), flush=True)
    assert rel_l2(got, ref) < 1e-5


@pytest.mark.parametrize("N,Y,X,s2", [(7, 49, 56, 15), (13, 65, 78, 27), (11, 60, 70, 23), (21, 84, 105, 43)])
def test_prime_pitch_small_direct(projector, N, Y, X, s2):
    """Pitches 7, 11, 13 and 21 (= 3 x 7) against the direct per-offset
    convolution, several planes of two grids, both directions."""
    hm = _hm_random(N, [s2, s2 + 2 * N], seed=N)
    rng = np.random.default_rng(N)
    x = rng.random((2, 1, Y, X)).astype(np.float32)
    got = projector.ForwardProjectACC(hm, x, [0, 1])
    ref = sum(oproj.project_direct(x[z], opsf.expand_quadrant(hm.Hcc(z, False).astype(np.float64))) for z in range(2))
    print("N %d grids %s %s forward rel-L2 %.3e" % (N, projector.fshape(0), projector.fshape(1), rel_l2(got, ref)),
          flush=True)
    assert rel_l2(got, ref) < 1e-5
    back = projector.BackwardProjectACC(hm, got, [0, 1])
    for z in range(2):
        refb = oproj.project_direct(got, opsf.expand_quadrant(hm.Hcc(z, True).astype(np.float64)))
        assert rel_l2(back[z], refb) < 1e-5


@pytest.mark.parametrize("order", [0, 1, 2, 3])
@pytest.mark.parametrize("N,Y,X,s2s", [(15, 90, 150, [31, 61]), (13, 65, 78, [27, 53])])
def test_plan_options_radix_order_and_ipt(projector, order, N, Y, X, s2s):
    """Every FFT pass order (radix_order 0..3) with one, two and four items
    per MAC lane group (mac_ipt), on grids of lengths 120..225 (radices 8, 6, 5, 3,
    2) and of Nnum 13 (prime pass last): forward and backward projection of
    two planes against the FP64 oracle.  Own projectors: the plan options
    stay off the session's projector."""
    from flfhip.projector import Projector_hip
    hm = _hm_random(N, s2s, seed=N + order)
    rng = np.random.default_rng(order)
    x = rng.random((2, 1, Y, X)).astype(np.float32)
    ref = sum(oproj.project_special(x[z], hm.Hcc(z, False)) for z in range(2))
    for ipt in (1, 2, 4):
        p = Projector_hip()
        p.set_option("radix_order", order)
        p.set_option("mac_ipt", ipt)
        got = p.ForwardProjectACC(hm, x, [0, 1])
        assert p.get_option("radix_order") == order and p.get_option("mac_ipt") == ipt
        e = rel_l2(got, ref)
        back = p.BackwardProjectACC(hm, got, [0, 1])
        eb = max(rel_l2(back[z], oproj.project_special(got, hm.Hcc(z, True))) for z in range(2))
        print("radix_order %d ipt %d grids %s %s: forward %.2e backward %.2e"
              % (order, ipt, p.fshape(0), p.fshape(1), e, eb), flush=True)
        assert e < 1e-5 and eb < 1e-5
        del p


def test_nils_geometry_plane(projector):
    """The reference benchmark's geometry (Nnum = 19, 1273 x 1463 px;
    benchmarking.ipynb cell 0) with a 229 px ring PSF (grid 1520 x 1710):
    forward and backward projection of one plane against the FP64 oracle."""
    from flfhip import synthetic
    rng = np.random.default_rng(22)
    Q = synthetic.ring_quadrant(19, 229, 60.0, 2.0, rng)
    from flfhip.psfmatrix import MemoryHMatrix, adjoint_quadrant
    hm = MemoryHMatrix([Q], [adjoint_quadrant(Q)])
    x = synthetic.blob_volume(1, 1273, 1463, 300, seed=9)
    got = projector.ForwardProjectACC(hm, x, [0])
    assert projector.fshape(0) == (1520, 1710)
    ref = oproj.project_special(x[0], Q)
    e = rel_l2(got, ref)
    print("Nils geometry forward rel-L2 %.3e" % e, flush=True)
    assert e < 1e-5
    back = projector.BackwardProjectACC(hm, got, [0])[0]
    refb = oproj.project_special(got, hm.Hcc(0, True))
    eb = rel_l2(back, refb)
    print("Nils geometry backward rel-L2 %.3e" % eb, flush=True)
    assert eb < 1e-5


def test_pitch_above_31_rejected(projector):
    """Nnum above 31 is refused at plan time with a clear error, before any
    kernel runs."""
    hm = _hm_random(33, [67], seed=33)
    with pytest.raises(ValueError, match="Nnum"):          # FLF_EINVAL
        projector.prepare(hm, 1, 4 * 33, 4 * 33)


def test_empty_plane_list(projector):
    """planes = []: BackwardProjectACC returns an all-zero volume, like the
    reference (projector.py:807); ForwardProjectACC refuses, where the
    reference fails on its empty plane sum (projector.py:854, 895); the
    device-level forward returns the empty sum, a zero image."""
    import torch
    hm = _hm_random(15, [31, 45], seed=51)
    rng = np.random.default_rng(3)
    img = rng.random((1, 90, 105)).astype(np.float32) + 0.1
    vol = rng.random((2, 1, 90, 105)).astype(np.float32)
    back = projector.BackwardProjectACC(hm, img, [])
    assert back.shape == (2, 1, 90, 105) and not back.any()
    with pytest.raises(ValueError):
        projector.ForwardProjectACC(hm, vol, [])
    out = projector.forward_device(hm, projector.asnative(vol) + 1.0, planes=[])
    torch.cuda.synchronize()
    assert out.shape == (1, 90, 105) and float(out.abs().max()) == 0.0


def test_largest_grid_adjoint_and_linearity(projector):
    """The largest Fourier grid of the register-held transforms (1920 = 15 x
    128: a 1575 px image with a 345 px PSF): back projection through Ht =
    adjoint_quadrant(H) is the adjoint of forward projection, <H x, y> =
    <x, Ht y>, and forward projection is linear; an image that needs a
    reduced grid above 256 points is refused at plan time.  The PSFs are rings (radially symmetric, as
    physical light-field PSFs are): only then does the adjoint set keep the
    8-fold symmetry that quadrant storage of Ht assumes (for an arbitrary
    quadrant it does not, in the reference's storage as here)."""
    import torch
    from flfhip import synthetic
    from flfhip.psfmatrix import MemoryHMatrix, adjoint_quadrant
    rng = np.random.default_rng(61)
    Q = synthetic.ring_quadrant(15, 345, 100.0, 1.0, rng).astype(np.float32)
    hm = MemoryHMatrix([Q], [adjoint_quadrant(Q)])
    Y = X = 1575
    projector.prepare(hm, 1, Y, X)
    assert projector.fshape(0) == (1920, 1920)
    g = torch.Generator(device=projector.device)
    g.manual_seed(9)
    x = torch.rand((1, 1, Y, X), generator=g, device=projector.device)
    x2 = torch.rand((1, 1, Y, X), generator=g, device=projector.device)
    y = torch.rand((1, Y, X), generator=g, device=projector.device)
    Hx = projector.forward_device(hm, x)
    Hty = projector.backward_device(hm, y)
    lhs = float((Hx.double() * y.double()).sum())
    rhs = float((x.double() * Hty.double()).sum())
    print("largest grid: <Hx,y> %.10e <x,Hty> %.10e" % (lhs, rhs), flush=True)
    assert abs(lhs - rhs) <= 1e-6 * abs(lhs)
    lin = projector.forward_device(hm, 2.0 * x + 3.0 * x2)
    ref = 2.0 * Hx.double() + 3.0 * projector.forward_device(hm, x2).double()
    assert float(torch.linalg.vector_norm(lin.double() - ref) / torch.linalg.vector_norm(ref)) < 1e-6
    with pytest.raises(ValueError, match="exceeds"):
        projector.prepare(hm, 1, 3750, 3750)        # (3750 + 344) / 15 -> 274 > 256


@pytest.mark.parametrize("N,Y,X,s2,grid", [(15, 150, 2000, 345, (540, 2400)), (15, 2000, 150, 345, (2400, 540)),
                                           (19, 190, 2200, 171, (380, 2375)), (15, 1800, 1800, 345, (2160, 2160))])
def test_large_grids_vs_oracle(projector, N, Y, X, s2, grid):
    """Fourier grids above 2048 points (out-of-place wave transforms in the
    column and row transforms) and reduced grids above 128 (the 256-point
    small FFT): rows of 2400 (even) and 2375 (odd, Nnum 19: with the prime
    pass), columns of 2400, and a square 2160 grid; forward and backward
    projection of one ring-PSF plane within 1e-5 of the FP64 oracle."""
    from flfhip.psfmatrix import MemoryHMatrix, adjoint_quadrant
    rng = np.random.default_rng(N + Y)
    Q = synthetic.ring_quadrant(N, s2, s2 / 4.0, 1.5, rng).astype(np.float32)
    hm = MemoryHMatrix([Q], [adjoint_quadrant(Q)])
    x = synthetic.blob_volume(1, Y, X, 60, seed=N + X)
    got = projector.ForwardProjectACC(hm, x, [0])
    assert projector.fshape(0) == grid
    ref = oproj.project_special(x[0], Q)
    e = rel_l2(got, ref)
    back = projector.BackwardProjectACC(hm, got, [0])[0]
    eb = rel_l2(back, oproj.project_special(got, hm.Hcc(0, True)))
    print("N %d %dx%d s2 %d grid %s: forward %.2e backward %.2e" % (N, Y, X, s2, grid, e, eb), flush=True)
    assert e < 1e-5 and eb < 1e-5


def test_camera_frame_2048(projector):
    """A 2048 x 2048 camera frame at Nnum 15 with a 345 px PSF (grid 2400 x
    2400, reduced 160): forward projection of one plane against the FP64
    oracle (about 20 s of CPU) and an RL iteration that stays finite."""
    from flfhip.lfdeconv import DeconvRL
    from flfhip.psfmatrix import MemoryHMatrix, adjoint_quadrant
    rng = np.random.default_rng(2048)
    Q = synthetic.ring_quadrant(15, 345, 90.0, 2.0, rng).astype(np.float32)
    hm = MemoryHMatrix([Q], [adjoint_quadrant(Q)])
    x = synthetic.blob_volume(1, 2048, 2048, 400, seed=3)
    got = projector.ForwardProjectACC(hm, x, [0])
    assert projector.fshape(0) == (2400, 2400)
    e = rel_l2(got, oproj.project_special(x[0], Q))
    print("2048^2 frame: forward rel-L2 %.2e" % e, flush=True)
    assert e < 1e-5
    X1 = DeconvRL(hm, None, 1, None, im=got + 0.01, projector=projector, logPrint=False)
    assert np.isfinite(X1).all() and X1.shape == (1, 1, 2048, 2048)


def test_full_size_plane_vs_oracle(projector):
    """The 1455 x 1455 / 41-plane workload's largest-PSF plane (z = -20,
    345 px, Fourier grid 1800): forward and back projection of one plane
    within 1e-5 relative L2 of the FP64 oracle (about 10 s of CPU)."""
    import torch
    from flfhip import workloads
    hm, zs, sides = workloads.psf_set("rl1455")
    assert sides[0] == 345
    Y = X = 1455
    rng = np.random.default_rng(71)
    x = rng.random((1, Y, X)).astype(np.float32)
    vol = np.zeros((hm.numZ, 1, Y, X), np.float32)
    vol[0] = x
    fwd = projector.ForwardProjectACC(hm, vol, [0])
    back = projector.BackwardProjectACC(hm, x, [0])
    ef = rel_l2(fwd, oproj.project_special(x.astype(np.float64), hm.Hcc(0, False)))
    eb = rel_l2(back[0], oproj.project_special(x.astype(np.float64), hm.Hcc(0, True)))
    print("1455^2 plane z=-20: forward rel-L2 %.2e, backward %.2e" % (ef, eb), flush=True)
    assert ef < 1e-5 and eb < 1e-5


def test_full_size_rl_two_planes_vs_oracle(projector):
    """Prevedel RL at 1455 x 1455 restricted to two planes of the 41-plane
    workload (z = -20, 345 px PSF, grid 1800; z = 0, 15 px PSF, grid 1500):
    Htf, two iterations from X = 1 and the per-iteration data fit within the
    bar of the FP64 oracle (volume 1e-4 relative L2, fit 1e-5)."""
    import torch
    from flfhip import workloads
    from flfhip.lfdeconv import deconv_rl_device
    from oracle import rl as orl
    hm, zs, sides = workloads.psf_set("rl1455")
    planes = [0, 20]
    assert (sides[0], sides[20]) == (345, 15)
    Y = X = 1455
    rng = np.random.default_rng(72)
    im = (rng.random((1, Y, X)) + 0.5).astype(np.float32)
    imt = projector.asnative(im)
    Htf = projector.backward_device(hm, imt, planes=planes)
    Xv = torch.ones((hm.numZ, 1, Y, X), dtype=torch.float32, device=projector.device)
    fits = []
    deconv_rl_device(projector, hm, Htf, 2, Xv, im=imt, planes=planes, fitOut=fits)
    Qf = [hm.Hcc(z, False) for z in planes]
    Qb = [hm.Hcc(z, True) for z in planes]
    Xr, rec = orl.deconv_rl(lambda v: oproj.forward_project(v, Qf), lambda i: oproj.backward_project(i, Qb), None, 2,
                            Xguess=np.ones((2, 1, Y, X)), im=im.astype(np.float64), record_fit=True)
    Xg = Xv.cpu().numpy()
    for k, z in enumerate(planes):
        e = rel_l2(Xg[z, 0], Xr[k, 0])
        print("1455^2 RL(2) plane %d rel-L2 %.2e" % (z, e), flush=True)
        assert e < 1e-4
    got = np.array([float(f[0]) for f in fits[:2]])
    ref = np.ravel(rec)[:2]
    assert np.max(np.abs(got - ref) / ref) < 1e-5
