"""Autotuner: times one forward and one backward projection per candidate
setting at the caller's shape and keeps the fastest (the role FFTW_MEASURE
planning plays for the reference, py_light_field.cpp:25-31, and its
CalibrateBlockFactors, projector.py:628-675).

Plan options come first, each candidate a fresh plan (spectra recomputed):
the pass order of every FFT (``radix_order``: largest radices first,
reversed, without radix 8, odd radices first) and the items per MAC thread
(``mac_ipt``).  Then the launch shapes: the matrix-core MAC (``mac_mfma``,
8 frames per workgroup; only when a call holds several frames), one dispatch per stage for all grid
groups or one per group (``grouped_launch``), the forward MAC's plane ranges
(``fwd_split``), the frames per MAC workgroup (``mac_frames``, only when a
call holds several frames), the columns per column-transform workgroup
(``ifft_col_batch``) and the planes per back-projection MAC workgroup
(``bwd_planes``).  ``bwd_frame_group`` is not searched: at 32 x 1455^2 it
saves about 5% of the back-projection MAC but multiplies that launch's HBM
reads several times over, for a whole-step difference within run-to-run
noise.  Each option is searched with the others fixed (coordinate
descent, one pass); results depend on them only through FP32 rounding.
"""
import torch

PLAN_CANDIDATES = {
    "radix_order": [0, 1, 2, 3],
    "mac_ipt": [2, 4, 1],
}

CANDIDATES = {
    "mac_mfma": [0, 1],
    "grouped_launch": [1, 0],
    "fwd_split": [1, 2, 3, 4],
    "mac_frames": [1, 2, 3, 4],
    "ifft_col_batch": [2, 4, 8],
    "bwd_planes": [1, 2, 3, 4],
}


def _time_pass(projector, hMatrix, vol, img, reps):
    projector.forward_device(hMatrix, vol)
    projector.backward_device(hMatrix, img)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        projector.forward_device(hMatrix, vol)
        projector.backward_device(hMatrix, img)
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def tune(projector, hMatrix, frames, Y, X, reps=3, log=None, plan_options=True):
    """Tune the projector for (hMatrix, frames, Y, X); returns
    {option: value} and the timings (ms per forward + backward).  The plan
    keeps the frame capacity the projector had (at least frames)."""
    cap = max(int(frames), int(getattr(projector, "_tcap", 0) or 0))

    def plan():
        projector.prepare(hMatrix, frames, Y, X, frames_capacity=cap)

    plan()
    g = torch.Generator(device=projector.device)
    g.manual_seed(0)
    vol = torch.rand((hMatrix.numZ, frames, Y, X), generator=g, device=projector.device)
    img = torch.rand((frames, Y, X), generator=g, device=projector.device)
    chosen, timings = {}, {}

    def search(name, values, replan):
        best, best_ms = None, None
        for v in values:
            try:
                projector.set_option(name, v)
                if replan:
                    plan()
            except Exception:
                continue                       # e.g. a column batch the LDS cannot hold
            ms = _time_pass(projector, hMatrix, vol, img, reps)
            timings["%s=%d" % (name, v)] = ms
            if log:
                log("autotune %s=%d: %.3f ms" % (name, v, ms))
            if best_ms is None or ms < best_ms:
                best, best_ms = v, ms
        if best is not None:
            projector.set_option(name, best)
            if replan:
                plan()
            chosen[name] = best

    if plan_options:
        for name, values in PLAN_CANDIDATES.items():
            search(name, values, True)
    for name, values in CANDIDATES.items():
        if name in ("mac_frames", "mac_mfma") and frames < 2:
            continue
        search(name, values, False)
    del vol, img
    return chosen, timings
