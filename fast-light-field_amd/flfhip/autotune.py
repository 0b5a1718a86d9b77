"""Launch-shape autotuner: times one forward and one backward projection per
candidate setting at the caller's shape and keeps the fastest.

The tuned options (flf_set_option) are the forward MAC's plane ranges
(``fwd_split``), the frames per MAC workgroup (``mac_frames``, only when a call
holds several frames), the columns per column-transform workgroup
(``ifft_col_batch``) and the planes per back-projection MAC workgroup
(``bwd_planes``).  Each is searched with the others fixed (coordinate
descent, one pass); results depend on them only through FP32 rounding.
"""
import torch

CANDIDATES = {
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


def tune(projector, hMatrix, frames, Y, X, reps=3, log=None):
    """Tune the projector for (hMatrix, frames, Y, X); returns
    {option: value} and the timings (ms per forward + backward)."""
    projector.prepare(hMatrix, frames, Y, X)
    g = torch.Generator(device=projector.device)
    g.manual_seed(0)
    vol = torch.rand((hMatrix.numZ, frames, Y, X), generator=g, device=projector.device)
    img = torch.rand((frames, Y, X), generator=g, device=projector.device)
    chosen, timings = {}, {}
    for name, values in CANDIDATES.items():
        if name == "mac_frames" and frames < 2:
            continue
        best, best_ms = None, None
        for v in values:
            try:
                projector.set_option(name, v)
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
            chosen[name] = best
    del vol, img
    return chosen, timings
