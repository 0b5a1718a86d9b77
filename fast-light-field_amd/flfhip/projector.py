"""Projector_hip: the reference's projector interface on the MI355X path.

Same methods, argument meaning and return types as the reference projectors
(projector.py:769-790 Projector_base, 798-896 Projector_allC):
ForwardProjectACC / BackwardProjectACC plus asnative / asnumpy / nativeZeros /
nativeDivide / assertNative and the attributes cacheFH, storeVolumesOnGPU and
name.  Arithmetic runs only in libflf_hip.so (hand-written HIP kernels);
PyTorch is used for device memory and the current stream.
"""
import ctypes

import numpy as np
import torch

from . import _lib


def _stream_handle():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _shape_str(shape):
    return "(" + ", ".join(str(int(v)) for v in shape) + ")"


class Projector_hip(object):
    def __init__(self, device=0):
        if not torch.cuda.is_available():
            raise RuntimeError("Projector_hip needs a ROCm GPU (torch.cuda.is_available() is False)")
        self.lib = _lib.load()
        self.device = torch.device("cuda", device)
        self.name = "HIP gfx950"
        self.storeVolumesOnGPU = False
        self.cacheFH = True          # spectra always stay resident in HBM
        ctx = ctypes.c_void_p()
        rc = self.lib.flf_create(device, ctypes.byref(ctx))
        if rc != 0:
            raise _lib.FLFError("flf_create failed (%d)" % rc)
        self._ctx = ctx
        self._key = None
        self._tcap = self._frames = 0
        self._loaded = set()
        # plan_generation counts flf_plan calls (each frees the previous
        # plan's tables); _dims is the per-frame size setting (None: full
        # size), whose values the library uploads on every ragged call.  A
        # captured HIP graph holds pointers to the plan's tables and reads the
        # uploaded sizes, so it is replayed only while both are unchanged
        # (lfdeconv.RLGraph).
        self.plan_generation = 0
        self._dims = None

    def __del__(self):
        lib = getattr(self, "lib", None)
        ctx = getattr(self, "_ctx", None)
        if lib is not None and ctx:
            lib.flf_destroy(ctx)
            self._ctx = None

    # ---- reference helper interface (projector.py:777-790) ----
    def asnative(self, m):
        if isinstance(m, torch.Tensor):
            return m.to(self.device, dtype=torch.float32).contiguous()
        return torch.from_numpy(np.ascontiguousarray(m, dtype=np.float32)).to(self.device)

    def asnumpy(self, m):
        if isinstance(m, torch.Tensor):
            return m.detach().cpu().numpy()
        return np.asarray(m)

    def nativeZeros(self, shape, dtype=float):
        return torch.zeros(shape, dtype=torch.float32, device=self.device)

    def nativeDivide(self, a, b, out):
        return torch.div(a, b, out=out)

    def assertNative(self, m):
        # Volumes handed to rlUpdateFunc are device tensors when
        # storeVolumesOnGPU is set and numpy arrays otherwise.
        if self.storeVolumesOnGPU:
            assert isinstance(m, torch.Tensor) and m.is_cuda
        else:
            assert isinstance(m, np.ndarray)

    # ---- planning ----
    def _check(self, rc, what):
        _lib.check(self.lib, self._ctx, rc, what)

    def _ensure_plan(self, hMatrix, T, Y, X):
        """Plan for (hMatrix, Y, X) with room for at least T frames; a call
        with fewer frames than planned only selects the frame count (the
        spectrum cache stays)."""
        key = (id(hMatrix), getattr(hMatrix, "HPathFormat", None), hMatrix.numZ, getattr(hMatrix, "zStart", 0),
               hMatrix.Nnum, Y, X)
        if key == self._key and T <= self._tcap:
            if T != self._frames:
                self._check(self.lib.flf_set_frames(self._ctx, int(T)), "flf_set_frames")
                self._frames = T
                self._set_dims_state(None)      # flf_set_frames returns to full-size frames
            return
        s2 = []
        for cc in range(hMatrix.numZ):
            ps = hMatrix.PSFShape(cc)
            if ps[0] != ps[1]:
                raise ValueError("PSF must be square")
            s2.append(int(ps[0]))
        arr = (ctypes.c_int * len(s2))(*s2)
        self._key = None
        self._loaded = set()
        self.plan_generation += 1
        self._dims = None
        self._check(self.lib.flf_plan(self._ctx, int(hMatrix.Nnum), int(Y), int(X), int(T), len(s2), arr), "flf_plan")
        self._key = key
        self._tcap = self._frames = int(T)
        self._hMatrix = hMatrix

    def _ensure_psf(self, hMatrix, planes, direction):
        for cc in planes:
            if (cc, direction) in self._loaded:
                continue
            Q = np.ascontiguousarray(hMatrix.Hcc(cc, direction == 1), dtype=np.float32)
            self._check(self.lib.flf_set_psf(self._ctx, int(cc), int(direction), Q.ctypes.data_as(ctypes.c_void_p)),
                        "flf_set_psf")
            self._loaded.add((cc, direction))

    def prepare(self, hMatrix, T, Y, X, planes=None, directions=(0, 1), frames_capacity=None):
        """Plan and compute all spectra up front (the first-iteration cost of
        the reference's F(H) computation, done once).  frames_capacity
        (>= T) reserves buffers for that many frames per call."""
        planes = range(hMatrix.numZ) if planes is None else planes
        cap = max(int(T), int(frames_capacity or 0))
        if cap > T:
            self._ensure_plan(hMatrix, cap, Y, X)
        self._ensure_plan(hMatrix, T, Y, X)
        for d in directions:
            self._ensure_psf(hMatrix, planes, d)

    def set_frame_dims(self, dims):
        """Ragged batch: dims = [(Y_t, X_t)] per active frame, each within
        the planned Y x X (frame t in the top-left corner of its slot), or
        None for full-size frames (flf_set_frame_dims).  Outputs are zero
        outside each frame's region; each frame's result equals the
        projection at its own size."""
        if dims is None:
            self._check(self.lib.flf_set_frame_dims(self._ctx, None, int(self._frames), _stream_handle()),
                        "flf_set_frame_dims")
            self._set_dims_state(None)
            return
        flat = [int(v) for d in dims for v in d]
        if len(flat) != 2 * len(dims):
            raise ValueError("dims must be (Y, X) pairs")
        arr = (ctypes.c_int * len(flat))(*flat)
        self._check(self.lib.flf_set_frame_dims(self._ctx, arr, len(dims), _stream_handle()), "flf_set_frame_dims")
        full = self._key is not None and all((int(y), int(x)) == tuple(self._key[-2:]) for y, x in dims)
        self._set_dims_state(None if full else tuple(flat))

    def _set_dims_state(self, dims):
        self._dims = dims

    def state_key(self):
        """Identifies the device state a captured graph of this projector
        refers to (plan tables, active frame count, per-frame sizes): it
        changes whenever that state is freed or rewritten."""
        return (self.plan_generation, self._frames, self._dims)

    # options that the library applies at its next plan (flf_set_option)
    PLAN_OPTIONS = ("radix_order", "mac_ipt")

    def set_option(self, name, value):
        """Launch-shape option of the library (flf_set_option).  A plan
        option (PLAN_OPTIONS) takes effect from the next prepare(), which
        then re-plans and recomputes the spectra."""
        self._check(self.lib.flf_set_option(self._ctx, name.encode(), int(value)), "flf_set_option(%s)" % name)
        if name in self.PLAN_OPTIONS:
            self._key = None

    def get_option(self, name):
        v = ctypes.c_int()
        self._check(self.lib.flf_get_option(self._ctx, name.encode(), ctypes.byref(v)), "flf_get_option(%s)" % name)
        return v.value

    def fshape(self, cc):
        fy, fx = ctypes.c_int(), ctypes.c_int()
        self._check(self.lib.flf_plane_fshape(self._ctx, int(cc), ctypes.byref(fy), ctypes.byref(fx)), "fshape")
        return (fy.value, fx.value)

    @staticmethod
    def _plane_array(planes, numZ):
        planes = list(range(numZ)) if planes is None else [int(p) for p in planes]
        return planes, (ctypes.c_int * len(planes))(*planes)

    # ---- device-level entry points (tensors on this device) ----
    def _need(self, t, name, shape):
        """t must be a contiguous float32 tensor on this projector's device
        with exactly `shape`; checked before anything is queued."""
        if not isinstance(t, torch.Tensor):
            raise TypeError("%s must be a torch tensor on %s" % (name, self.device))
        if t.dtype != torch.float32:
            raise TypeError("%s must be float32, not %s" % (name, t.dtype))
        if t.device != self.device:
            raise ValueError("%s is on %s, the projector on %s" % (name, t.device, self.device))
        if tuple(t.shape) != tuple(int(v) for v in shape):
            raise ValueError("%s has shape %s, expected %s" % (name, _shape_str(t.shape), _shape_str(shape)))
        if not t.is_contiguous():
            raise ValueError("%s must be contiguous" % name)
        return t

    def _image(self, img, name="image"):
        if not isinstance(img, torch.Tensor) or img.dim() != 3:
            raise ValueError("%s must be a (T, Y, X) tensor" % name)
        img = img.contiguous() if img.dtype == torch.float32 else img
        return self._need(img, name, img.shape)

    def forward_device(self, hMatrix, vol, planes=None):
        """vol: (numZ, T, Y, X) float32 cuda tensor -> image (T, Y, X) tensor."""
        if not isinstance(vol, torch.Tensor) or vol.dim() != 4:
            raise ValueError("volume must be a (numZ, T, Y, X) tensor")
        nz, T, Y, X = vol.shape
        if nz != hMatrix.numZ:
            raise ValueError("volume has %d planes, HMatrix has %d" % (nz, hMatrix.numZ))
        vol = self._need(vol.contiguous() if vol.dtype == torch.float32 else vol, "volume", vol.shape)
        self._ensure_plan(hMatrix, T, Y, X)
        planes, arr = self._plane_array(planes, hMatrix.numZ)
        self._ensure_psf(hMatrix, planes, 0)
        img = torch.empty((T, Y, X), dtype=torch.float32, device=self.device)
        self._check(self.lib.flf_forward(self._ctx, _ptr(vol), _ptr(img), arr, len(planes), _stream_handle()),
                    "flf_forward")
        return img

    def forward_ratio_device(self, hMatrix, vol, im, planes=None, store_image=True):
        """Forward projection HX of vol and standard RL's error image im / HX
        from one pass (flf_forward_ratio).  Returns (HX or None, ratio)."""
        if not isinstance(vol, torch.Tensor) or vol.dim() != 4:
            raise ValueError("volume must be a (numZ, T, Y, X) tensor")
        nz, T, Y, X = vol.shape
        if nz != hMatrix.numZ:
            raise ValueError("volume has %d planes, HMatrix has %d" % (nz, hMatrix.numZ))
        vol = self._need(vol.contiguous() if vol.dtype == torch.float32 else vol, "volume", vol.shape)
        im = self._need(im, "im", (T, Y, X))
        self._ensure_plan(hMatrix, T, Y, X)
        planes, arr = self._plane_array(planes, hMatrix.numZ)
        self._ensure_psf(hMatrix, planes, 0)
        img = torch.empty((T, Y, X), dtype=torch.float32, device=self.device) if store_image else None
        ratio = torch.empty((T, Y, X), dtype=torch.float32, device=self.device)
        self._check(self.lib.flf_forward_ratio(self._ctx, _ptr(vol), _ptr(im),
                                               _ptr(img) if img is not None else ctypes.c_void_p(0), _ptr(ratio),
                                               arr, len(planes), _stream_handle()), "flf_forward_ratio")
        return img, ratio

    def forward_pair_device(self, hMatrix, volA, volB, planes=None):
        """Forward projection of two volumes (numZ, S, Y, X) as one call of 2S
        frames (flf_forward_pair): image (2S, Y, X), frames of volA first."""
        if not isinstance(volA, torch.Tensor) or volA.dim() != 4:
            raise ValueError("volA must be a (numZ, S, Y, X) tensor")
        nz, S, Y, X = volA.shape
        if nz != hMatrix.numZ:
            raise ValueError("volume has %d planes, HMatrix has %d" % (nz, hMatrix.numZ))
        volA = self._need(volA, "volA", volA.shape)
        volB = self._need(volB, "volB", volA.shape)
        self._ensure_plan(hMatrix, 2 * S, Y, X)
        planes, arr = self._plane_array(planes, hMatrix.numZ)
        self._ensure_psf(hMatrix, planes, 0)
        img = torch.empty((2 * S, Y, X), dtype=torch.float32, device=self.device)
        self._check(self.lib.flf_forward_pair(self._ctx, _ptr(volA), _ptr(volB), _ptr(img), arr, len(planes),
                                              _stream_handle()), "flf_forward_pair")
        return img

    def backward_device(self, hMatrix, img, planes=None, out=None):
        """img: (T, Y, X) cuda tensor -> volume (numZ, T, Y, X); planes not
        listed are zero (Projector_allC.BackwardProjectACC, projector.py:807).
        out: a (numZ, T, Y, X) float32 tensor to write instead (planes not
        listed are left as they are)."""
        img = self._image(img)
        T, Y, X = img.shape
        if out is not None:
            self._need(out, "out", (hMatrix.numZ, T, Y, X))
        self._ensure_plan(hMatrix, T, Y, X)
        planes, arr = self._plane_array(planes, hMatrix.numZ)
        self._ensure_psf(hMatrix, planes, 1)
        if out is None:
            out = torch.zeros((hMatrix.numZ, T, Y, X), dtype=torch.float32, device=self.device)
        self._check(self.lib.flf_backward(self._ctx, _ptr(img), _ptr(out), arr, len(planes), _stream_handle()),
                    "flf_backward")
        return out

    def rl_update_device(self, hMatrix, img, X, Htf=None, trust=None, mode=1, planes=None):
        """Back-project img and update X in place (mode 1 Prevedel: X *= Htf /
        B(img); mode 2 standard: X *= B(img) [/ trust]).  X and Htf are
        (numZ, T, Y, X), trust (numZ, 1, Y, X), all float32 and contiguous on
        this device."""
        img = self._image(img)
        T, Y, Xw = img.shape
        nz = hMatrix.numZ
        self._need(X, "X", (nz, T, Y, Xw))
        if mode == 1:
            if Htf is None:
                raise ValueError("the Prevedel update needs Htf")
            self._need(Htf, "Htf", (nz, T, Y, Xw))
        elif mode == 2:
            if trust is not None:
                self._need(trust, "trust", (nz, 1, Y, Xw))
        else:
            raise ValueError("mode must be 1 (Prevedel) or 2 (standard)")
        self._ensure_plan(hMatrix, T, Y, Xw)
        planes, arr = self._plane_array(planes, hMatrix.numZ)
        self._ensure_psf(hMatrix, planes, 1)
        hp = _ptr(Htf) if (Htf is not None and mode == 1) else ctypes.c_void_p(0)
        tp = _ptr(trust) if (trust is not None and mode == 2) else ctypes.c_void_p(0)
        self._check(self.lib.flf_backward_rl_update(self._ctx, _ptr(img), _ptr(X), hp, tp, int(mode),
                                                    arr, len(planes), _stream_handle()), "flf_backward_rl_update")

    def _scale_args(self, X, Htf, v):
        if not isinstance(X, torch.Tensor) or X.dim() != 4:
            raise ValueError("X must be a (numZ, T, Y, X) tensor")
        nz, T, Y, Xw = X.shape
        self._need(X, "X", X.shape)
        self._need(Htf, "Htf", X.shape)
        if not isinstance(v, torch.Tensor) or v.dim() != 4 or v.shape[1] not in (1, T):
            raise ValueError("v must be (numZ, 1 or T, Y, X)")
        self._need(v, "v", (nz, v.shape[1], Y, Xw))
        return nz, T, Y * Xw, int(v.shape[1])

    def rl_scale_device(self, X, Htf, v):
        """X *= Htf / v in place, NaN -> 0 (the Prevedel update with a known
        denominator v: (numZ, T, Y, X), or (numZ, 1, Y, X) shared by the
        frames)."""
        nz, T, npix, vf = self._scale_args(X, Htf, v)
        self._check(self.lib.flf_rl_scale(_ptr(Htf), _ptr(v), _ptr(X), nz, T, npix, vf, _stream_handle()),
                    "flf_rl_scale")

    def rl_ratio_device(self, X, Htf, v):
        """X = Htf / v, NaN -> 0 (the first Prevedel iterate from X = 1)."""
        nz, T, npix, vf = self._scale_args(X, Htf, v)
        self._check(self.lib.flf_rl_ratio(_ptr(Htf), _ptr(v), _ptr(X), nz, T, npix, vf, _stream_handle()),
                    "flf_rl_ratio")

    def datafit_device(self, img, im):
        """Per-frame FP64 (sum((img-im)^2), sum(im^2)).  im may hold fewer
        frames than img when img's frame count is a multiple of it: frame t
        of img is compared with frame t % len(im)."""
        img = self._image(img)
        im = self._image(im, "im")
        T, Y, X = img.shape
        Ti = im.shape[0]
        if im.shape[1:] != img.shape[1:] or Ti < 1 or T % Ti:
            raise ValueError("im %s does not match image %s" % (_shape_str(im.shape), _shape_str(img.shape)))
        if self._key is None or (Y, X) != tuple(self._key[-2:]):
            raise ValueError("datafit of (%d, %d) frames on a plan for %s" % (Y, X, self._key and self._key[-2:]))
        if T != self._frames:
            if not (1 <= T <= self._tcap):
                raise ValueError("datafit of %d frames on a plan for %d" % (T, self._tcap))
            self._check(self.lib.flf_set_frames(self._ctx, int(T)), "flf_set_frames")
            self._frames = T
            self._set_dims_state(None)
        out = (ctypes.c_double * (2 * T))()
        self._check(self.lib.flf_datafit(self._ctx, _ptr(img), _ptr(im), int(Ti), out, _stream_handle()),
                    "flf_datafit")
        return np.array(out[:], dtype=np.float64).reshape(T, 2)

    def profile(self, enable):
        """Bracket every multiply-accumulate launch with HIP events."""
        self._check(self.lib.flf_profile_enable(self._ctx, 1 if enable else 0), "flf_profile_enable")

    def profile_read(self):
        """(summed MAC kernel ms, launches, algorithmic bytes) since the last read."""
        ms, n, by = ctypes.c_double(), ctypes.c_longlong(), ctypes.c_double()
        self._check(self.lib.flf_profile_read(self._ctx, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(by)),
                    "flf_profile_read")
        return ms.value, n.value, by.value

    # ---- reference interface (projector.py:806-896) ----
    def BackwardProjectACC(self, hMatrix, projection, planes, progress=None, logPrint=False, numjobs=1,
                           rlUpdateFunc=None):
        if len(projection.shape) != 3:
            raise ValueError("projection must be (T, Y, X)")
        if isinstance(projection, np.ndarray) and projection.dtype != np.float32:
            raise TypeError("projection must be float32")
        img = self.asnative(projection)
        vol = self.backward_device(hMatrix, img, planes)
        if rlUpdateFunc is not None:
            rlUpdateFunc(vol if self.storeVolumesOnGPU else self.asnumpy(vol), slice(None))
            return None
        return vol if self.storeVolumesOnGPU else self.asnumpy(vol)

    def ForwardProjectACC(self, hMatrix, realspace, planes, progress=None, logPrint=False, numjobs=1):
        if len(realspace.shape) != 4:
            raise ValueError("realspace must be (numZ, T, Y, X)")
        if planes is not None:
            planes = list(planes)            # any iterable of plane indices, read once
            if not planes:
                # the reference has no projection to return then (its plane
                # sum stays None, projector.py:854, 895); forward_device
                # returns zeros
                raise ValueError("ForwardProjectACC needs at least one plane")
        if isinstance(realspace, np.ndarray) and realspace.dtype != np.float32:
            raise TypeError("realspace must be float32")
        vol = self.asnative(realspace)
        img = self.forward_device(hMatrix, vol, planes)
        return img if self.storeVolumesOnGPU else self.asnumpy(img)
