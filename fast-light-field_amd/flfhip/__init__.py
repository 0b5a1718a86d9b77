"""flfhip - MI355X-native light-field projector and Richardson-Lucy loop.

Drop-in for the projector / lfdeconv call surface of fast-light-field:
    from flfhip import Projector_hip, lfdeconv, psfmatrix
    X = lfdeconv.DeconvRL(hMatrix, Htf=None, maxIter=8, Xguess=None, im=image)
All arithmetic runs in libflf_hip.so (HIP kernels for gfx950).
"""
from . import psfmatrix, synthetic  # noqa: F401

__all__ = ["Projector_hip", "lfdeconv", "psfmatrix", "synthetic"]


def __getattr__(name):
    # The projector and driver import torch and the HIP library lazily, so
    # that the PSF containers and generators stay importable without them.
    import importlib
    if name == "Projector_hip":
        return importlib.import_module(".projector", __name__).Projector_hip
    if name == "lfdeconv":
        return importlib.import_module(".lfdeconv", __name__)
    raise AttributeError(name)
