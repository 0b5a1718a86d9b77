"""Fourier-grid shape rules of the reference projector (oracle; test use only).

Restates special_fftconvolve.py:71-80 (``_centered``), :107-140
(``prime_factors``/``betterPrimes``/``BetterPrimeTable``) and :142-174
(``convolutionShape``), plus the 5-smooth padding the HIP path uses.
"""
import numpy as np


def prime_factors(n):
    """special_fftconvolve.py:107-120."""
    i, out = 2, []
    while i * i <= n:
        if n % i:
            i += 1
        else:
            n //= i
            out.append(i)
    if n > 1:
        out.append(n)
    return out


def smooth_at_least(n, max_prime):
    """Smallest m >= n whose prime factors are all <= max_prime
    (special_fftconvolve.py:122-129 with maxAcceptable=max_prime)."""
    m = max(int(n), 1)
    while m > 1 and max(prime_factors(m)) > max_prime:
        m += 1
    return m


def convolution_shape(in1_shape, in2_shape, nnum, pad="7smooth"):
    """special_fftconvolve.py:142-174.

    pad = "minimal": multiple of Nnum (padToSmallPrimes=False, :154-158);
          "7smooth": Nnum * 7-smooth (padToSmallPrimes=True, the CPU default,
                     :159-171, via BetterPrimeTable);
          "5smooth": Nnum * 5-smooth (the HIP path's choice; same linear
                     convolution, different zero padding).
    Returns (fshape, fslice, s1) like the reference.
    """
    s1 = np.array(in1_shape)
    s2 = np.array(in2_shape)
    if len(s1) == 3:
        s1 = s1[1:]
    shape = s1 + s2 - 1
    red = [int(np.ceil(d / float(nnum))) for d in shape]
    if pad == "minimal":
        fshape = [r * nnum for r in red]
    elif pad == "7smooth":
        # BetterPrimeTable(Nnum)[i] = betterPrimes(i*Nnum) for i >= 2, else i*Nnum
        fshape = [(smooth_at_least(r, 7) if r >= 2 else r) * nnum for r in red]
    elif pad == "5smooth":
        fshape = [smooth_at_least(r, 5) * nnum for r in red]
    else:
        raise ValueError(pad)
    fslice = tuple(slice(0, int(sz)) for sz in shape)
    return fshape, fslice, s1


def centered(arr, newsize):
    """special_fftconvolve.py:71-80: the centre ``newsize`` portion (the
    'same' crop of fftconvolve)."""
    currsize = np.array(arr.shape)
    newsize = np.asarray(newsize)
    if len(currsize) > len(newsize):
        newsize = np.append(currsize[:len(currsize) - len(newsize)], newsize)
    start = (currsize - newsize) // 2
    end = start + newsize
    return arr[tuple(slice(start[k], end[k]) for k in range(len(end)))]
