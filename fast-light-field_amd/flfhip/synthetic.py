"""Seeded synthetic inputs of the benchmark workloads.

* PSFs: per plane z and lenslet offset (b, a) a Gaussian-blurred ring of
  radius 3 + 6|z| px, centred on the lenslet centre seen from that offset,
  with a seeded radial modulation; radially symmetric, so the full set has the
  8-fold mirror/transpose symmetry the reference's quadrant storage assumes
  (py_light_field.cpp:1514-1562).  Each PSF sums to 1 (the per-PSF
  normalisation of generate_psf.py:606-613).  Ht is derived as the adjoint.
* Volumes: Gaussian blobs (sigma 1-3 px, amplitude U[0.5, 2]) on a 0.01
  background; images: Poisson counts of the forward projection scaled to a
  given peak.
"""
import numpy as np

from .psfmatrix import MemoryHMatrix, adjoint_quadrant


def psf_sides(zs, s_min, s_max):
    """Odd PSF sides growing linearly in |z| from s_min (z=0) to s_max (max |z|)."""
    zs = np.asarray(list(zs))
    zmax = max(1, int(np.max(np.abs(zs))))
    half = (s_max - s_min) / 2.0
    return [int(s_min + 2 * round(half * abs(int(z)) / zmax)) for z in zs]


def ring_quadrant(nnum, s2, radius, blur, rng):
    """(h, h, s2, s2) float32 quadrant of ring PSFs for one plane."""
    h = (nnum + 1) // 2
    c = nnum // 2
    c2 = (s2 - 1) // 2
    # Seeded radial modulation (a function of radius only keeps the symmetry).
    k = rng.uniform(0.5, 3.0, size=3)
    ph = rng.uniform(0, 2 * np.pi, size=3)
    amp = rng.uniform(0.0, 0.25, size=3)
    y = np.arange(s2, dtype=np.float64)
    Q = np.zeros((h, h, s2, s2), dtype=np.float32)
    for b in range(h):
        for a in range(h):
            cy = c2 - (b - c)
            cx = c2 - (a - c)
            rho = np.sqrt((y[:, None] - cy) ** 2 + (y[None, :] - cx) ** 2)
            prof = np.exp(-0.5 * ((rho - radius) / blur) ** 2)
            mod = 1.0 + np.sum(amp[:, None, None] * np.cos(k[:, None, None] * rho[None] / max(radius, 1.0) * np.pi
                                                          + ph[:, None, None]), axis=0)
            p = prof * np.clip(mod, 0.05, None)
            Q[b, a] = (p / p.sum()).astype(np.float32)
    return Q


def make_psf_set(nnum, zs, sides, seed=0, radius=lambda z: 3.0 + 6.0 * abs(z), blur=lambda z: 1.0 + 0.05 * abs(z)):
    """Quadrants H and adjoint Ht for every plane; returns a MemoryHMatrix."""
    rng = np.random.default_rng(seed)
    H, Ht = [], []
    for z, s2 in zip(zs, sides):
        r = min(radius(z), (s2 - 1) / 2.0 - nnum / 2.0 - 2.0 * blur(z))
        Q = ring_quadrant(nnum, s2, max(r, 0.0), blur(z), rng)
        H.append(Q)
        Ht.append(adjoint_quadrant(Q).astype(np.float32))
    return MemoryHMatrix(H, Ht, identifier="synthetic-ring:N%d:seed%d:%s" % (nnum, seed, ",".join(map(str, sides))))


def blob_volume(nz, Y, X, nblobs, seed, background=0.01, sigma=(1.0, 3.0), amp=(0.5, 2.0), T=1):
    """(nz, T, Y, X) float32 volume of Gaussian blobs."""
    rng = np.random.default_rng(seed)
    vol = np.full((nz, T, Y, X), background, dtype=np.float32)
    for t in range(T):
        zc = rng.uniform(0, nz, nblobs)
        yc = rng.uniform(0, Y, nblobs)
        xc = rng.uniform(0, X, nblobs)
        sg = rng.uniform(sigma[0], sigma[1], nblobs)
        am = rng.uniform(amp[0], amp[1], nblobs)
        for i in range(nblobs):
            r = int(np.ceil(3 * sg[i]))
            y0, y1 = max(0, int(yc[i]) - r), min(Y, int(yc[i]) + r + 1)
            x0, x1 = max(0, int(xc[i]) - r), min(X, int(xc[i]) + r + 1)
            z0, z1 = max(0, int(zc[i]) - r), min(nz, int(zc[i]) + r + 1)
            zz = np.arange(z0, z1)[:, None, None]
            yy = np.arange(y0, y1)[None, :, None]
            xx = np.arange(x0, x1)[None, None, :]
            g = am[i] * np.exp(-((zz - zc[i]) ** 2 + (yy - yc[i]) ** 2 + (xx - xc[i]) ** 2) / (2 * sg[i] ** 2))
            vol[z0:z1, t, y0:y1, x0:x1] += g.astype(np.float32)
    return vol


def poisson_image(clean, peak, seed):
    """Poisson counts of a clean forward projection scaled so its max is `peak`."""
    rng = np.random.default_rng(seed)
    clean = np.asarray(clean, dtype=np.float64)
    scale = peak / max(float(clean.max()), 1e-30)
    return rng.poisson(np.clip(clean, 0, None) * scale).astype(np.float32)


def drifting_blob_frames(nz, shapes, nblobs, seed, drift=0.5, field=None, background=0.01, sigma=(1.0, 3.0),
                         amp=(0.5, 2.0)):
    """A time series of blob volumes seen through regions of uneven size.

    One seeded set of max(nblobs) blobs fills a field (default: the largest
    frame); blob i moves by `drift` px per frame in a seeded direction.
    Frame t shows the first nblobs[t] blobs inside its region, the top-left
    shapes[t] = (Y_t, X_t) of the field.  Returns [(nz, Y_t, X_t) float32].
    """
    T = len(shapes)
    nblobs = np.broadcast_to(np.asarray(nblobs, int), (T,))
    FY = field[0] if field else max(s[0] for s in shapes)
    FX = field[1] if field else max(s[1] for s in shapes)
    nmax = int(nblobs.max())
    rng = np.random.default_rng(seed)
    zc = rng.uniform(0, nz, nmax)
    yc = rng.uniform(0, FY, nmax)
    xc = rng.uniform(0, FX, nmax)
    sg = rng.uniform(sigma[0], sigma[1], nmax)
    am = rng.uniform(amp[0], amp[1], nmax)
    ang = rng.uniform(0, 2 * np.pi, nmax)
    out = []
    for t, (Y, X) in enumerate(shapes):
        vol = np.full((nz, Y, X), background, dtype=np.float32)
        for i in range(int(nblobs[t])):
            cy = yc[i] + drift * t * np.sin(ang[i])
            cx = xc[i] + drift * t * np.cos(ang[i])
            r = int(np.ceil(3 * sg[i]))
            y0, y1 = max(0, int(cy) - r), min(Y, int(cy) + r + 1)
            x0, x1 = max(0, int(cx) - r), min(X, int(cx) + r + 1)
            z0, z1 = max(0, int(zc[i]) - r), min(nz, int(zc[i]) + r + 1)
            if y0 >= y1 or x0 >= x1 or z0 >= z1:
                continue
            zz = np.arange(z0, z1)[:, None, None]
            yy = np.arange(y0, y1)[None, :, None]
            xx = np.arange(x0, x1)[None, None, :]
            g = am[i] * np.exp(-((zz - zc[i]) ** 2 + (yy - cy) ** 2 + (xx - cx) ** 2) / (2 * sg[i] ** 2))
            vol[z0:z1, y0:y1, x0:x1] += g.astype(np.float32)
        out.append(vol)
    return out
