"""CPU oracle for the light-field projector and Richardson-Lucy loop.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (``fast-light-field_amd/``)
imports this package; only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker or as the
CPU baseline that is timed beside the GPU path.

Everything here is an FP64 numpy restatement of the reference algorithm, with
each function citing the reference file:line it follows:

* ``shapes``    - Fourier-grid shape rules (special_fftconvolve.py:71-80,
                  122-174).
* ``psf``       - quadrant <-> full PSF expansion by the 8-fold symmetry
                  (py_light_field.cpp:1514-1562, projector.py:159-189), the
                  adjoint back-projection PSF set, and a restatement of the
                  reference's own Ht construction, calcHt_new
                  (generate_psf.py:646-719).
* ``projector`` - the direct per-offset convolution projector
                  (projector.py:1084-1126) and a literal restatement of the
                  special-FFT projector (py_light_field.cpp:284-341, 711-1102,
                  1417-1740; projector.py:103-230, 792-896).
* ``rl``        - Richardson-Lucy, Prevedel and standard updates
                  (lfdeconv.py:72-202).

How each function is pinned (none is "parity unpinned"):

* ``shapes.convolution_shape``: the reference's own known answers
  (py_light_field.cpp:1904-1925) and the closed-form definition.
* ``psf.expand_quadrant``: every one of the N*N offsets is produced exactly
  once from the quadrant; flips/transposes checked element-wise.
* ``psf.adjoint_psf``: <F x, y> == <x, B y> to 1e-12 with B built from it,
  on random inputs, through the direct projector.
* ``projector.project_direct``: against an explicit shift-and-add loop
  (brute force) on tiny inputs, and against its own linearity.
* ``projector.project_special``: against ``project_direct`` (a different
  algorithm: masked full-grid convolution vs lenslet-subsampled Fourier
  tiling with mirrored/transposed spectra) to 1e-12, for square and
  non-square grids, odd and even Fourier sizes.
* ``psf.calc_ht_stamp``: equals the exact adjoint (``adjoint_psf``) on every
  pixel it keeps, zero on the h-1-b leading rows and h-1-a leading columns
  it drops; a planted missing rotation fails.
* ``rl.deconv_rl`` (Prevedel): closed form of one iteration on a 1-plane
  identity-PSF system, and monotone data-fit on a small synthetic problem.
* ``rl.deconv_rl`` (standard, ``prevedel=False``) and
  ``rl.rl_update_standard`` with a trust mask: identity-PSF closed form
  (X1 = im, a fixed point), a hand-built N = 3 trust-mask case checked
  against X0 * B(im / F X0) / B(trust) written with explicit loops, the
  0/0 -> 0 and x/0 -> inf semantics of lfdeconv.py:156-164, and a planted
  double trust division that must fail.

The compiled reference (py_light_field + vendored FFTW 3.3.8) is not built
here: FFTW needs the headers its own configure step generates, which the
rules for this oracle forbid producing or standing in for.
"""
