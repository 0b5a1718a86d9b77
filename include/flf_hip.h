/*
 * flf_hip.h - C ABI of the MI355X (gfx950) light-field projector and
 * Richardson-Lucy loop.
 *
 * One context per device.  The caller owns every image/volume buffer (device
 * pointers); the context owns the PSF spectrum cache and its work buffers.
 * Every entry point returns FLF_OK (0) or a negative FLF_E* code and never
 * aborts; flf_last_error() gives the message.  A context is not re-entrant:
 * calls on one context must not overlap (distinct contexts are independent).
 *
 * Every array argument that a kernel reads or writes must be device memory
 * (hipMalloc, including a caching allocator's blocks) of the context's
 * device, large enough for the layout below; anything else (host memory,
 * pinned or not, another device's memory, a too-small allocation) returns
 * FLF_EINVAL before any work is queued.  Host arrays are named *_host.
 *
 * Array layouts (C order, float32):
 *   volume  [nz][T][Y][X]      image  [T][Y][X]
 *   PSF quadrant of one plane [h][h][s2][s2], h = (Nnum+1)/2, entry [b][a]
 *   used for b <= a (the reference's HMatrix.Hcc layout, psfmatrix.py:46-66).
 *
 * Each entry point names the reference interface it replaces.
 */
#ifndef FLF_HIP_H
#define FLF_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FLF_OK 0
#define FLF_EINVAL (-1)   /* bad argument, shape or stride  (reference: TypeError/ValueError) */
#define FLF_ENOMEM (-2)   /* device allocation failed        (reference: MemoryError)         */
#define FLF_EHIP (-3)     /* HIP runtime / launch failure                                      */
#define FLF_ESTATE (-4)   /* not planned, or PSF of a requested plane not uploaded             */

typedef struct flf_ctx flf_ctx;

/* Library/ABI version (no reference counterpart; lets bindings check the .so). */
int flf_version(void);

/* Create a context on HIP device `device`.
 * Replaces the module-global state of py_light_field (py_light_field.cpp:
 * 15-37, 1319-1363: thread count, planning mode, F(H) cache). */
int flf_create(int device, flf_ctx** out);

/* Destroy a context and free its device memory (ClearFHCache,
 * py_light_field.cpp:1323-1335). */
void flf_destroy(flf_ctx* ctx);

/* Message of the last error on this context ("" if none).  With ctx NULL:
 * the last error of this host thread's context-free calls (flf_rl_scale,
 * flf_rl_ratio, flf_net_upsample, flf_mfma_gemm_nt). */
const char* flf_last_error(const flf_ctx* ctx);

/* Plan the geometry: lenslet pitch nnum (odd), image Y x X, T frames per
 * batch, nz planes with odd PSF sides s2[z].  Chooses the Fourier grid of
 * each plane (convolutionShape, special_fftconvolve.py:142-174, padded to a
 * 5-smooth multiple of nnum), builds the kernel tables and allocates the
 * spectrum cache and work buffers.  Replaces the per-plane set-up of
 * ProjectorForZ_base.__init__ (projector.py:106-151).  Drops any uploaded
 * PSFs.  At most 256 planes, and nz * T <= 65535. */
int flf_plan(flf_ctx* ctx, int nnum, int Y, int X, int T, int nz, const int* s2);

/* Frames per call.  flf_plan's T is the capacity the work buffers hold;
 * flf_set_frames selects how many frames (1..capacity) the following calls
 * process, without touching the spectrum cache.  The reference plans per
 * call shape (ProjectorForZ_base.__init__, projector.py:106-151) and has no
 * counterpart; callers that alternate between batch sizes (a two-frame
 * check, then one-frame iterations) use it to keep one cache. */
int flf_set_frames(flf_ctx* ctx, int T);
int flf_frame_capacity(const flf_ctx* ctx);

/* Forward projection of 2S active frames held in two buffers: frames
 * 0..S-1 from volA, frames S..2S-1 from volB, each (nz, S, Y, X); img is
 * (2S, Y, X).  Same result as flf_forward on the two volumes concatenated
 * along frames (ForwardProjectACC, projector.py:849-896), without the copy:
 * the learned initialiser projects its guess and the first plain iterate
 * together this way.  Not with per-frame sizes. */
int flf_forward_pair(flf_ctx* ctx, const float* volA, const float* volB, float* img, const int* planes,
                     int nplanes, void* stream);

/* Ragged batches: dims[2t], dims[2t+1] = valid rows and columns of frame t
 * (1..Y, 1..X) for the T active frames (T must equal the active count);
 * NULL returns to full-size frames.  Frame t occupies the top-left corner
 * of its Y x X slot: forward projection reads only that region of every
 * plane and writes zeros outside it, back projection reads only that region
 * of the image and writes (or, in the update modes, sets) the volume to
 * zero outside it.  Each frame's result equals the projection at its own
 * size, because zero padding of a linear convolution does not change the
 * cropped output.  The reference refuses mixed shapes in one call
 * (deconvolve_time_series.py:21-23) and plans one convolution shape per
 * image size (special_fftconvolve.py:142-174); here frames of uneven
 * lenslet grids share one plan and one spectrum cache.  flf_set_frames
 * resets to full-size frames.  The upload is ordered on stream, which is
 * synchronised before returning. */
int flf_set_frame_dims(flf_ctx* ctx, const int* dims, int T, void* stream);

/* Launch-shape options (no reference counterpart; the reference's
 * SetNumThreadsToUse / SetPlanningMode, py_light_field.cpp:216-241, are its
 * tuning knobs).  Names and values:
 *   "fwd_split"      0 (automatic) or 1..4 plane ranges of the forward MAC
 *   "mac_frames"     1..4 frames per MAC workgroup (upper bound)
 *   "mac_mfma"       0/1: matrix-core MAC (FP32 MFMA, 8 frames per
 *                    workgroup) for calls of >= 2 frames
 *   "bwd_planes"     1..8 planes per back-projection MAC workgroup
 *   "bwd_frame_group" 0..64 frame blocks of a chunk the back-projection MAC
 *                    dispatches side by side (0: all; bit-identical results)
 *   "ifft_col_batch" 1..8 columns per column-transform workgroup (after plan)
 *   "grouped_launch" 0/1 (default 1): the small FFTs, the multiply-accumulate
 *                    and the column transforms of all Fourier-grid groups of
 *                    a projection go out as one dispatch each (bit-identical
 *                    results); not with group_streams or mac_mfma
 *   "group_streams"  0/1 (default 0): the Fourier-grid groups of a projection
 *                    run side by side on library-owned streams, forked from
 *                    and joined back to the caller's stream by events (one
 *                    stream while the caller's stream is captured)
 * Plan options, applied by the next flf_plan (the FFTW_MEASURE-style
 * choices of py_light_field.cpp:25-31):
 *   "radix_order"    0..3 pass order of every FFT: 0 largest radices first,
 *                    1 reversed, 2 no radix 8 (wave-level transforms), 3 odd
 *                    radices first (wave-level); a prime pass stays last
 *   "mac_ipt"        1, 2 or 4 items per MAC lane group (2 / 4: lane pairs / quads share them)
 * Results do not depend on them beyond FP32 rounding (group_streams: not at
 * all).  flfhip.autotune picks them by timing. */
int flf_set_option(flf_ctx* ctx, const char* name, int value);
int flf_get_option(const flf_ctx* ctx, const char* name, int* value);

/* Fourier grid of plane z after flf_plan (fshape of projector.py:123). */
int flf_plane_fshape(const flf_ctx* ctx, int z, int* fY, int* fX);

/* Bytes of device memory held by the spectrum cache and work buffers. */
long long flf_device_bytes(const flf_ctx* ctx);

/* Upload the quadrant PSF of plane z for direction dir (0: H, used by the
 * forward projection; 1: Ht, used by the backward projection) from HOST
 * memory, and compute its spectra into the cache.  Replaces FHWorkItem /
 * TransposeWorkItem / MirrorWorkItem with caching enabled
 * (py_light_field.cpp:711-870, 1365-1385): spectra persist until the next
 * flf_plan or flf_destroy. */
int flf_set_psf(flf_ctx* ctx, int z, int dir, const float* quadrant_host);

/* Forward projection: img[T][Y][X] = sum over the listed planes of the
 * projection of vol[z] (device pointers).  planes == NULL means all nz.
 * Replaces ProjectForZList(backwards=0) + InverseRFFTList + CropAfterFFT and
 * the plane sum of Projector_allC.ForwardProjectACC (projector.py:851-896;
 * py_light_field.cpp:1417-1740).  stream: hipStream_t or NULL. */
int flf_forward(flf_ctx* ctx, const float* vol, float* img, const int* planes, int nplanes, void* stream);

/* Forward projection fused with standard RL's error image: ratio[T][Y][X] =
 * im / (sum over the listed planes of the projection of vol), IEEE division
 * (0/0 = NaN, x/0 = inf, as numpy), written by the pass that writes the
 * image; img may be NULL (image not stored).  Replaces ForwardProjectACC
 * followed by `error = im / HXguess` in DeconvRL's standard branch
 * (lfdeconv.py:189-192). */
int flf_forward_ratio(flf_ctx* ctx, const float* vol, const float* im, float* img, float* ratio, const int* planes,
                      int nplanes, void* stream);

/* Backward projection: vol[z] for each listed plane (others untouched) from
 * img[T][Y][X].  Replaces Projector_allC.BackwardProjectACC with
 * rlUpdateFunc=None (projector.py:806-849). */
int flf_backward(flf_ctx* ctx, const float* img, float* vol, const int* planes, int nplanes, void* stream);

/* Backward projection of img fused with the Richardson-Lucy update of X:
 *   mode 1 (Prevedel, lfdeconv.py:112-126):  X[z] *= Htf[z] / B(img)[z]
 *   mode 2 (standard, lfdeconv.py:151-164):  X[z] *= B(img)[z] [/ trust[z]]
 * then NaN -> 0.  Replaces BackwardProjectACC(..., rlUpdateFunc=...)
 * (projector.py:844-847).  Htf is required for mode 1; trust (nz,1,Y,X) is
 * optional for mode 2. */
int flf_backward_rl_update(flf_ctx* ctx, const float* img, float* X, const float* Htf, const float* trust,
                           int mode, const int* planes, int nplanes, void* stream);

/* Forward-model data fit per frame, FP64 sums: out[2t] = sum((img-im)^2),
 * out[2t+1] = sum(im^2) (host array of 2*T doubles), where frame t of img
 * (T active frames) is compared with frame t % im_frames of im (im_frames 0:
 * T; a guard that projects two volumes per image passes im once).  The
 * reference has no counterpart (DeconvRL always runs maxIter,
 * lfdeconv.py:185-196); this is the convergence measure of this library's RL
 * driver.  Synchronises `stream`. */
int flf_datafit(flf_ctx* ctx, const float* img, const float* im, int im_frames, double* out_host, void* stream);

/* The data fit of flf_datafit without any allocation, for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  No reference counterpart.
 * img: T frames of npix floats, im: im_frames frames (frame t of img against
 * frame t % im_frames), out: 2*T doubles in DEVICE memory, written on
 * `stream` (no synchronisation, no host copy): out[2t] = sum((img-im)^2),
 * out[2t+1] = sum(im^2), bit-identical to flf_datafit.  T <= 512.  The
 * partial sums live in one fixed 2 MiB device array per GPU (held until the
 * process ends) that every call on that GPU shares: calls that overlap on
 * one GPU (several streams or threads) race and must be serialised by the
 * caller.  Arrays on the current HIP device; errors via flf_last_error(NULL). */
int flf_datafit_noalloc(const float* img, const float* im, int T, int im_frames, long long npix, double* out,
                        void* stream);

/* X *= Htf / v, NaN -> 0, on `stream` (device of the current HIP device):
 * the Prevedel RL update of lfdeconv.py:112-126 with a denominator the
 * caller already holds.  X and Htf are (nz, T, npix); v is (nz, v_frames,
 * npix) with v_frames = T, or 1 for one slab per plane shared by every
 * frame.  With X = 1 and v = Ht(H 1) (which does not depend on the image) it
 * gives the first plain RL iterate from a uniform start without projecting;
 * the learned initialiser's guard compares its guess against that iterate.
 * nz * T <= 65535. */
int flf_rl_scale(const float* Htf, const float* v, float* X, int nz, int T, long long npix, int v_frames,
                 void* stream);

/* X = Htf / v, NaN -> 0: flf_rl_scale applied to X = 1 without reading X
 * (the first plain RL iterate from a uniform start, lfdeconv.py:112-126 with
 * Xguess = 1). */
int flf_rl_ratio(const float* Htf, const float* v, float* X, int nz, int T, long long npix, int v_frames,
                 void* stream);

/* FLF_OK if [p, p + bytes) is device memory of `device` in one allocation,
 * else FLF_EINVAL with the reason in flf_last_error(NULL): the check every
 * array entry point applies to its arguments, exposed so that a binding can
 * test a buffer before handing it over.  Host-side only (queues nothing).
 * No reference counterpart (the reference's py_light_field checks numpy
 * dtypes and strides, py_light_field.cpp:1489-1493). */
int flf_check_device_buffer(int device, const void* p, long long bytes);

/* Kernel timing for roofline reporting (no reference counterpart; the
 * reference times whole calls with time.time(), lf_performance_analysis.py:
 * 259-268).  While enabled, every multiply-accumulate launch is bracketed by
 * HIP events on its own stream.  flf_profile_read synchronises those events
 * and returns the summed kernel time, the number of launches and their
 * algorithmic bytes (compulsory spectrum reads), then resets the record. */
int flf_profile_enable(flf_ctx* ctx, int enable);
int flf_profile_read(flf_ctx* ctx, double* mac_ms, long long* mac_launches, double* mac_bytes);

/* Host-only planning queries (no GPU needed).
 * flf_fourier_len: Fourier grid length of one axis for image side s1 and PSF
 * side s2 (convolutionShape, special_fftconvolve.py:142-174, with 5-smooth
 * padding of the reduced length).
 * flf_grid_check: builds the kernel tables of an fY x fX grid (as planned:
 * items in pairs of one residue class) and verifies that every half-plane
 * bin is written exactly once and that every pair shares its residues;
 * stats[6] = {workgroups,
 * items, max residues per orbit, stored spectra per plane, slots per item,
 * max items per workgroup}.  Returns FLF_EINVAL if the tables are invalid.
 * flf_radix_plan: the passes of a length-L transform for plan option
 * radix_order = order (wide = 1: a wave-level transform) into radices[16];
 * *npass receives their count.  FLF_EINVAL if L has no such plan. */
int flf_fourier_len(int s1, int s2, int nnum);
int flf_grid_check(int nnum, int fY, int fX, long long* stats);
int flf_radix_plan(int L, int wide, int order, int* radices, int* npass);

/* ---- Learned initialiser (no reference counterpart: the reference starts
 * RL from Htf or a caller-supplied Xguess, lfdeconv.py:97-104) ----
 * A lenslet-aware conv net on BF16 MFMA kernels with FP32 accumulation:
 * pixel-unshuffled image (nnum^2 channels at lenslet resolution) -> 3x3 conv
 * to `width` channels -> `nblocks` residual blocks -> 3x3 conv to nz planes ->
 * softplus, scaled by the frame mean.  Parameters are one flat float32
 * vector: per layer W[Co][9][Ci] then bias[Co], layers conv0, (a, b) per
 * block, head. */
typedef struct flf_net flf_net;
int flf_net_create(int device, int nnum, int nz, int width, int nblocks, flf_net** out);
void flf_net_destroy(flf_net* net);
const char* flf_net_last_error(const flf_net* net);
long long flf_net_num_params(const flf_net* net);
int flf_net_set_params(flf_net* net, const float* host);   /* also resets the Adam state */
int flf_net_get_params(flf_net* net, float* host);
int flf_net_get_grads(flf_net* net, float* host);          /* gradient of the last train step */
/* img: device [T][Ly*nnum][Lx*nnum]; out: device [T][nz][Ly][Lx] (volume units) */
int flf_net_forward(flf_net* net, const float* img, int T, int Ly, int Lx, float* out, void* stream);
/* One Adam step (lr > 0) or gradient only (lr == 0) on the MSE between the
 * normalised output and target [T][nz][Ly][Lx]; loss_host may be NULL. */
int flf_net_train_step(flf_net* net, const float* img, const float* target, int T, int Ly, int Lx, float lr,
                       double* loss_host, void* stream);
/* Nearest upsampling of [T][nz][Ly][Lx] to a volume [nz][T][Ly*nnum][Lx*nnum]. */
int flf_net_upsample(const float* g, int T, int nz, int Ly, int Lx, int nnum, float* vol, void* stream);
/* C[M][N] = A[M][K] * B[N][K]^T (bf16 in, fp32 out, K % 16 == 0): the MFMA
 * GEMM the network uses, exposed for its layout test. */
int flf_mfma_gemm_nt(const void* A, const void* B, float* C, int M, int N, int K, void* stream);

/* ---- Light-field PSF synthesis (GeneratePSF, generate_psf.py:63-798) ----
 * The wave-optics model of the reference's PSF generator in FP64 kernels,
 * called per plane by flfhip.psfgen (which holds the host-side bookkeeping of
 * the generator's grids and windows).  Arrays are device memory of the
 * current HIP device; errors via flf_last_error(NULL).  Each call
 * synchronises `stream`.  Optics in SI units, the GeneratePSF arguments. */
typedef struct flf_psf_optics {
    double M, NA, MLPitch, n, fml, lam;
    int Nnum, OSR;                 /* both odd */
} flf_psf_optics;

/* |Koi I0|^2 (the intensity of the Debye field) at the npts points x1 =
 * j * MLPitch / Nnum / OSR, x2 = 0 of the plane at depth p3: the profile
 * calcPSFFT computes to size the PSF (generate_psf.py:103-142). */
int flf_psf_line(const flf_psf_optics* o, double p3, int npts, double* out, void* stream);

/* The wave PSF of the plane at depth p3 on the oversampled L x L grid, L =
 * 2 * half * OSR + 1 (half = IMG_HALFWIDTH): Debye integrals on rows lo..C-1,
 * columns row..C-1 (C = (L+1)/2), completed by the mirror and rotations of
 * calcPSF_p (generate_psf.py:310-350).  pattern: L*L complex (interleaved
 * re, im doubles). */
int flf_psf_wave(const flf_psf_optics* o, double p3, int half, int lo, double* pattern, void* stream);

/* The light-field PSFs of every lenslet offset (a, b) of the quadrant for
 * one plane (ComputeLFpsf, generate_psf.py:533-567) from its wave PSF:
 * H[b][a][q][p] (h*h*s2*s2 doubles, the HMatrix quadrant layout, s2 = 2 *
 * Nnum * imgsize_ref + 1), unnormalised.  lo: the wave PSF's first nonzero
 * row (as passed to flf_psf_wave); the grid needs L <= 4095. */
int flf_psf_lightfield(const flf_psf_optics* o, const double* pattern, int half, int imgsize_ref, int lo,
                       double* H, void* stream);

/* One plane's normalisation (generate_psf.py:594-613): values below
 * clip * (the plane's maximum) become 0, then (per_psf) each of the h*h PSFs
 * is divided by its sum; out: h*h*s2*s2 float32. */
int flf_psf_normalise(const double* H, int h, int s2, double clip, int per_psf, float* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* FLF_HIP_H */
