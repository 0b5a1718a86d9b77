// Light-field PSF synthesis on the GPU: the wave-optics model of GeneratePSF
// (generate_psf.py:63-798, after Prevedel et al.) in FP64 HIP kernels.
//
//   stage 1  Debye integral of every point of one eighth of each plane's
//            oversampled wave PSF (generate_psf.py:264-308), composite
//            Gauss-Legendre quadrature, one wavefront per point; the other
//            seven eighths by the mirror and rotations of calcPSF_p
//            (generate_psf.py:310-350).
//   stage 2  for every lenslet offset (a, b) of the quadrant: the wave PSF
//            shifted by OSR*(a, b) lenslet pixels, times the microlens phase
//            array, propagated by the Fresnel transfer function over fml
//            (generate_psf.py:245-259, 533-567).  The transfer function is
//            separable and even, so the propagation is P f0 P^T with P the
//            circulant matrix of p = IDFT(transfer function): two circulant
//            passes restricted to the rows and columns that survive the
//            aperture, binning and crop.  Then |f|^2, OSR x OSR binning
//            (pixelBinning, generate_psf.py:480-513) and the crop to the
//            PSF window, written in the HMatrix quadrant layout.
//   stage 3  per plane: values below clip * plane max set to zero, then
//            each PSF divided by its sum (generate_psf.py:594-613).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/flf_hip.h"
#include "flf_devptr.h"

namespace {

constexpr int GL_N = 16;          // Gauss-Legendre nodes per panel

struct DebyeArgs {
    double k, alpha, M, fobj, lam, p3;
    double dx;                    // oversampled pixel pitch (pixelPitch / OSR)
    double gx[GL_N], gw[GL_N];    // nodes and weights on [-1, 1]
};

// sqrt(cos t) (1 + cos t) sin t J0(vF sin t) exp(-i uO2 aF sin^2(t/2)),
// the Debye integrand (light_field_integrands.pyx:12-39).
__device__ inline double2 debye_integrand(double t, double aF, double uO2, double vF) {
    double st, ct;
    sincos(t, &st, &ct);
    const double s2 = sin(0.5 * t);
    const double amp = sqrt(ct) * (1.0 + ct) * st * j0(st * vF);
    double sp, cp;
    sincos(-uO2 * s2 * s2 * aF, &sp, &cp);
    return make_double2(amp * cp, amp * sp);
}

// Koi * I0 at radius r (image-space distance from the axis), one wavefront:
// the lanes take the panels in turn; a fixed-order butterfly sum.
__device__ double2 point_field(const DebyeArgs& a, double r_img, int lane) {
    const double sa = sin(a.alpha), sh = sin(0.5 * a.alpha);
    const double aF = 1.0 / (sh * sh);
    const double v = a.k * (r_img / a.M) * sa;
    const double u = 4.0 * a.k * a.p3 * sh * sh;
    const double uO2 = 0.5 * u, vF = v / sa;
    // bound of the integrand's angular frequency: J0 argument rate vF cos t
    // plus the phase rate |uO2| aF sin(t)/2; panels of at most 2 rad of it
    const double w = vF + 0.5 * fabs(uO2) * aF;
    const int panels = 4 + (int)ceil(w * a.alpha * 0.5);
    const double hp = a.alpha / panels;
    double re = 0.0, im = 0.0;
    for (int p = lane; p < panels; p += 64) {
        const double c = (p + 0.5) * hp, half = 0.5 * hp;
        double pr = 0.0, pi = 0.0;
#pragma unroll 4
        for (int j = 0; j < GL_N; ++j) {
            const double2 f = debye_integrand(c + half * a.gx[j], aF, uO2, vF);
            pr = fma(a.gw[j], f.x, pr);
            pi = fma(a.gw[j], f.y, pi);
        }
        re = fma(half, pr, re);
        im = fma(half, pi, im);
    }
    for (int off = 32; off > 0; off >>= 1) {
        re += __shfl_xor(re, off);
        im += __shfl_xor(im, off);
    }
    // Koi = M / (fobj lam)^2 exp(-i u / (4 sin^2(alpha/2))) = ... exp(-i k p3)
    const double kmag = a.M / ((a.fobj * a.lam) * (a.fobj * a.lam));
    double sk, ck;
    sincos(-u * 0.25 * aF, &sk, &ck);
    return make_double2(kmag * (ck * re - sk * im), kmag * (ck * im + sk * re));
}

// |Koi I0|^2 at x1 = j dx, x2 = 0, j < n (calcPSFFT's line).  One wave per point.
__global__ __launch_bounds__(256) void k_psf_line(DebyeArgs a, int n, double* out) {
    const int lane = threadIdx.x & 63;
    const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (j >= n) return;
    const double2 f = point_field(a, j * a.dx, lane);
    if (lane == 0) out[j] = f.x * f.x + f.y * f.y;
}

// Stage 1a: Debye values on the wedge rows lo..C-1, columns row..C-1 of the
// L x L grid (C = (L+1)/2, 1-based centre; 0-based index C-1 is the axis).
// Point (i, j) lies at image position (x[i], x[j]), x[m] = (m - (C-1)) dx.
// wedge: [C][C] complex, zero elsewhere.  One wave per point.
__global__ __launch_bounds__(256) void k_psf_wedge(DebyeArgs a, int C, int lo, const int2* pts, int npts,
                                                    double2* wedge) {
    const int lane = threadIdx.x & 63;
    const int q = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (q >= npts) return;
    const int i = pts[q].x, j = pts[q].y;
    const double xi = (double)(i - (C - 1)) * a.dx, xj = (double)(j - (C - 1)) * a.dx;
    const double2 f = point_field(a, sqrt(xi * xi + xj * xj), lane);
    if (lane == 0) wedge[(long long)i * C + j] = f;
}

// p0 of calcPSF_p: the wedge on the top rows, mirrored about the centre
// column (np.fliplr of the top-left quadrant), zero below row C-1.
__device__ inline double2 p0_at(const double2* wedge, int C, int L, int lo, int i, int j) {
    if (i >= C) return make_double2(0.0, 0.0);
    const int jj = j < C - 1 ? j : 2 * (C - 1) - j;
    (void)L;
    if (i < lo || jj < i) return make_double2(0.0, 0.0);
    return wedge[(long long)i * C + jj];
}

__device__ inline bool is_zero(double2 v) { return v.x == 0.0 && v.y == 0.0; }

// Stage 1b: the full wave PSF: p0, else its rotation by -90, -180, -270
// degrees (np.rot90(p0, k)[i][j] = p0[L-1-j][i], p0[L-1-i][L-1-j],
// p0[j][L-1-i]), taken in that order where still zero.
__global__ void k_psf_fill(const double2* wedge, int C, int L, int lo, double2* pat) {
    const long long n = (long long)L * L;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(e / L), j = (int)(e - (long long)i * L);
        double2 v = p0_at(wedge, C, L, lo, i, j);
        if (is_zero(v)) v = p0_at(wedge, C, L, lo, L - 1 - j, i);
        if (is_zero(v)) v = p0_at(wedge, C, L, lo, L - 1 - i, L - 1 - j);
        if (is_zero(v)) v = p0_at(wedge, C, L, lo, j, L - 1 - i);
        pat[e] = v;
    }
}

// Microlens phase array (calcML, generate_psf.py:169-205): every pixel takes
// the lens of its nearest lenslet centre (centres every K = Nnum*OSR pixels
// from index 0), exp(-i k0 (d1^2 + d2^2) dx^2 / (2 fml)), d = offset from it.
__global__ void k_ml(int L, int K, double phase, double2* ml) {
    const long long n = (long long)L * L;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(e / L), j = (int)(e - (long long)i * L);
        const int hk = (K - 1) / 2;
        const int d1 = (i + hk) % K - hk, d2 = (j + hk) % K - hk;
        double s, c;
        sincos(-phase * (double)(d1 * d1 + d2 * d2), &s, &c);
        ml[e] = make_double2(c, s);
    }
}

// p[m] = (1/L) sum_j exp(-i 2 pi^2 u_j^2 z / k) exp(+2 pi i j m / L), u_j the
// FFT frequency of bin j (j or j - L) / (L dx): the first column of the
// circulant 1-D Fresnel propagator (fresnel2D, generate_psf.py:245-259).
__global__ void k_prop_kernel(int L, double coef, double2* p) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= L) return;
    double re = 0.0, im = 0.0;
    for (int j = 0; j < L; ++j) {
        const int f = j < (L + 1) / 2 ? j : j - L;          // np.fft frequency order
        const double ph = -coef * (double)f * (double)f;
        const long long jm = ((long long)j * m) % L;
        double s1, c1, s2, c2;
        sincos(ph, &s1, &c1);
        sincospi(2.0 * (double)jm / (double)L, &s2, &c2);
        re += c1 * c2 - s1 * s2;
        im += c1 * s2 + s1 * c2;
    }
    p[m] = make_double2(re / L, im / L);
}

struct LfGeom {
    int L, h, OSR, s2;
    int lo_pat, hi_pat;           // rows/cols of the wave PSF that can be nonzero: [lo_pat, hi_pat)
    int ap_lo, ap_hi;             // aperture window (inclusive) in the shifted-back frame
    int crop0;                    // first oversampled index of binned pixel 0 (pixelBinning)
    int cp0;                      // first binned index kept (CP_p[0] - shift)
};

__device__ inline double2 cmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }

// Stage 2, pass 1 (along axis 0): G[z][m][n] = sum_k p[(R0 + m + sa - k) mod L]
// f0[k][n], f0[k][n] = pat[k - sa][n - sb] ml[k][n], for the rows m < nR
// that survive (R0 + m in the shifted-back frame) and every column n.
// Thread: RB consecutive rows m for one column n; lanes along n.
constexpr int RB = 4;
__global__ __launch_bounds__(256) void k_prop_rows(const double2* pat, const double2* ml, const double2* p, LfGeom g,
                                                    int R0, int nR, const int2* shifts, double2* G) {
    extern __shared__ double2 sp[];
    const int L = g.L;
    for (int i = threadIdx.x; i < L; i += blockDim.x) sp[i] = p[i];
    __syncthreads();
    const int z = blockIdx.z;
    const int sa = shifts[z].x, sb = shifts[z].y;
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    const int m0 = blockIdx.y * RB;
    if (n >= L || m0 >= nR) return;
    const int ns = n - sb;
    double2 acc[RB];
#pragma unroll
    for (int r = 0; r < RB; ++r) acc[r] = make_double2(0.0, 0.0);
    if (ns >= g.lo_pat && ns < g.hi_pat) {
        const int k_lo = max(g.lo_pat + sa, 0), k_hi = min(g.hi_pat + sa, L);
        for (int k = k_lo; k < k_hi; ++k) {
            const double2 f = cmul(pat[(long long)(k - sa) * L + ns], ml[(long long)k * L + n]);
            int idx = R0 + m0 + sa - k;
            idx %= L;
            if (idx < 0) idx += L;
#pragma unroll
            for (int r = 0; r < RB; ++r) {
                const double2 w = sp[idx];
                acc[r].x = fma(w.x, f.x, fma(-w.y, f.y, acc[r].x));
                acc[r].y = fma(w.x, f.y, fma(w.y, f.x, acc[r].y));
                idx = idx + 1 == L ? 0 : idx + 1;
            }
        }
    }
#pragma unroll
    for (int r = 0; r < RB; ++r)
        if (m0 + r < nR) G[((long long)z * nR + m0 + r) * L + n] = acc[r];
}

// Stage 2, pass 2 (along axis 1): F[z][m][c] = sum_n p[(C0 + c + sb - n) mod
// L] G[z][m][n] for the columns c < nC that survive.  Thread: RB columns for
// one row m; lanes along c.  The |.|^2 of F is stored (the propagation's
// constant phase exp(i k z) drops out).
__global__ __launch_bounds__(256) void k_prop_cols(const double2* G, const double2* p, LfGeom g, int nR, int C0,
                                                    int nC, const int2* shifts, double* I) {
    extern __shared__ double2 sp[];
    const int L = g.L;
    for (int i = threadIdx.x; i < L; i += blockDim.x) sp[i] = p[i];
    __syncthreads();
    const int z = blockIdx.z;
    const int sb = shifts[z].y;
    const int c0 = (blockIdx.x * blockDim.x + threadIdx.x) * RB;
    const int m = blockIdx.y;
    if (c0 >= nC || m >= nR) return;
    const double2* row = G + ((long long)z * nR + m) * L;
    double2 acc[RB];
#pragma unroll
    for (int r = 0; r < RB; ++r) acc[r] = make_double2(0.0, 0.0);
    for (int n = 0; n < L; ++n) {
        const double2 f = row[n];
        int idx = C0 + c0 + sb - n;
        idx %= L;
        if (idx < 0) idx += L;
#pragma unroll
        for (int r = 0; r < RB; ++r) {
            const double2 w = sp[idx];
            acc[r].x = fma(w.x, f.x, fma(-w.y, f.y, acc[r].x));
            acc[r].y = fma(w.x, f.y, fma(w.y, f.x, acc[r].y));
            idx = idx + 1 == L ? 0 : idx + 1;
        }
    }
#pragma unroll
    for (int r = 0; r < RB; ++r)
        if (c0 + r < nC) I[((long long)z * nR + m) * nC + c0 + r] = acc[r].x * acc[r].x + acc[r].y * acc[r].y;
}

// Stage 2 epilogue: binned and cropped intensity of offset z = a * h + b.
// Oversampled index r (both axes) of binned pixel t is crop0 + OSR t + u;
// values outside the aperture [ap_lo, ap_hi] or outside the grid after the
// shift back are zero.  I holds rows R0.. and columns C0.. of the
// shifted-back frame.  Output: Hcc[b][a][q][p] = binned[p][q] (axis 0 of the
// grid is p), double.
__global__ void k_bin(const double* I, LfGeom g, int R0, int nR, int C0, int nC, const int2* shifts, double* H) {
    const int z = blockIdx.y;
    const int a = z / g.h, b = z - a * g.h;
    const int sa = shifts[z].x, sb = shifts[z].y;
    const int n = g.s2 * g.s2;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
        const int pp = e / g.s2, qq = e - pp * g.s2;
        double s = 0.0;
        for (int u = 0; u < g.OSR; ++u) {
            const int r = g.crop0 + g.OSR * (g.cp0 + pp) + u;
            if (r < g.ap_lo || r > g.ap_hi || r + sa < 0 || r + sa >= g.L) continue;
            for (int v = 0; v < g.OSR; ++v) {
                const int c = g.crop0 + g.OSR * (g.cp0 + qq) + v;
                if (c < g.ap_lo || c > g.ap_hi || c + sb < 0 || c + sb >= g.L) continue;
                s += I[((long long)z * nR + (r - R0)) * nC + (c - C0)];
            }
        }
        H[(((long long)b * g.h + a) * g.s2 + qq) * g.s2 + pp] = s;
    }
}

// Stage 3: per plane, clip below clip * max and divide each PSF by its sum
// (in FP64), stored as float32.  One workgroup per PSF for the sums; the
// plane maximum comes first from one workgroup.
__global__ __launch_bounds__(256) void k_plane_max(const double* H, long long n, double* out) {
    __shared__ double s[256];
    double m = 0.0;
    for (long long e = threadIdx.x; e < n; e += blockDim.x) m = fmax(m, H[e]);
    s[threadIdx.x] = m;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + k]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = s[0];
}

__global__ __launch_bounds__(256) void k_psf_normalise(const double* H, int per, const double* pmax, double clip,
                                                        int per_psf, float* out) {
    __shared__ double s[256];
    const double* h = H + (long long)blockIdx.x * per;
    const double thr = clip * pmax[0];
    double sum = 0.0;
    for (int e = threadIdx.x; e < per; e += blockDim.x) sum += h[e] < thr ? 0.0 : h[e];
    s[threadIdx.x] = sum;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) s[threadIdx.x] += s[threadIdx.x + k];
        __syncthreads();
    }
    const double tot = per_psf ? s[0] : 1.0;
    for (int e = threadIdx.x; e < per; e += blockDim.x)
        out[(long long)blockIdx.x * per + e] = (float)(h[e] < thr ? 0.0 : h[e] / tot);
}

// Gauss-Legendre nodes and weights on [-1, 1] (Newton on the Legendre
// recurrence, double precision).
void gauss_legendre(double* x, double* w) {
    const int n = GL_N;
    for (int i = 0; i < n; ++i) {
        double t = std::cos(3.14159265358979323846 * (i + 0.75) / (n + 0.5));
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = t;
            for (int k = 2; k <= n; ++k) {
                const double p2 = ((2.0 * k - 1.0) * t * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            const double dp = n * (t * p1 - p0) / (t * t - 1.0);
            const double dt = p1 / dp;
            t -= dt;
            if (std::fabs(dt) < 1e-16) {
                x[i] = t;
                w[i] = 2.0 / ((1.0 - t * t) * dp * dp);
                break;
            }
        }
    }
}

bool optics_ok(const flf_psf_optics* o) {
    return o && o->M > 0 && o->NA > 0 && o->n > 0 && o->NA < o->n && o->MLPitch > 0 && o->fml > 0 && o->lam > 0 &&
           o->Nnum >= 1 && o->Nnum % 2 == 1 && o->OSR >= 1 && o->OSR % 2 == 1;
}

DebyeArgs debye_args(const flf_psf_optics* o, double p3) {
    DebyeArgs a{};
    a.k = 2.0 * 3.14159265358979323846 * o->n / o->lam;
    a.alpha = std::asin(o->NA / o->n);
    a.M = o->M;
    a.fobj = 200e-3 / o->M;
    a.lam = o->lam;
    a.p3 = p3;
    a.dx = o->MLPitch / o->Nnum / o->OSR;
    gauss_legendre(a.gx, a.gw);
    return a;
}

int fail_tls(int code, const std::string& msg) {
    flf::tls_error() = msg;
    return code;
}

#define PCHK(call)                                                                              \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess) return fail_tls(FLF_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// The oversampled grid of the generator (generate_psf.py:207-221):
// L = 2 * half * OSR + 1 points, half = IMG_HALFWIDTH.
int grid_len(int half, int OSR) { return 2 * half * OSR + 1; }

}  // namespace

extern "C" {

int flf_psf_line(const flf_psf_optics* o, double p3, int npts, double* out, void* stream) {
    flf::tls_error().clear();
    if (!optics_ok(o) || npts < 1) return fail_tls(FLF_EINVAL, "flf_psf_line: bad optics or point count");
    int dev = 0;
    PCHK(hipGetDevice(&dev));
    std::string why;
    if (!flf::device_span(dev, out, (size_t)npts * sizeof(double), "out", why)) return fail_tls(FLF_EINVAL, why);
    const DebyeArgs a = debye_args(o, p3);
    hipLaunchKernelGGL(k_psf_line, dim3((npts + 3) / 4), dim3(256), 0, (hipStream_t)stream, a, npts, out);
    PCHK(hipGetLastError());
    return FLF_OK;
}

int flf_psf_wave(const flf_psf_optics* o, double p3, int half, int lo, double* pattern, void* stream) {
    flf::tls_error().clear();
    if (!optics_ok(o) || half < 1) return fail_tls(FLF_EINVAL, "flf_psf_wave: bad optics or grid");
    const int L = grid_len(half, o->OSR), C = (L + 1) / 2;
    if (lo < 0 || lo >= C) return fail_tls(FLF_EINVAL, "flf_psf_wave: first wedge row outside 0..centre");
    int dev = 0;
    PCHK(hipGetDevice(&dev));
    std::string why;
    if (!flf::device_span(dev, pattern, (size_t)L * L * 2 * sizeof(double), "pattern", why))
        return fail_tls(FLF_EINVAL, why);
    hipStream_t st = (hipStream_t)stream;
    std::vector<int2> pts;
    for (int i = lo; i < C; ++i)
        for (int j = i; j < C; ++j) pts.push_back(make_int2(i, j));
    int2* dpts = nullptr;
    double2* wedge = nullptr;
    auto cleanup = [&]() {
        if (dpts) (void)hipFree(dpts);
        if (wedge) (void)hipFree(wedge);
    };
    if (hipMalloc((void**)&dpts, pts.size() * sizeof(int2)) != hipSuccess ||
        hipMalloc((void**)&wedge, (size_t)C * C * sizeof(double2)) != hipSuccess) {
        (void)hipGetLastError();
        cleanup();
        return fail_tls(FLF_ENOMEM, "flf_psf_wave: device allocation failed");
    }
    hipError_t e = hipMemcpy(dpts, pts.data(), pts.size() * sizeof(int2), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemsetAsync(wedge, 0, (size_t)C * C * sizeof(double2), st);
    const DebyeArgs a = debye_args(o, p3);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_psf_wedge, dim3((unsigned)((pts.size() + 3) / 4)), dim3(256), 0, st, a, C, lo, dpts,
                           (int)pts.size(), wedge);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_psf_fill, dim3(1024), dim3(256), 0, st, wedge, C, L, lo, (double2*)pattern);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    cleanup();
    if (e != hipSuccess) return fail_tls(FLF_EHIP, std::string("flf_psf_wave: ") + hipGetErrorString(e));
    return FLF_OK;
}

int flf_psf_lightfield(const flf_psf_optics* o, const double* pattern, int half, int imgsize_ref, int lo,
                       double* H, void* stream) {
    flf::tls_error().clear();
    if (!optics_ok(o) || half < 1 || imgsize_ref < 0) return fail_tls(FLF_EINVAL, "flf_psf_lightfield: bad arguments");
    const int OSR = o->OSR, N = o->Nnum, h = (N + 1) / 2;
    const int L = grid_len(half, OSR), C = (L + 1) / 2;
    if (lo < 0 || lo >= C) return fail_tls(FLF_EINVAL, "flf_psf_lightfield: first wedge row outside 0..centre");
    if (L > 4095) return fail_tls(FLF_EINVAL, "flf_psf_lightfield: grid longer than 4095 points");
    // crop windows (generate_psf.py:352-355, 517-519, 553-566, 480-513)
    const int hw = N * imgsize_ref * OSR;
    const int s2 = 2 * N * imgsize_ref + 1;                // len(CP_p)
    const int c0 = C - 1;
    LfGeom g{};
    g.L = L; g.h = h; g.OSR = OSR; g.s2 = s2;
    g.lo_pat = lo;                                         // wave PSF support: rows/cols lo..L-1-lo
    g.hi_pat = L - lo;
    g.ap_lo = std::max(c0 - hw, 0);
    g.ap_hi = std::min(c0 + hw, L - 1);
    {
        const int xc = (L - 1) / 2 + 1, xci = xc - (OSR - 1) / 2;
        int init = xci - (xci / OSR) * OSR, shift = 0;
        if (init < 1) { init += OSR; shift = 1; }
        g.crop0 = init - 1;
        g.cp0 = (int)((double)c0 / OSR - (double)hw / OSR) - shift;    // CP_p[0] - shift
    }
    if (g.cp0 < 0) return fail_tls(FLF_EINVAL, "flf_psf_lightfield: PSF window outside the grid");
    // rows (and columns) of the shifted-back frame the binning reads
    const int R0 = g.crop0 + OSR * g.cp0, nR = OSR * s2;
    if (R0 < 0 || R0 + nR > L) return fail_tls(FLF_EINVAL, "flf_psf_lightfield: binning window outside the grid");
    int dev = 0;
    PCHK(hipGetDevice(&dev));
    std::string why;
    if (!flf::device_span(dev, pattern, (size_t)L * L * 2 * sizeof(double), "pattern", why) ||
        !flf::device_span(dev, H, (size_t)h * h * s2 * s2 * sizeof(double), "H", why))
        return fail_tls(FLF_EINVAL, why);
    hipStream_t st = (hipStream_t)stream;
    const int nab = h * h;
    std::vector<int2> sh(nab);
    for (int a = 0; a < h; ++a)
        for (int b = 0; b < h; ++b) sh[a * h + b] = make_int2(OSR * (a + 1 - h), OSR * (b + 1 - h));
    int2* dsh = nullptr;
    double2 *ml = nullptr, *p = nullptr, *G = nullptr;
    double* I = nullptr;
    auto cleanup = [&]() {
        for (void* q : {(void*)dsh, (void*)ml, (void*)p, (void*)G, (void*)I})
            if (q) (void)hipFree(q);
    };
    if (hipMalloc((void**)&dsh, nab * sizeof(int2)) != hipSuccess ||
        hipMalloc((void**)&ml, (size_t)L * L * sizeof(double2)) != hipSuccess ||
        hipMalloc((void**)&p, (size_t)L * sizeof(double2)) != hipSuccess ||
        hipMalloc((void**)&G, (size_t)nab * nR * L * sizeof(double2)) != hipSuccess ||
        hipMalloc((void**)&I, (size_t)nab * nR * nR * sizeof(double)) != hipSuccess) {
        (void)hipGetLastError();
        cleanup();
        return fail_tls(FLF_ENOMEM, "flf_psf_lightfield: device allocation failed");
    }
    hipError_t e = hipMemcpy(dsh, sh.data(), nab * sizeof(int2), hipMemcpyHostToDevice);
    const double pi = 3.14159265358979323846;
    const double dx = o->MLPitch / N / OSR;
    const double k0 = 2.0 * pi / o->lam;
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_ml, dim3(1024), dim3(256), 0, st, L, N * OSR, k0 / (2.0 * o->fml) * dx * dx, ml);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        // transfer function exp(-i 2 pi^2 u^2 z / k0), u = f / (L dx)
        const double coef = 2.0 * pi * pi * o->fml / k0 / ((double)L * dx * (double)L * dx);
        hipLaunchKernelGGL(k_prop_kernel, dim3((L + 255) / 256), dim3(256), 0, st, L, coef, p);
        e = hipGetLastError();
    }
    const size_t lds = (size_t)L * sizeof(double2);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_prop_rows, dim3((L + 255) / 256, (nR + RB - 1) / RB, nab), dim3(256), lds, st,
                           (const double2*)pattern, ml, p, g, R0, nR, dsh, G);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_prop_cols, dim3((nR / RB + 1 + 255) / 256, nR, nab), dim3(256), lds, st, G, p, g, nR, R0,
                           nR, dsh, I);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_bin, dim3((s2 * s2 + 255) / 256, nab), dim3(256), 0, st, I, g, R0, nR, R0, nR, dsh, H);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    cleanup();
    if (e != hipSuccess) return fail_tls(FLF_EHIP, std::string("flf_psf_lightfield: ") + hipGetErrorString(e));
    return FLF_OK;
}

int flf_psf_normalise(const double* H, int h, int s2, double clip, int per_psf, float* out, void* stream) {
    flf::tls_error().clear();
    if (h < 1 || s2 < 1 || clip < 0) return fail_tls(FLF_EINVAL, "flf_psf_normalise: bad arguments");
    const long long n = (long long)h * h * s2 * s2;
    int dev = 0;
    PCHK(hipGetDevice(&dev));
    std::string why;
    if (!flf::device_span(dev, H, (size_t)n * sizeof(double), "H", why) ||
        !flf::device_span(dev, out, (size_t)n * sizeof(float), "out", why))
        return fail_tls(FLF_EINVAL, why);
    hipStream_t st = (hipStream_t)stream;
    double* pm = nullptr;
    if (hipMalloc((void**)&pm, sizeof(double)) != hipSuccess) {
        (void)hipGetLastError();
        return fail_tls(FLF_ENOMEM, "flf_psf_normalise: device allocation failed");
    }
    hipLaunchKernelGGL(k_plane_max, dim3(1), dim3(256), 0, st, H, n, pm);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_psf_normalise, dim3(h * h), dim3(256), 0, st, H, s2 * s2, pm, clip, per_psf, out);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(pm);
    if (e != hipSuccess) return fail_tls(FLF_EHIP, std::string("flf_psf_normalise: ") + hipGetErrorString(e));
    return FLF_OK;
}

}  // extern "C"
