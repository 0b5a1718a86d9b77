// C ABI of libflf_hip: context, planning, PSF spectrum cache, projections and
// the fused Richardson-Lucy update.  Declarations: include/flf_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/flf_hip.h"
#include "flf_devptr.h"
#include "flf_kernels.h"
#include "flf_plan.h"

using namespace flf;

namespace {

struct Group {
    GridTables gt;
    std::vector<int> planes;          // global plane ids, ascending
    RadixPlan rpY, rpX, rpYr, rpXr;   // fY, fX (block-level PSF transforms), rY, rX (small FFT)
    RadixPlan rpYw, rpXw, rpXhw;      // fY, fX, fX/2 for the wave-level inverse transforms (wide radices)
    int RB = 1, CBp = 1, CBi = 1;     // FFT batching (psf rows, psf cols, ifft cols)
    ChunkDesc* d_chunks = nullptr;
    ItemDesc* d_items = nullptr;
    TileDesc* d_tiles = nullptr;
    int* d_bu = nullptr;
    int* d_au = nullptr;
    int* d_uoff = nullptr;
    int* d_s2 = nullptr;
    std::vector<unsigned char> enable_host;   // planes of the current projection (kernel-argument mask)
    float2* d_G[2] = {nullptr, nullptr};
    long long* d_srcoff_fwd = nullptr;   // active frame count's tables (point into the *_T lists)
    long long* d_Soff_fwd = nullptr;
    std::vector<long long*> srcoff_T, soff_T;    // [T - 1]: tables for every frame count <= Tcap
    std::vector<long long> Tbwd_off_T;
    float2* d_S_fwd = nullptr;
    float2* d_S_bwd = nullptr;
    float2* d_A_fwd = nullptr;
    float2* d_A_bwd = nullptr;
    float2* d_T_fwd = nullptr;
    float2* d_tw = nullptr;           // twiddle tables: fY | fX | fX/2 | rY | rX
    const float2 *twY = nullptr, *twX = nullptr, *twXh = nullptr, *twRY = nullptr, *twRX = nullptr;
    RadixPlan rpXh{};                 // fX/2 (even fX)
    long long Tbwd_off = 0;           // offset of this group's planes in the shared backward buffer
    std::vector<char> have[2];
};

}  // namespace

struct flf_ctx {
    int device = 0;
    std::string err;
    bool planned = false;
    int N = 0, Y = 0, X = 0, T = 0, nz = 0;
    int Tcap = 0;                     // frames the buffers hold; T <= Tcap is the active count
    std::vector<RowPlaneDesc*> rowdesc_T;
    std::vector<int> s2, group_of, idx_in_group;
    std::vector<Group> groups;
    RowPlaneDesc* d_rowdesc = nullptr;
    // Ragged batches: per active frame the valid (Y, X) inside the planned
    // canvas (flf_set_frame_dims); null pointer in the kernels = uniform.
    int2* d_fdims = nullptr;
    std::vector<int2> fdims_host;
    bool ragged = false;
    float2* d_T_bwd = nullptr;
    long long* d_zeros = nullptr;     // zero offsets (length max planes per group)
    double* d_fitpart = nullptr;
    double* d_fitout = nullptr;
    size_t max_lds = 65536;
    int ipt = 2;                      // MAC items per lane pair's thread: 2 = lane pairs share two items
                                      // and split the spectra (FLF_ITEMS_PER_THREAD=1 selects one item
                                      // per thread); plan time
    int radix_order = 0;              // pass order of every FFT plan (make_radix_plan); plan time
    int fb_max = 2;                   // frames per MAC workgroup at T >= 2 (FLF_MAC_FRAMES overrides)
    int fwd_split = 0;                // forward plane ranges, 0 = automatic (FLF_FWD_SPLIT overrides)
    int bwd_planes = 1;               // planes per backward MAC workgroup
    int bwd_fgroup = 0;               // backward MAC frame blocks dispatched side by side (0: all)
    bool mfma = false;                // matrix-core MAC (k_mac_tc, 8 frames per workgroup) for calls of >= 2
                                      // frames (option mac_mfma, FLF_MAC_MFMA=1; chosen by the autotuner)
    // Fourier-grid groups of one projection side by side on streams of
    // their own (default priority, non-blocking), forked from and joined
    // back to the caller's stream by events; option "group_streams", off by
    // default (at 1455^2 x 41 planes: two-frame forward -1.8%, forward,
    // backward and RL iteration unchanged; the MAC launches already fill
    // the GPU).
    bool group_streams = false;
    // Grouped launches (option "grouped_launch", on by default): the small
    // FFTs, the MAC and the column transforms of all active grid groups of a
    // projection each go out as one dispatch, every group's arguments by
    // value in the kernel arguments.
    bool grouped = true;
    std::vector<hipStream_t> gstreams;
    std::vector<hipEvent_t> gevents;  // [0] fork, [1 + i] join of gstreams[i]
    long long bytes = 0;
    std::vector<void*> allocs;
    // Optional timing of the multiply-accumulate launches (flf_profile_*).
    bool prof = false;
    std::vector<hipEvent_t> ev_pool;
    struct ProfRec { hipEvent_t a, b; double bytes; };
    std::vector<ProfRec> prof_recs;
};

// ---------------------------------------------------------------------------
// Scratch of flf_datafit_noalloc: the per-block partial sums of the data fit
// (2 doubles per block and frame), in a file-scope device array instead of a
// buffer the library allocates.  Each GPU holds its own copy, from module
// load until the process ends.  In host code `scratch` names a host-side
// stand-in, never the device memory: kernels get the address that
// hipGetSymbolAddress returns for the current device (noalloc_scratch).
// ---------------------------------------------------------------------------
#define FLF_NOALLOC_FIT_BLOCKS 256
#define FLF_NOALLOC_MAX_FRAMES 512
#define FLF_NOALLOC_MAX_DEVICES 64
static __device__ __align__(256) double scratch[2 * FLF_NOALLOC_FIT_BLOCKS * FLF_NOALLOC_MAX_FRAMES];

namespace {

std::mutex g_scratch_mu;
double* g_scratch_ptr[FLF_NOALLOC_MAX_DEVICES] = {};

// Device address of `scratch` on `device`, looked up once per device with
// that device current (the caller's current device is restored).
hipError_t noalloc_scratch(int device, double** out) {
    *out = nullptr;
    if (device < 0 || device >= FLF_NOALLOC_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device && (e = hipSetDevice(device)) != hipSuccess) return e;
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = (double*)p;
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

}  // namespace

namespace {

const int FIT_BLOCKS = 256;
static_assert(FIT_BLOCKS == FLF_NOALLOC_FIT_BLOCKS, "the no-allocation data fit shares flf_datafit's reduction");
const int FWD_SPLIT_MAX = 4;
const int GROUP_STREAMS_MAX = 3;
const int FFT_THREADS = 512;
const int FFT_MAXB = 16;

int fail(flf_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

#define HIPCHK(call)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return fail(ctx, FLF_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));        \
    } while (0)

template <class T>
int dalloc(flf_ctx* ctx, T** p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void**)p, count * sizeof(T));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return fail(ctx, FLF_ENOMEM, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed");
    }
    ctx->allocs.push_back((void*)*p);
    ctx->bytes += (long long)(count * sizeof(T));
    return FLF_OK;
}

template <class T>
int upload(flf_ctx* ctx, T** p, const std::vector<T>& v) {
    int rc = dalloc(ctx, p, v.size());
    if (rc) return rc;
    if (!v.empty()) HIPCHK(hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return FLF_OK;
}

void free_all(flf_ctx* ctx) {
    for (void* p : ctx->allocs) (void)hipFree(p);
    ctx->allocs.clear();
    ctx->groups.clear();
    ctx->rowdesc_T.clear();
    ctx->Tcap = 0;
    ctx->bytes = 0;
    ctx->planned = false;
    ctx->d_rowdesc = nullptr;
    ctx->d_fdims = nullptr;
    ctx->ragged = false;
    ctx->d_T_bwd = nullptr;
    ctx->d_zeros = nullptr;
    ctx->d_fitpart = nullptr;
    ctx->d_fitout = nullptr;
}

// exp(-2 pi i m / L), m < L, in double then rounded to float.
void append_twiddles(std::vector<float2>& v, int L) {
    for (int m = 0; m < L; ++m) {
        double ang = -2.0 * 3.14159265358979323846 * (double)m / (double)L;
        v.push_back(make_float2((float)std::cos(ang), (float)std::sin(ang)));
    }
}

int min_radix(const RadixPlan& rp) {
    int m = 5;
    for (int i = 0; i < rp.n; ++i) m = std::min(m, rp.r[i]);
    return rp.n ? m : 5;
}

// Largest batch of length-L transforms one FFT workgroup can hold.
int fft_batch_cap(const RadixPlan& rp, size_t max_lds) {
    long long by_regs = (long long)min_radix(rp) * FFT_MAXB * FFT_THREADS / rp.L;
    long long by_lds = (long long)(max_lds / (sizeof(float2) * (size_t)rp.L));
    return (int)std::max<long long>(0, std::min(by_regs, by_lds));
}

bool fft_fits(const RadixPlan& rp, int nbatch, size_t lds_bytes, size_t max_lds) {
    long long per_pass = (long long)nbatch * rp.L / min_radix(rp);
    return per_pass <= (long long)FFT_MAXB * FFT_THREADS && lds_bytes <= max_lds;
}

int resolve_planes(flf_ctx* ctx, const int* planes, int nplanes, std::vector<char>& sel) {
    sel.assign(ctx->nz, 0);
    if (!planes) {
        std::fill(sel.begin(), sel.end(), 1);
        return FLF_OK;
    }
    if (nplanes < 0) return fail(ctx, FLF_EINVAL, "negative plane count");
    for (int i = 0; i < nplanes; ++i) {
        if (planes[i] < 0 || planes[i] >= ctx->nz)
            return fail(ctx, FLF_EINVAL, "plane index " + std::to_string(planes[i]) + " out of range");
        sel[planes[i]] = 1;
    }
    return FLF_OK;
}

// Planes of group g selected for this projection (sel: per global plane);
// `any` tells whether the group has work.  The selection reaches the kernels
// as a by-value mask (plane_mask), never through device memory.
int set_enable(flf_ctx* ctx, Group& g, const std::vector<char>& sel, int dir, bool& any) {
    std::vector<unsigned char> en(g.planes.size());
    any = false;
    for (size_t p = 0; p < g.planes.size(); ++p) {
        en[p] = sel[g.planes[p]] ? 1 : 0;
        if (en[p]) {
            if (!g.have[dir][p])
                return fail(ctx, FLF_ESTATE, std::string("PSF ") + (dir ? "Ht" : "H") + " of plane " +
                                                 std::to_string(g.planes[p]) + " not uploaded");
            any = true;
        }
    }
    g.enable_host = en;
    return FLF_OK;
}

PlaneMask plane_mask(const Group& g) {
    PlaneMask m{};
    for (size_t p = 0; p < g.enable_host.size(); ++p)
        if (g.enable_host[p]) m.w[p >> 5] |= 1u << (p & 31);
    return m;
}

// Device memory checks of the caller's arrays (flf_devptr.h), before any
// launch.  Sizes in float elements.
int need_dev(flf_ctx* ctx, const void* p, long long nfloat, const char* what) {
    std::string why;
    if (!device_span(ctx->device, p, (size_t)nfloat * sizeof(float), what, why)) return fail(ctx, FLF_EINVAL, why);
    return FLF_OK;
}

long long volume_floats(const flf_ctx* ctx, int frames) { return (long long)ctx->nz * frames * ctx->Y * ctx->X; }
long long image_floats(const flf_ctx* ctx, int frames) { return (long long)frames * ctx->Y * ctx->X; }

// Streams for the Fourier-grid groups of one projection.  Every group reads
// and writes buffers of its own (small-FFT output, accumulators, its column
// transform output or its slice of d_T_bwd), so groups may run side by side:
// the caller's stream forks to the group streams after its queued work and
// joins them before the step that reads every group.  One stream (st) while
// st is being captured into a graph, with fewer than two active groups, or
// with the option off.
struct GroupFork {
    hipStream_t st = nullptr;
    bool on = false;
    std::vector<hipStream_t>* gs = nullptr;
    hipStream_t stream(int k) const { return on ? (*gs)[k % gs->size()] : st; }
};

int fork_groups(flf_ctx* ctx, hipStream_t st, int nactive, GroupFork& f) {
    f.st = st;
    f.on = false;
    f.gs = &ctx->gstreams;
    if (!ctx->group_streams || nactive < 2) return FLF_OK;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    HIPCHK(hipStreamIsCapturing(st, &cs));
    if (cs != hipStreamCaptureStatusNone) return FLF_OK;
    const size_t want = (size_t)std::min(nactive, GROUP_STREAMS_MAX);
    if (ctx->gstreams.size() < want || ctx->gevents.size() < want + 1) {
        // created on the context's device; the caller's current device is kept
        int cur = -1;
        HIPCHK(hipGetDevice(&cur));
        if (cur != ctx->device) HIPCHK(hipSetDevice(ctx->device));
        hipError_t e = hipSuccess;
        while (e == hipSuccess && ctx->gstreams.size() < want) {
            hipStream_t s = nullptr;
            e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
            if (e == hipSuccess) ctx->gstreams.push_back(s);
        }
        while (e == hipSuccess && ctx->gevents.size() < ctx->gstreams.size() + 1) {
            hipEvent_t ev = nullptr;
            e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
            if (e == hipSuccess) ctx->gevents.push_back(ev);
        }
        if (cur != ctx->device) HIPCHK(hipSetDevice(cur));
        HIPCHK(e);
    }
    HIPCHK(hipEventRecord(ctx->gevents[0], st));
    for (hipStream_t s : ctx->gstreams) HIPCHK(hipStreamWaitEvent(s, ctx->gevents[0], 0));
    f.on = true;
    return FLF_OK;
}

// Called on every path after fork_groups, errors included, so that st never
// runs ahead of work still queued on the group streams.
int join_groups(flf_ctx* ctx, const GroupFork& f) {
    if (!f.on) return FLF_OK;
    for (size_t i = 0; i < ctx->gstreams.size(); ++i) {
        HIPCHK(hipEventRecord(ctx->gevents[i + 1], ctx->gstreams[i]));
        HIPCHK(hipStreamWaitEvent(f.st, ctx->gevents[i + 1], 0));
    }
    return FLF_OK;
}

hipEvent_t take_event(flf_ctx* ctx) {
    if (!ctx->ev_pool.empty()) {
        hipEvent_t e = ctx->ev_pool.back();
        ctx->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Frames per MAC workgroup: each spectrum value read serves fb frames.
int mac_frames(const flf_ctx* ctx, const GridTables& gt) {
    int fb = std::max(1, std::min(ctx->fb_max, ctx->T));
    while (fb > 1 && mac_lds(gt.slots, gt.nu, ctx->N, fb) > ctx->max_lds) --fb;
    return fb;
}

// Plane ranges of the forward MAC: about 4096 workgroups (4 per slot of the
// 256 CUs x 4 resident workgroups), at most FWD_SPLIT_MAX ranges and one
// plane per range (FLF_FWD_SPLIT overrides).
int fwd_split(const flf_ctx* ctx, const Group& g, int nP) {
    const long long per = (long long)g.gt.chunks.size() * ((ctx->T + mac_frames(ctx, g.gt) - 1) / mac_frames(ctx, g.gt));
    int sp = (int)std::max<long long>(1, (4096 + per - 1) / std::max<long long>(per, 1));
    if (ctx->fwd_split > 0) sp = ctx->fwd_split;
    return std::max(1, std::min(std::min(sp, FWD_SPLIT_MAX), nP));
}

// Launch the multiply-accumulate kernel; when profiling, bracket it with
// events on the same stream and record its algorithmic bytes (the compulsory
// read of the cached spectra of the planes it covers).
int run_mac(flf_ctx* ctx, const MacArgs& m, const Group& g, int nplanes_grid, hipStream_t st) {
    const GridTables& gt = g.gt;
    hipEvent_t ea = nullptr, eb = nullptr;
    if (ctx->prof) {
        ea = take_event(ctx);
        eb = take_event(ctx);
        if (ea && eb) HIPCHK(hipEventRecord(ea, st));
    }
    // the matrix-core kernel (8 frames per workgroup) for calls of several frames
    const bool tc = ctx->mfma && m.T >= 2 && m.tiles && mac_tc_lds(gt.slots, m.N) <= ctx->max_lds;
    HIPCHK(launch_mac(m, gt.slots, gt.ipt, mac_frames(ctx, gt), (int)gt.chunks.size(), nplanes_grid, tc, st));
    if (ctx->prof && ea && eb) {
        HIPCHK(hipEventRecord(eb, st));
        int np = 0;
        for (unsigned char e : g.enable_host) np += e ? 1 : 0;
        double bytes = (double)np * gt.nu * gt.fY * gt.rfX * sizeof(float2);
        ctx->prof_recs.push_back({ea, eb, bytes});
    }
    return FLF_OK;
}

// Elements of one frame's small-FFT output: [chunk][N*N][NH], NH = slots
// (half of each orbit's residue list; the rest follows by conjugation).
long long s_frame_stride(const flf_ctx* ctx, const Group& g) {
    return (long long)g.gt.chunks.size() * ctx->N * ctx->N * g.gt.slots;
}

const int2* frame_dims(const flf_ctx* ctx) { return ctx->ragged ? ctx->d_fdims : nullptr; }

// Small FFTs of the backward source (the image) for one group.
int run_small_fft_image(flf_ctx* ctx, Group& g, const float* img, hipStream_t st) {
    SmallFFTArgs a{};
    a.src = img;
    a.src_off = ctx->d_zeros;
    a.src_tstride = (long long)ctx->Y * ctx->X;
    a.src_pitch = ctx->X;
    a.Y = ctx->Y; a.X = ctx->X; a.N = ctx->N;
    a.rpY = g.rpYr; a.rpX = g.rpXr;
    a.twY = g.twRY; a.twX = g.twRX;
    a.S = g.d_S_bwd;
    a.S_off = ctx->d_zeros;
    a.chunks = g.d_chunks; a.nchunks = (int)g.gt.chunks.size(); a.NH = g.gt.slots;
    a.fY = g.gt.fY; a.fX = g.gt.fX;
    a.S_tstride = s_frame_stride(ctx, g);
    a.fdims = frame_dims(ctx);
    HIPCHK(launch_small_fft(a, 1, ctx->T, st));
    return FLF_OK;
}

// The work of one grid group in one projection, as the kernels take it:
// its small FFTs (nsrc source planes), its MAC (nranges plane ranges) and
// its column transforms (nbatch transforms).
struct GroupWork {
    size_t gi;
    SmallFFTArgs sa;
    int nsrc;
    MacArgs ma;
    int nranges;
    IfftColArgs ca;
    int nbatch;
};

// One group's launches on st (the path of group streams, of a single group,
// and of graphs captured before their argument tables exist).
int launch_group_work(flf_ctx* ctx, const GroupWork& w, hipStream_t st) {
    Group& g = ctx->groups[w.gi];
    HIPCHK(launch_small_fft(w.sa, w.nsrc, ctx->T, st));
    int rc = run_mac(ctx, w.ma, g, w.nranges, st);
    if (rc) return rc;
    HIPCHK(launch_ifft_cols(w.ca, w.nbatch, st));
    return FLF_OK;
}

// All groups' stages as one dispatch each on st.  done = false when this is
// not possible (fewer than two groups, the matrix-core MAC, mixed item
// shapes, LDS): the caller then launches group by group.
int launch_grouped(flf_ctx* ctx, const std::vector<GroupWork>& ws, hipStream_t st, bool& done) {
    done = false;
    if (!ctx->grouped || ws.size() < 2 || ws.size() > FLF_MAX_GROUPS || ctx->mfma) return FLF_OK;
    const GridTables& g0 = ctx->groups[ws[0].gi].gt;
    int fb = 4, CB = 8;
    bool prime = false;
    for (const GroupWork& w : ws) {
        const Group& g = ctx->groups[w.gi];
        if (g.gt.slots != g0.slots || g.gt.ipt != g0.ipt) return FLF_OK;
        // the grouped kernels hold the register-held transforms only
        if (small_fft_big(g.gt.rY, g.gt.rX) || wave_big(g.gt.fY)) return FLF_OK;
        fb = std::min(fb, mac_frames(ctx, g.gt));
        CB = std::min(CB, g.CBi);
        prime = prime || plan_has_prime(w.ca.rpY);
    }
    SmallFFTGrouped sg{};
    MacGrouped mg{};
    IfftColGrouped cg{};
    sg.gg.ngroups = mg.gg.ngroups = cg.gg.ngroups = (int)ws.size();
    mg.nfb = (ctx->T + fb - 1) / fb;
    size_t lds_s = 0, lds_m = 0, lds_c = 0;
    double bytes = 0;
    for (size_t k = 0; k < ws.size(); ++k) {
        const GroupWork& w = ws[k];
        const Group& g = ctx->groups[w.gi];
        sg.a[k] = w.sa;
        sg.gg.start[k + 1] = sg.gg.start[k] + w.nsrc;
        sg.gg.nchunks[k] = sg.gg.nranges[k] = 1;
        lds_s = std::max(lds_s, small_fft_lds(w.sa.rpY.L, w.sa.rpX.L));
        mg.a[k] = w.ma;
        const int nch = (int)g.gt.chunks.size();
        mg.gg.start[k + 1] = mg.gg.start[k] + (nch * w.nranges * mg.nfb + 7) / 8 * 8;   // XCD-aligned (k_mac_grouped)
        mg.gg.nchunks[k] = nch;
        mg.gg.nranges[k] = w.nranges;
        lds_m = std::max(lds_m, mac_lds(g.gt.slots, g.gt.nu, ctx->N, fb));
        cg.a[k] = w.ca;
        cg.a[k].CB = CB;
        const int ncb = (w.ca.rfX + CB - 1) / CB;
        cg.gg.start[k + 1] = cg.gg.start[k] + ncb * w.nbatch;
        cg.gg.nchunks[k] = ncb;
        cg.gg.nranges[k] = w.nbatch;
        lds_c = std::max(lds_c, ifft_cols_lds(CB, w.ca.rpY.L));
        int np = 0;
        for (unsigned char e : g.enable_host) np += e ? 1 : 0;
        bytes += (double)np * g.gt.nu * g.gt.fY * g.gt.rfX * sizeof(float2);
    }
    if (lds_s > ctx->max_lds || lds_m > ctx->max_lds || lds_c > ctx->max_lds) return FLF_OK;
    HIPCHK(launch_small_fft_grouped(sg, ctx->N, ctx->T, lds_s, st));
    hipEvent_t ea = nullptr, eb = nullptr;
    if (ctx->prof) {
        ea = take_event(ctx);
        eb = take_event(ctx);
        if (ea && eb) HIPCHK(hipEventRecord(ea, st));
    }
    HIPCHK(launch_mac_grouped(mg, g0.slots, g0.ipt, fb, mg.nfb, lds_m, st));
    if (ctx->prof && ea && eb) {
        HIPCHK(hipEventRecord(eb, st));
        ctx->prof_recs.push_back({ea, eb, bytes});
    }
    HIPCHK(launch_ifft_cols_grouped(cg, CB, prime, lds_c, st));
    done = true;
    return FLF_OK;
}

// The groups' work of one projection: grouped dispatches on the caller's
// stream when possible, else per group on the group streams (fk).
int run_group_work(flf_ctx* ctx, const std::vector<GroupWork>& ws, const GroupFork& fk) {
    if (!fk.on) {
        bool done = false;
        int rc = launch_grouped(ctx, ws, fk.st, done);
        if (rc || done) return rc;
    }
    for (size_t k = 0; k < ws.size(); ++k) {
        int rc = launch_group_work(ctx, ws[k], fk.stream((int)k));
        if (rc) return rc;
    }
    return FLF_OK;
}

// Per active group: small FFTs of the image, the per-plane MAC and the
// column transforms into the group's slice of d_T_bwd.
int run_backward_groups(flf_ctx* ctx, const float* img, const std::vector<char>& active, const GroupFork& fk) {
    const int N = ctx->N, T = ctx->T;
    std::vector<GroupWork> ws;
    for (size_t gi = 0; gi < ctx->groups.size(); ++gi) {
        if (!active[gi]) continue;
        Group& g = ctx->groups[gi];
        const GridTables& gt = g.gt;
        GroupWork w{};
        w.gi = gi;
        SmallFFTArgs& a = w.sa;
        a.src = img;
        a.src_off = ctx->d_zeros;
        a.src_tstride = (long long)ctx->Y * ctx->X;
        a.src_pitch = ctx->X;
        a.Y = ctx->Y; a.X = ctx->X; a.N = N;
        a.rpY = g.rpYr; a.rpX = g.rpXr;
        a.twY = g.twRY; a.twX = g.twRX;
        a.S = g.d_S_bwd;
        a.S_off = ctx->d_zeros;
        a.chunks = g.d_chunks; a.nchunks = (int)gt.chunks.size(); a.NH = gt.slots;
        a.fY = gt.fY; a.fX = gt.fX;
        a.S_tstride = s_frame_stride(ctx, g);
        a.fdims = frame_dims(ctx);
        a.tsplit = T;
        w.nsrc = 1;
        MacArgs& m = w.ma;
        m.G = g.d_G[1];
        m.chunks = g.d_chunks; m.items = g.d_items; m.u_bu = g.d_bu; m.u_au = g.d_au; m.uoff = g.d_uoff;
        m.tiles = g.d_tiles;
        m.S = g.d_S_bwd; m.S_off = ctx->d_zeros;
        m.S_tstride = s_frame_stride(ctx, g);
        m.s2 = g.d_s2; m.enable = plane_mask(g);
        m.nP = (int)g.planes.size();
        m.A = g.d_A_bwd;
        m.A_tstride = (long long)gt.fY * gt.rfX;
        m.A_pstride = (long long)T * gt.fY * gt.rfX;
        m.N = N; m.fY = gt.fY; m.fX = gt.fX; m.rY = gt.rY; m.rX = gt.rX; m.rfX = gt.rfX; m.nu = gt.nu;
        m.T = T;
        m.rs = mac_row_stride(gt.slots, gt.nu);
        // Each workgroup takes bwd_planes consecutive planes (loads run on
        // across the plane boundary) and stores every plane's accumulator.
        m.per_plane = 1;
        m.fgroup = ctx->bwd_fgroup;
        const int bp = std::max(1, ctx->bwd_planes);
        w.nranges = (m.nP + bp - 1) / bp;
        IfftColArgs& ic = w.ca;
        ic.A = g.d_A_bwd;
        ic.A_bstride = (long long)gt.fY * gt.rfX;
        ic.nsum = 1;
        ic.rfX = gt.rfX; ic.CB = g.CBi; ic.Yout = ctx->Y;
        ic.rpY = g.rpYw;
        ic.twY = g.twY;
        ic.out = ctx->d_T_bwd + g.Tbwd_off;
        ic.out_bstride = (long long)ctx->Y * gt.rfX;
        w.nbatch = m.nP * T;
        ws.push_back(w);
    }
    return run_group_work(ctx, ws, fk);
}

int run_backward(flf_ctx* ctx, const float* img, float* vol, const float* Htf, const float* trust, int mode,
                 const int* planes, int nplanes, hipStream_t st) {
    if (!ctx->planned) return fail(ctx, FLF_ESTATE, "flf_plan has not been called");
    if (!img || !vol) return fail(ctx, FLF_EINVAL, "null image or volume pointer");
    if (mode == 1 && !Htf) return fail(ctx, FLF_EINVAL, "Prevedel update needs Htf");
    int rc;
    if ((rc = need_dev(ctx, img, image_floats(ctx, ctx->T), "image"))) return rc;
    if ((rc = need_dev(ctx, vol, volume_floats(ctx, ctx->T), "volume"))) return rc;
    if (mode == 1 && (rc = need_dev(ctx, Htf, volume_floats(ctx, ctx->T), "Htf"))) return rc;
    if (mode == 2 && trust && (rc = need_dev(ctx, trust, volume_floats(ctx, 1), "trust"))) return rc;
    std::vector<char> sel;
    rc = resolve_planes(ctx, planes, nplanes, sel);
    if (rc) return rc;
    // enable flags of every group on st, then the groups side by side
    std::vector<char> active(ctx->groups.size(), 0);
    int nactive = 0;
    for (size_t gi = 0; gi < ctx->groups.size(); ++gi) {
        bool any = false;
        rc = set_enable(ctx, ctx->groups[gi], sel, 1, any);
        if (rc) return rc;
        active[gi] = any ? 1 : 0;
        nactive += any ? 1 : 0;
    }
    GroupFork fk;
    rc = fork_groups(ctx, st, nactive, fk);
    if (rc) return rc;
    rc = run_backward_groups(ctx, img, active, fk);
    const int rj = join_groups(ctx, fk);
    if (rc) return rc;
    if (rj) return rj;
    // every plane's row transform (and RL update) in one launch on st
    std::vector<int> ids;
    int wbuf = 0, prime = 0, big = 0;
    for (int z = 0; z < ctx->nz; ++z)
        if (sel[z]) {
            ids.push_back(z);
            const Group& gz = ctx->groups[ctx->group_of[z]];
            wbuf = std::max(wbuf, c2r_wave_buffer(gz.gt.fX));
            if (wave_big(gz.gt.fX)) big = 1;
            if (plan_has_prime(gz.rpXw) || plan_has_prime(gz.rpXhw)) prime = 1;
        }
    if (ids.empty()) return FLF_OK;
    IfftRowPlaneArgs r{};
    r.pd = ctx->d_rowdesc;
    for (size_t i = 0; i < ids.size(); ++i) r.plane_ids[i] = (unsigned short)ids[i];
    r.Tbuf = ctx->d_T_bwd;
    r.vol = vol; r.Htf = Htf; r.trust = trust;
    r.mode = mode; r.T = ctx->T; r.Y = ctx->Y; r.X = ctx->X;
    r.wbuf = wbuf;
    r.prime = prime;
    r.big = big;
    r.fdims = frame_dims(ctx);
    HIPCHK(launch_ifft_rows_plane(r, (int)ids.size(), st));
    return FLF_OK;
}

}  // namespace

extern "C" {

int flf_version(void) { return 1; }

int flf_create(int device, flf_ctx** out) {
    if (!out) return FLF_EINVAL;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return FLF_EHIP;
    if (device < 0 || device >= ndev) return FLF_EINVAL;
    flf_ctx* c = new flf_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess) { delete c; return FLF_EHIP; }
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && lds > 0)
        c->max_lds = (size_t)lds;
    if (configure_kernels(c->max_lds) != hipSuccess) c->max_lds = 65536;
    if (const char* e = getenv("FLF_ITEMS_PER_THREAD")) {
        int v = atoi(e);
        if (v == 1 || v == 2) c->ipt = v;
    }
    if (const char* e = getenv("FLF_MAC_MFMA")) c->mfma = atoi(e) != 0;
    if (const char* e = getenv("FLF_FWD_SPLIT")) {
        int v = atoi(e);
        if (v >= 1 && v <= 4) c->fwd_split = v;
    }
    if (const char* e = getenv("FLF_MAC_FRAMES")) {
        int v = atoi(e);
        if (v >= 1 && v <= 4) c->fb_max = v;
    }
    // set-up of the no-allocation data fit's scratch address on this device
    // (not required by the context: on failure flf_datafit_noalloc retries
    // and reports the error itself)
    double* sp = nullptr;
    if (noalloc_scratch(device, &sp) != hipSuccess) (void)hipGetLastError();
    *out = c;
    return FLF_OK;
}

void flf_destroy(flf_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    free_all(ctx);
    for (auto& r : ctx->prof_recs) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
    for (auto e : ctx->ev_pool) (void)hipEventDestroy(e);
    for (auto s : ctx->gstreams) (void)hipStreamSynchronize(s);
    for (auto s : ctx->gstreams) (void)hipStreamDestroy(s);
    for (auto e : ctx->gevents) (void)hipEventDestroy(e);
    delete ctx;
}

const char* flf_last_error(const flf_ctx* ctx) { return ctx ? ctx->err.c_str() : tls_error().c_str(); }

int flf_plan(flf_ctx* ctx, int nnum, int Y, int X, int T, int nz, const int* s2) {
    if (!ctx) return FLF_EINVAL;
    HIPCHK(hipSetDevice(ctx->device));
    free_all(ctx);
    ctx->err.clear();
    if (nnum < 3 || nnum % 2 == 0 || nnum > FLF_MAX_N) return fail(ctx, FLF_EINVAL, "Nnum must be odd, 3..31");
    if (Y < 1 || X < 1 || T < 1 || nz < 1 || !s2) return fail(ctx, FLF_EINVAL, "bad geometry");
    if (nz > FLF_MAX_PLANES) return fail(ctx, FLF_EINVAL, "more than " + std::to_string(FLF_MAX_PLANES) + " planes");
    if ((long long)nz * T > 65535) return fail(ctx, FLF_EINVAL, "planes x frames above 65535");
    for (int z = 0; z < nz; ++z)
        if (s2[z] < 1 || s2[z] % 2 == 0) return fail(ctx, FLF_EINVAL, "PSF side must be odd and positive");
    ctx->N = nnum; ctx->Y = Y; ctx->X = X; ctx->T = T; ctx->Tcap = T; ctx->nz = nz;
    ctx->s2.assign(s2, s2 + nz);
    ctx->group_of.assign(nz, -1);
    ctx->idx_in_group.assign(nz, -1);

    // Group planes by Fourier grid.
    for (int z = 0; z < nz; ++z) {
        int fY = fourier_len(Y, s2[z], nnum), fX = fourier_len(X, s2[z], nnum);
        if (fY > FLF_MAX_FOURIER || fX > FLF_MAX_FOURIER)
            return fail(ctx, FLF_EINVAL, "Fourier grid longer than " + std::to_string(FLF_MAX_FOURIER));
        int gi = -1;
        for (size_t i = 0; i < ctx->groups.size(); ++i)
            if (ctx->groups[i].gt.fY == fY && ctx->groups[i].gt.fX == fX) gi = (int)i;
        if (gi < 0) {
            if ((int)ctx->groups.size() >= FLF_MAX_GROUPS) return fail(ctx, FLF_EINVAL, "too many distinct Fourier grids");
            ctx->groups.emplace_back();
            Group& g = ctx->groups.back();
            std::string e = build_grid_tables(nnum, fY, fX, g.gt, ctx->ipt);
            if (!e.empty()) return fail(ctx, FLF_EINVAL, e);
            const int ro = ctx->radix_order;
            if (!make_radix_plan(fY, g.rpY, false, ro) || !make_radix_plan(fX, g.rpX, false, ro) ||
                !make_radix_plan(g.gt.rY, g.rpYr, false, ro) || !make_radix_plan(g.gt.rX, g.rpXr, false, ro) ||
                !make_radix_plan(fY, g.rpYw, true, ro) || !make_radix_plan(fX, g.rpXw, true, ro) ||
                (fX % 2 == 0 &&
                 (!make_radix_plan(fX / 2, g.rpXh, false, ro) || !make_radix_plan(fX / 2, g.rpXhw, true, ro))))
                return fail(ctx, FLF_EINVAL, "Fourier grid length has a prime factor other than 2, 3, 5 and one odd "
                                             "prime up to 31");
            gi = (int)ctx->groups.size() - 1;
        }
        ctx->group_of[z] = gi;
        ctx->idx_in_group[z] = (int)ctx->groups[gi].planes.size();
        ctx->groups[gi].planes.push_back(z);
    }

    // Device tables and buffers.
    int maxP = 1;
    long long Tbwd_total = 0;
    std::vector<RowPlaneDesc> rowdesc(nz);
    for (auto& g : ctx->groups) {
        GridTables& gt = g.gt;
        const int nP = (int)g.planes.size();
        maxP = std::max(maxP, nP);
        // FFT workgroup shapes.
        g.RB = std::min(16, fft_batch_cap(g.rpX, ctx->max_lds));
        g.CBp = std::min(16, fft_batch_cap(g.rpY, ctx->max_lds));
        g.CBi = 8;
        while (g.CBi > 1 && ifft_cols_lds(g.CBi, gt.fY) > ctx->max_lds) --g.CBi;
        if (g.RB < 1 || g.CBp < 1 || ifft_cols_lds(g.CBi, gt.fY) > ctx->max_lds)
            return fail(ctx, FLF_EINVAL, "Fourier grid too large for the LDS FFT");
        // Wave-level transforms: up to 2048 points in registers, up to 4096
        // out of place (two LDS buffers); reduced grids up to 256.
        if (gt.fY > 4096 || gt.fX > 4096 || gt.rY > 256 || gt.rX > 256)
            return fail(ctx, FLF_EINVAL, "Fourier grid " + std::to_string(gt.fY) + "x" + std::to_string(gt.fX) +
                                             " exceeds the transform sizes this library supports (4096, reduced 256)");
        if (ifft_rows_plane_lds(c2r_wave_buffer(gt.fX), wave_big(gt.fX)) > ctx->max_lds)
            return fail(ctx, FLF_EINVAL, "image row too long for the row transform");
        if (small_fft_lds(gt.rY, gt.rX) > ctx->max_lds)
            return fail(ctx, FLF_EINVAL, "reduced grid too large for the small-FFT kernel");
        {
            std::vector<float2> tw;
            append_twiddles(tw, gt.fY);
            append_twiddles(tw, gt.fX);
            append_twiddles(tw, std::max(1, gt.fX / 2));
            append_twiddles(tw, gt.rY);
            append_twiddles(tw, gt.rX);
            int rc2 = upload(ctx, &g.d_tw, tw);
            if (rc2) return rc2;
            g.twY = g.d_tw;
            g.twX = g.twY + gt.fY;
            g.twXh = g.twX + gt.fX;
            g.twRY = g.twXh + std::max(1, gt.fX / 2);
            g.twRX = g.twRY + gt.rY;
        }
        int rc;
        if ((rc = upload(ctx, &g.d_chunks, gt.chunks))) return rc;
        if ((rc = upload(ctx, &g.d_items, gt.items))) return rc;
        if ((rc = upload(ctx, &g.d_tiles, gt.tiles))) return rc;
        if ((rc = upload(ctx, &g.d_bu, gt.u_bu))) return rc;
        if ((rc = upload(ctx, &g.d_au, gt.u_au))) return rc;
        if ((rc = upload(ctx, &g.d_uoff, gt.uoff))) return rc;
        std::vector<int> s2g;
        for (int p = 0; p < nP; ++p) s2g.push_back(s2[g.planes[p]]);
        if ((rc = upload(ctx, &g.d_s2, s2g))) return rc;
        // Volume-plane and small-FFT offsets depend on the frame count: one
        // table per count up to the capacity, selected by flf_set_frames.
        const long long Sper = (long long)T * s_frame_stride(ctx, g);
        for (int tf = 1; tf <= T; ++tf) {
            std::vector<long long> srcoff, soff;
            for (int p = 0; p < nP; ++p) {
                srcoff.push_back((long long)g.planes[p] * tf * Y * X);
                soff.push_back((long long)p * tf * s_frame_stride(ctx, g));
            }
            long long *ps = nullptr, *pso = nullptr;
            if ((rc = upload(ctx, &ps, srcoff))) return rc;
            if ((rc = upload(ctx, &pso, soff))) return rc;
            g.srcoff_T.push_back(ps);
            g.soff_T.push_back(pso);
        }
        g.d_srcoff_fwd = g.srcoff_T[T - 1];
        g.d_Soff_fwd = g.soff_T[T - 1];
        g.enable_host.assign(nP, 1);
        const long long items = (long long)gt.items.size();
        for (int d = 0; d < 2; ++d) {
            if ((rc = dalloc(ctx, &g.d_G[d], (size_t)items * gt.slots * gt.nu * nP))) return rc;
            g.have[d].assign(nP, 0);
        }
        if ((rc = dalloc(ctx, &g.d_S_fwd, (size_t)(nP * Sper)))) return rc;
        if ((rc = dalloc(ctx, &g.d_S_bwd, (size_t)Sper))) return rc;
        if ((rc = dalloc(ctx, &g.d_A_fwd, (size_t)FWD_SPLIT_MAX * T * gt.fY * gt.rfX))) return rc;
        if ((rc = dalloc(ctx, &g.d_A_bwd, (size_t)nP * T * gt.fY * gt.rfX))) return rc;
        if ((rc = dalloc(ctx, &g.d_T_fwd, (size_t)T * Y * gt.rfX))) return rc;
        // Offsets into the shared backward buffer scale with the frame count:
        // unit offsets here, multiplied per count below.
        g.Tbwd_off = Tbwd_total;
        for (int p = 0; p < nP; ++p) {
            RowPlaneDesc& rd = rowdesc[g.planes[p]];
            rd.T_off = Tbwd_total + (long long)p * Y * gt.rfX;
            rd.T_tstride = (long long)Y * gt.rfX;
            rd.rfX = gt.rfX;
            rd.scale = 1.0f / ((float)gt.fY * (float)gt.fX);
            rd.rp = g.rpXw;
            rd.rph = g.rpXhw;
            rd.twf = g.twX;
            rd.twh = g.twXh;
        }
        Tbwd_total += (long long)nP * Y * gt.rfX;
        // (rows of a grid up to 2048 points hold X < 2048 samples; longer
        // ones take the out-of-place kernels, which hold up to 4096)
        if (ifft_rows_sum_lds(gt.fX) > ctx->max_lds || X > 4096)
            return fail(ctx, FLF_EINVAL, "image row too long for the row transform");
    }
    int rc;
    if ((rc = dalloc(ctx, &ctx->d_T_bwd, (size_t)Tbwd_total * T))) return rc;
    for (auto& g : ctx->groups) {
        for (int tf = 1; tf <= T; ++tf) g.Tbwd_off_T.push_back(g.Tbwd_off * tf);
        g.Tbwd_off = g.Tbwd_off_T[T - 1];
    }
    for (int tf = 1; tf <= T; ++tf) {
        std::vector<RowPlaneDesc> rd = rowdesc;
        for (auto& r : rd) r.T_off *= tf;
        RowPlaneDesc* prd = nullptr;
        if ((rc = upload(ctx, &prd, rd))) return rc;
        ctx->rowdesc_T.push_back(prd);
    }
    ctx->d_rowdesc = ctx->rowdesc_T[T - 1];
    if ((rc = upload(ctx, &ctx->d_zeros, std::vector<long long>(maxP, 0)))) return rc;
    if ((rc = dalloc(ctx, &ctx->d_fitpart, (size_t)2 * FIT_BLOCKS * T))) return rc;
    if ((rc = dalloc(ctx, &ctx->d_fitout, (size_t)2 * T))) return rc;
    if ((rc = dalloc(ctx, &ctx->d_fdims, (size_t)T))) return rc;
    ctx->fdims_host.assign(T, make_int2(Y, X));
    ctx->ragged = false;
    ctx->planned = true;
    return FLF_OK;
}

int flf_set_frames(flf_ctx* ctx, int T) {
    if (!ctx) return FLF_EINVAL;
    if (!ctx->planned) return fail(ctx, FLF_ESTATE, "flf_plan has not been called");
    if (T < 1 || T > ctx->Tcap)
        return fail(ctx, FLF_EINVAL, "frame count " + std::to_string(T) + " outside 1.." + std::to_string(ctx->Tcap));
    ctx->T = T;
    for (auto& g : ctx->groups) {
        g.d_srcoff_fwd = g.srcoff_T[T - 1];
        g.d_Soff_fwd = g.soff_T[T - 1];
        g.Tbwd_off = g.Tbwd_off_T[T - 1];
    }
    ctx->d_rowdesc = ctx->rowdesc_T[T - 1];
    ctx->ragged = false;                  // frame sizes apply to one frame count
    return FLF_OK;
}

int flf_set_frame_dims(flf_ctx* ctx, const int* dims, int T, void* stream) {
    if (!ctx) return FLF_EINVAL;
    if (!ctx->planned) return fail(ctx, FLF_ESTATE, "flf_plan has not been called");
    if (T != ctx->T)
        return fail(ctx, FLF_EINVAL, "frame sizes for " + std::to_string(T) + " frames, " + std::to_string(ctx->T) +
                                         " active");
    if (!dims) {
        ctx->ragged = false;
        return FLF_OK;
    }
    std::vector<int2> d(T);
    bool full = true;
    for (int t = 0; t < T; ++t) {
        const int y = dims[2 * t], x = dims[2 * t + 1];
        if (y < 1 || y > ctx->Y || x < 1 || x > ctx->X)
            return fail(ctx, FLF_EINVAL, "frame " + std::to_string(t) + " size " + std::to_string(y) + " x " +
                                             std::to_string(x) + " outside 1.." + std::to_string(ctx->Y) + " x 1.." +
                                             std::to_string(ctx->X));
        d[t] = make_int2(y, x);
        full = full && y == ctx->Y && x == ctx->X;
    }
    if (full) {
        ctx->ragged = false;
        return FLF_OK;
    }
    // Ordered on the caller's stream after the work already queued (which may
    // still read the previous sizes); the stream is synchronised so that the
    // host copy can be rewritten by the next call.
    ctx->fdims_host.assign(d.begin(), d.end());
    HIPCHK(hipMemcpyAsync(ctx->d_fdims, ctx->fdims_host.data(), T * sizeof(int2), hipMemcpyHostToDevice,
                          (hipStream_t)stream));
    HIPCHK(hipStreamSynchronize((hipStream_t)stream));
    ctx->ragged = true;
    return FLF_OK;
}

int flf_frame_capacity(const flf_ctx* ctx) { return ctx && ctx->planned ? ctx->Tcap : 0; }

int flf_set_option(flf_ctx* ctx, const char* name, int value) {
    if (!ctx || !name) return FLF_EINVAL;
    const std::string n(name);
    if (n == "fwd_split") {
        if (value < 0 || value > FWD_SPLIT_MAX) return fail(ctx, FLF_EINVAL, "fwd_split must be 0..4");
        ctx->fwd_split = value;
    } else if (n == "mac_frames") {
        if (value < 1 || value > 4) return fail(ctx, FLF_EINVAL, "mac_frames must be 1..4");
        ctx->fb_max = value;
    } else if (n == "mac_mfma") {
        ctx->mfma = value != 0;
    } else if (n == "bwd_planes") {
        if (value < 1 || value > 8) return fail(ctx, FLF_EINVAL, "bwd_planes must be 1..8");
        ctx->bwd_planes = value;
    } else if (n == "bwd_frame_group") {
        if (value < 0 || value > 64) return fail(ctx, FLF_EINVAL, "bwd_frame_group must be 0..64");
        ctx->bwd_fgroup = value;
    } else if (n == "group_streams") {
        ctx->group_streams = value != 0;
    } else if (n == "grouped_launch") {
        ctx->grouped = value != 0;
    } else if (n == "radix_order") {             // takes effect at the next flf_plan
        if (value < 0 || value >= FLF_RADIX_ORDERS) return fail(ctx, FLF_EINVAL, "radix_order must be 0..3");
        ctx->radix_order = value;
    } else if (n == "mac_ipt") {                 // takes effect at the next flf_plan
        if (value != 1 && value != 2 && value != 4) return fail(ctx, FLF_EINVAL, "mac_ipt must be 1, 2 or 4");
        ctx->ipt = value;
    } else if (n == "ifft_col_batch") {
        if (!ctx->planned) return fail(ctx, FLF_ESTATE, "flf_plan has not been called");
        if (value < 1 || value > 8) return fail(ctx, FLF_EINVAL, "ifft_col_batch must be 1..8");
        for (auto& g : ctx->groups)
            if (ifft_cols_lds(value, g.gt.fY) > ctx->max_lds)
                return fail(ctx, FLF_EINVAL, "ifft_col_batch too large for the LDS at this Fourier grid");
        for (auto& g : ctx->groups) g.CBi = value;
    } else {
        return fail(ctx, FLF_EINVAL, "unknown option " + n);
    }
    return FLF_OK;
}

int flf_get_option(const flf_ctx* ctx, const char* name, int* value) {
    if (!ctx || !name || !value) return FLF_EINVAL;
    const std::string n(name);
    if (n == "fwd_split") *value = ctx->fwd_split;
    else if (n == "mac_frames") *value = ctx->fb_max;
    else if (n == "mac_mfma") *value = ctx->mfma ? 1 : 0;
    else if (n == "bwd_planes") *value = ctx->bwd_planes;
    else if (n == "bwd_frame_group") *value = ctx->bwd_fgroup;
    else if (n == "group_streams") *value = ctx->group_streams ? 1 : 0;
    else if (n == "grouped_launch") *value = ctx->grouped ? 1 : 0;
    else if (n == "radix_order") *value = ctx->radix_order;
    else if (n == "mac_ipt") *value = ctx->ipt;
    else if (n == "ifft_col_batch") *value = ctx->groups.empty() ? 0 : ctx->groups[0].CBi;
    else return FLF_EINVAL;
    return FLF_OK;
}

int flf_plane_fshape(const flf_ctx* ctx, int z, int* fY, int* fX) {
    if (!ctx || !ctx->planned || z < 0 || z >= ctx->nz || !fY || !fX) return FLF_EINVAL;
    const Group& g = ctx->groups[ctx->group_of[z]];
    *fY = g.gt.fY;
    *fX = g.gt.fX;
    return FLF_OK;
}

long long flf_device_bytes(const flf_ctx* ctx) { return ctx ? ctx->bytes : 0; }

int flf_set_psf(flf_ctx* ctx, int z, int dir, const float* quadrant_host) {
    if (!ctx) return FLF_EINVAL;
    if (!ctx->planned) return fail(ctx, FLF_ESTATE, "flf_plan has not been called");
    if (z < 0 || z >= ctx->nz || (dir != 0 && dir != 1) || !quadrant_host)
        return fail(ctx, FLF_EINVAL, "bad plane, direction or pointer");
    HIPCHK(hipSetDevice(ctx->device));
    Group& g = ctx->groups[ctx->group_of[z]];
    const GridTables& gt = g.gt;
    const int h = (ctx->N + 1) / 2, s2 = ctx->s2[z];
    const size_t qn = (size_t)h * h * s2 * s2;
    float* dQ = nullptr;
    float2 *dtmp = nullptr, *dspec = nullptr;
    auto cleanup = [&]() {
        if (dQ) (void)hipFree(dQ);
        if (dtmp) (void)hipFree(dtmp);
        if (dspec) (void)hipFree(dspec);
    };
    if (hipMalloc((void**)&dQ, qn * sizeof(float)) != hipSuccess ||
        hipMalloc((void**)&dtmp, (size_t)gt.nu * s2 * gt.rfX * sizeof(float2)) != hipSuccess ||
        hipMalloc((void**)&dspec, (size_t)gt.nu * gt.fY * gt.rfX * sizeof(float2)) != hipSuccess) {
        (void)hipGetLastError();
        cleanup();
        return fail(ctx, FLF_ENOMEM, "out of device memory computing PSF spectra");
    }
    hipError_t e = hipMemcpy(dQ, quadrant_host, qn * sizeof(float), hipMemcpyHostToDevice);
    PsfRowArgs ra{};
    ra.Q = dQ; ra.h = h; ra.s2 = s2; ra.RB = g.RB; ra.u_bu = g.d_bu; ra.u_au = g.d_au; ra.rpX = g.rpX; ra.tmp = dtmp;
    if (e == hipSuccess) e = launch_psf_rows(ra, gt.nu, 0);
    PsfColArgs ca{};
    ca.tmp = dtmp; ca.s2 = s2; ca.rfX = gt.rfX; ca.CB = g.CBp; ca.rpY = g.rpY; ca.spec = dspec;
    if (e == hipSuccess) e = launch_psf_cols(ca, gt.nu, 0);
    SpecPackArgs pa{};
    pa.spec = dspec; pa.chunks = g.d_chunks; pa.items = g.d_items; pa.u_bu = g.d_bu; pa.u_au = g.d_au;
    pa.fY = gt.fY; pa.rfX = gt.rfX; pa.N = ctx->N; pa.slots = gt.slots; pa.nu = gt.nu;
    pa.nP = (int)g.planes.size(); pa.p = ctx->idx_in_group[z]; pa.G = g.d_G[dir]; pa.s2 = s2;
    if (e == hipSuccess) e = launch_spec_pack(pa, (int)gt.chunks.size(), 0);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    cleanup();
    if (e != hipSuccess) return fail(ctx, FLF_EHIP, std::string("PSF spectra: ") + hipGetErrorString(e));
    g.have[dir][ctx->idx_in_group[z]] = 1;
    return FLF_OK;
}

namespace {
// Forward projection of the active frames; with vol2, frames t >= T/2 come
// from vol2 (frame t - T/2) and both sources hold T/2 frames.  img may be
// null when ratio is set (ratio = im / image).
int run_forward(flf_ctx* ctx, const float* vol, const float* vol2, float* img, const float* im, float* ratio,
                const int* planes, int nplanes, hipStream_t st);
}  // namespace

int flf_forward(flf_ctx* ctx, const float* vol, float* img, const int* planes, int nplanes, void* stream) {
    if (!ctx) return FLF_EINVAL;
    if (!ctx->planned) return fail(ctx, FLF_ESTATE, "flf_plan has not been called");
    int rc;
    if ((rc = need_dev(ctx, vol, volume_floats(ctx, ctx->T), "volume"))) return rc;
    if ((rc = need_dev(ctx, img, image_floats(ctx, ctx->T), "image"))) return rc;
    return run_forward(ctx, vol, nullptr, img, nullptr, nullptr, planes, nplanes, (hipStream_t)stream);
}

int flf_forward_ratio(flf_ctx* ctx, const float* vol, const float* im, float* img, float* ratio, const int* planes,
                      int nplanes, void* stream) {
    if (!ctx) return FLF_EINVAL;
    if (!ctx->planned) return fail(ctx, FLF_ESTATE, "flf_plan has not been called");
    int rc;
    if ((rc = need_dev(ctx, vol, volume_floats(ctx, ctx->T), "volume"))) return rc;
    if ((rc = need_dev(ctx, im, image_floats(ctx, ctx->T), "im"))) return rc;
    if ((rc = need_dev(ctx, ratio, image_floats(ctx, ctx->T), "ratio"))) return rc;
    if (img && (rc = need_dev(ctx, img, image_floats(ctx, ctx->T), "image"))) return rc;
    return run_forward(ctx, vol, nullptr, img, im, ratio, planes, nplanes, (hipStream_t)stream);
}

int flf_forward_pair(flf_ctx* ctx, const float* volA, const float* volB, float* img, const int* planes, int nplanes,
                     void* stream) {
    if (!ctx) return FLF_EINVAL;
    if (!ctx->planned) return fail(ctx, FLF_ESTATE, "flf_plan has not been called");
    if (ctx->T % 2) return fail(ctx, FLF_EINVAL, "flf_forward_pair needs an even active frame count");
    if (ctx->ragged) return fail(ctx, FLF_EINVAL, "flf_forward_pair does not take per-frame sizes");
    int rc;
    if ((rc = need_dev(ctx, volA, volume_floats(ctx, ctx->T / 2), "volA"))) return rc;
    if ((rc = need_dev(ctx, volB, volume_floats(ctx, ctx->T / 2), "volB"))) return rc;
    if ((rc = need_dev(ctx, img, image_floats(ctx, ctx->T), "image"))) return rc;
    return run_forward(ctx, volA, volB, img, nullptr, nullptr, planes, nplanes, (hipStream_t)stream);
}

namespace {
int run_forward_groups(flf_ctx* ctx, const float* vol, const float* vol2, const IfftRowSumArgs& rs,
                       const GroupFork& fk);

int run_forward(flf_ctx* ctx, const float* vol, const float* vol2, float* img, const float* im, float* ratio,
                const int* planes, int nplanes, hipStream_t st) {
    std::vector<char> sel;
    int rc = resolve_planes(ctx, planes, nplanes, sel);
    if (rc) return rc;
    const int T = ctx->T;
    IfftRowSumArgs rs{};
    rs.ngroups = (int)ctx->groups.size();
    rs.X = ctx->X;
    rs.Y = ctx->Y;
    rs.img = img;
    rs.im = im;
    rs.ratio = ratio;
    rs.img_tstride = (long long)ctx->Y * ctx->X;
    rs.fdims = frame_dims(ctx);
    rs.maxfX = 1;
    // enable flags of every group on st, then the groups side by side
    int nactive = 0;
    for (size_t gi = 0; gi < ctx->groups.size(); ++gi) {
        Group& g = ctx->groups[gi];
        const GridTables& gt = g.gt;
        bool any = false;
        rc = set_enable(ctx, g, sel, 0, any);
        if (rc) return rc;
        nactive += any ? 1 : 0;
        rs.grp[gi].active = any ? 1 : 0;
        rs.grp[gi].T = g.d_T_fwd;
        rs.grp[gi].tstride = (long long)ctx->Y * gt.rfX;
        rs.grp[gi].rfX = gt.rfX;
        rs.grp[gi].scale = 1.0f / ((float)gt.fY * (float)gt.fX);
        rs.grp[gi].rp = g.rpXw;
        rs.grp[gi].rph = g.rpXhw;
        rs.grp[gi].twf = g.twX;
        rs.grp[gi].twh = g.twXh;
        if (any) rs.maxfX = std::max(rs.maxfX, gt.fX);
    }
    GroupFork fk;
    rc = fork_groups(ctx, st, nactive, fk);
    if (rc) return rc;
    rc = run_forward_groups(ctx, vol, vol2, rs, fk);
    const int rj = join_groups(ctx, fk);
    if (rc) return rc;
    if (rj) return rj;
    HIPCHK(launch_ifft_rows_sum(rs, ctx->Y, T, st));
    return FLF_OK;
}

// Per active group: small FFTs of the volume planes, the MAC into the
// group's partial accumulators and the column transforms into the group's
// d_T_fwd (summed over groups by the row transform).
int run_forward_groups(flf_ctx* ctx, const float* vol, const float* vol2, const IfftRowSumArgs& rs,
                       const GroupFork& fk) {
    const int N = ctx->N, T = ctx->T;
    std::vector<GroupWork> ws;
    for (size_t gi = 0; gi < ctx->groups.size(); ++gi) {
        if (!rs.grp[gi].active) continue;
        Group& g = ctx->groups[gi];
        const GridTables& gt = g.gt;
        const int nP = (int)g.planes.size();
        GroupWork w{};
        w.gi = gi;
        SmallFFTArgs& a = w.sa;
        a.src = vol;
        // two sources: both hold T/2 frames, so the plane offsets are those of T/2 frames
        a.src_off = vol2 ? g.srcoff_T[T / 2 - 1] : g.d_srcoff_fwd;
        a.src2 = vol2;
        a.tsplit = vol2 ? T / 2 : T;
        a.src_tstride = (long long)ctx->Y * ctx->X;
        a.src_pitch = ctx->X;
        a.Y = ctx->Y; a.X = ctx->X; a.N = N;
        a.rpY = g.rpYr; a.rpX = g.rpXr;
        a.twY = g.twRY; a.twX = g.twRX;
        a.S = g.d_S_fwd;
        a.S_off = g.d_Soff_fwd;
        a.chunks = g.d_chunks; a.nchunks = (int)gt.chunks.size(); a.NH = gt.slots;
        a.fY = gt.fY; a.fX = gt.fX;
        a.S_tstride = s_frame_stride(ctx, g);
        a.fdims = frame_dims(ctx);
        w.nsrc = nP;
        MacArgs& m = w.ma;
        m.G = g.d_G[0];
        m.chunks = g.d_chunks; m.items = g.d_items; m.u_bu = g.d_bu; m.u_au = g.d_au; m.uoff = g.d_uoff;
        m.tiles = g.d_tiles;
        m.S = g.d_S_fwd; m.S_off = g.d_Soff_fwd; m.S_tstride = a.S_tstride;
        m.s2 = g.d_s2; m.enable = plane_mask(g);
        m.nP = nP;
        m.A = g.d_A_fwd;
        m.A_tstride = (long long)gt.fY * gt.rfX;
        m.A_pstride = (long long)T * gt.fY * gt.rfX;     // one partial accumulator per plane range
        m.N = N; m.fY = gt.fY; m.fX = gt.fX; m.rY = gt.rY; m.rX = gt.rX; m.rfX = gt.rfX; m.nu = gt.nu;
        m.T = T;
        m.rs = mac_row_stride(gt.slots, gt.nu);
        // Split the plane sum into ranges so that the launch has enough
        // workgroups to keep the tail short; the column transform adds the
        // partial accumulators in a fixed order.
        w.nranges = fwd_split(ctx, g, nP);
        IfftColArgs& ic = w.ca;
        ic.A = g.d_A_fwd;
        ic.A_bstride = (long long)gt.fY * gt.rfX;
        ic.nsum = w.nranges;
        ic.A_sstride = m.A_pstride;
        ic.rfX = gt.rfX; ic.CB = g.CBi; ic.Yout = ctx->Y;
        ic.rpY = g.rpYw;
        ic.twY = g.twY;
        ic.out = g.d_T_fwd;
        ic.out_bstride = (long long)ctx->Y * gt.rfX;
        w.nbatch = T;
        ws.push_back(w);
    }
    return run_group_work(ctx, ws, fk);
}
}  // namespace

int flf_backward(flf_ctx* ctx, const float* img, float* vol, const int* planes, int nplanes, void* stream) {
    if (!ctx) return FLF_EINVAL;
    return run_backward(ctx, img, vol, nullptr, nullptr, 0, planes, nplanes, (hipStream_t)stream);
}

int flf_backward_rl_update(flf_ctx* ctx, const float* img, float* X, const float* Htf, const float* trust,
                           int mode, const int* planes, int nplanes, void* stream) {
    if (!ctx) return FLF_EINVAL;
    if (mode != 1 && mode != 2) return fail(ctx, FLF_EINVAL, "mode must be 1 (Prevedel) or 2 (standard)");
    return run_backward(ctx, img, X, Htf, trust, mode, planes, nplanes, (hipStream_t)stream);
}

namespace {
// Shared checks of flf_rl_scale / flf_rl_ratio: every array on the current
// device and as large as the slabs it covers.
int rl_scale_call(const float* Htf, const float* v, float* X, int nz, int T, long long npix, int v_frames,
                  void* stream, bool ratio) {
    tls_error().clear();
    if (nz < 1 || T < 1 || npix < 0 || (v_frames != 1 && v_frames != T) || (long long)nz * T > 65535) {
        tls_error() = "flf_rl_scale: bad nz, T, npix or v_frames";
        return FLF_EINVAL;
    }
    if (npix == 0) return FLF_OK;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return FLF_EHIP;
    const size_t nx = (size_t)nz * T * npix * sizeof(float), nv = (size_t)nz * v_frames * npix * sizeof(float);
    std::string why;
    if (!device_span(dev, Htf, nx, "Htf", why) || !device_span(dev, v, nv, "v", why) ||
        !device_span(dev, X, nx, "X", why)) {
        tls_error() = why;
        return FLF_EINVAL;
    }
    return flf::launch_rl_scale(X, Htf, v, nz, T, npix, v_frames, (hipStream_t)stream, ratio) == hipSuccess
               ? FLF_OK
               : FLF_EHIP;
}
}  // namespace

int flf_rl_scale(const float* Htf, const float* v, float* X, int nz, int T, long long npix, int v_frames,
                 void* stream) {
    return rl_scale_call(Htf, v, X, nz, T, npix, v_frames, stream, false);
}

int flf_rl_ratio(const float* Htf, const float* v, float* X, int nz, int T, long long npix, int v_frames,
                 void* stream) {
    return rl_scale_call(Htf, v, X, nz, T, npix, v_frames, stream, true);
}

int flf_datafit(flf_ctx* ctx, const float* img, const float* im, int im_frames, double* out_host, void* stream) {
    if (!ctx) return FLF_EINVAL;
    if (!ctx->planned) return fail(ctx, FLF_ESTATE, "flf_plan has not been called");
    if (!out_host) return fail(ctx, FLF_EINVAL, "null output pointer");
    if (im_frames == 0) im_frames = ctx->T;
    if (im_frames < 1 || im_frames > ctx->T) return fail(ctx, FLF_EINVAL, "im_frames outside 1..active frames");
    int rc;
    if ((rc = need_dev(ctx, img, image_floats(ctx, ctx->T), "image"))) return rc;
    if ((rc = need_dev(ctx, im, image_floats(ctx, im_frames), "im"))) return rc;
    hipStream_t st = (hipStream_t)stream;
    HIPCHK(launch_fit(img, im, (long long)ctx->Y * ctx->X, ctx->T, im_frames, ctx->d_fitpart, FIT_BLOCKS,
                      ctx->d_fitout, st));
    HIPCHK(hipMemcpyAsync(out_host, ctx->d_fitout, sizeof(double) * 2 * ctx->T, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return FLF_OK;
}

int flf_check_device_buffer(int device, const void* p, long long bytes) {
    tls_error().clear();
    if (bytes < 0) {
        tls_error() = "negative size";
        return FLF_EINVAL;
    }
    std::string why;
    if (!device_span(device, p, (size_t)bytes, "buffer", why)) {
        tls_error() = why;
        return FLF_EINVAL;
    }
    return FLF_OK;
}

int flf_datafit_noalloc(const float* img, const float* im, int T, int im_frames, long long npix, double* out,
                        void* stream) {
    tls_error().clear();
    if (T < 1 || T > FLF_NOALLOC_MAX_FRAMES) {
        tls_error() = "flf_datafit_noalloc: T must be 1.." + std::to_string(FLF_NOALLOC_MAX_FRAMES);
        return FLF_EINVAL;
    }
    if (im_frames < 1 || im_frames > T || npix < 1) {
        tls_error() = "flf_datafit_noalloc: im_frames must be 1..T and npix >= 1";
        return FLF_EINVAL;
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return FLF_EHIP;
    std::string why;
    if (!device_span(dev, img, (size_t)T * npix * sizeof(float), "image", why) ||
        !device_span(dev, im, (size_t)im_frames * npix * sizeof(float), "im", why) ||
        !device_span(dev, out, (size_t)2 * T * sizeof(double), "out", why)) {
        tls_error() = why;
        return FLF_EINVAL;
    }
    double* part = nullptr;
    hipError_t e = noalloc_scratch(dev, &part);
    if (e != hipSuccess) {
        tls_error() = std::string("scratch address: ") + hipGetErrorString(e);
        return FLF_EHIP;
    }
    // every partial that k_fit_final reads is written by k_fit_partial in this
    // call, so the scratch needs no clearing
    e = launch_fit(img, im, npix, T, im_frames, part, FIT_BLOCKS, out, (hipStream_t)stream);
    if (e != hipSuccess) {
        tls_error() = std::string("flf_datafit_noalloc: ") + hipGetErrorString(e);
        return FLF_EHIP;
    }
    return FLF_OK;
}

int flf_profile_enable(flf_ctx* ctx, int enable) {
    if (!ctx) return FLF_EINVAL;
    ctx->prof = enable != 0;
    return FLF_OK;
}

int flf_profile_read(flf_ctx* ctx, double* mac_ms, long long* mac_launches, double* mac_bytes) {
    if (!ctx || !mac_ms || !mac_launches || !mac_bytes) return FLF_EINVAL;
    double ms = 0.0, by = 0.0;
    long long n = 0;
    for (auto& r : ctx->prof_recs) {
        HIPCHK(hipEventSynchronize(r.b));
        float t = 0.f;
        HIPCHK(hipEventElapsedTime(&t, r.a, r.b));
        ms += t;
        by += r.bytes;
        ++n;
        ctx->ev_pool.push_back(r.a);
        ctx->ev_pool.push_back(r.b);
    }
    ctx->prof_recs.clear();
    *mac_ms = ms;
    *mac_launches = n;
    *mac_bytes = by;
    return FLF_OK;
}

int flf_radix_plan(int L, int wide, int order, int* radices, int* npass) {
    if (!radices || !npass || L < 1) return FLF_EINVAL;
    RadixPlan rp;
    if (!make_radix_plan(L, rp, wide != 0, order)) return FLF_EINVAL;
    for (int i = 0; i < rp.n; ++i) radices[i] = rp.r[i];
    *npass = rp.n;
    return FLF_OK;
}

int flf_grid_check(int nnum, int fY, int fX, long long* stats) {
    if (!stats) return FLF_EINVAL;
    GridTables gt;
    // the tables the projector plans by default: lane pairs of two items
    std::string e = build_grid_tables(nnum, fY, fX, gt, 2);
    if (!e.empty()) return FLF_EINVAL;
    // every workgroup holds whole pairs, and the two items of a pair share
    // their residues (one P/M table row serves both)
    for (auto& c : gt.chunks) {
        if (c.n_items % 2) return FLF_EINVAL;
        for (int i = 0; i < c.n_items; i += 2) {
            const ItemDesc& x = gt.items[c.item_base + i];
            const ItemDesc& y = gt.items[c.item_base + i + 1];
            for (int j = 0; j < 4; ++j)
                if (x.res[j] != y.res[j]) return FLF_EINVAL;
        }
    }
    int max_items = 0;
    for (auto& c : gt.chunks) max_items = std::max(max_items, c.n_items);
    stats[0] = (long long)gt.chunks.size();
    stats[1] = (long long)gt.items.size();
    stats[2] = gt.max_res;
    stats[3] = gt.nu;
    stats[4] = gt.slots;
    stats[5] = max_items;
    return FLF_OK;
}

int flf_fourier_len(int s1, int s2, int nnum) {
    if (s1 < 1 || s2 < 1 || nnum < 1) return FLF_EINVAL;
    return fourier_len(s1, s2, nnum);
}

}  // extern "C"
